"""Checkpoint / restart and VTK output of the resident state (src/mod_io.f90) on the GPU.

* Restart replay (tests/test_week6_advanced.f90:126-152): checkpoint at step 3, two more steps,
  then a restart from the checkpoint replays them -- every padded array, ghosts included,
  bit-identical to the continuous run (the reference asks for < 1e-14), both into a fresh context
  and into the same one.
* Byte-exact files: the GPU's checkpoint is byte for byte the restatement's (oracle/io_ref.py, pinned
  to a Fortran runtime by tests/test_io_formats.py) of the downloaded state, and so is its VTK file.
* CPU <-> GPU restart: the oracle continues from the GPU's checkpoint and the GPU from the oracle's,
  with bit-identical states after the next steps (reference summation order); other dimensions are
  refused."""
import numpy as np
import pytest

import io_ref as IO
import pyoracle as O
from cases import droplet_c, phys

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")

PH = dict(rho1=1000.0, rho2=1.0, mu1=1.0, mu2=0.01, sigma=1.0, time_order=2)
OKW = dict(use_bicgstab2=False, tol=1e-8, max_iter=60)
DT = 1e-4


def make(n=32):
    g = O.Grid(n, n)
    ph = phys(**PH)
    ctx = abi.Context(n, n, params=abi.FuguParams(1000.0, 1.0, 1.0, 0.01, 1.0, 0.0, 0.0, 2, 1e12, 0.5),
                      sum_order=abi.SUM_REFERENCE)
    return g, ph, ctx


def start(g, ctx):
    z = [np.zeros(g.shape(k)) for k in ("u", "v", "u", "v", "p")]
    ctx.upload(*z, droplet_c(g, 0.25))
    ctx.init_properties()


def state(g, ctx):
    out = [np.zeros(g.shape(k)) for k in ("u", "v", "u", "v", "p", "p")]
    ctx.download(*out)
    return out


def test_restart_replay_bitwise(tmp_path):
    g, ph, ctx = make()
    start(g, ctx)
    opts = abi.step_opts(**OKW)
    for s in (1, 2, 3):
        ctx.step(DT, s, opts)
    prefix = str(tmp_path / "chk")
    ctx.checkpoint_write(prefix, 3, 3 * DT)
    for s in (4, 5):
        ctx.step(DT, s, opts)
    ref = state(g, ctx)
    for target in ("fresh", "same"):
        c2 = make()[2] if target == "fresh" else ctx
        ok, step, t = c2.checkpoint_read(prefix)
        assert ok and step == 3 and t == 3 * DT
        for s in (step + 1, step + 2):
            c2.step(DT, s, opts)
        got = state(g, c2)
        for a, b, name in zip(got, ref, ("u", "v", "u_prev", "v_prev", "p", "c")):
            assert np.array_equal(a, b), (target, name)
        if target == "fresh":
            c2.close()
    # other dimensions are refused and leave the state alone
    g24, _, c24 = make(24)
    start(g24, c24)
    before = state(g24, c24)
    ok, step, t = c24.checkpoint_read(prefix)
    assert not ok and step == 3
    assert all(np.array_equal(a, b) for a, b in zip(state(g24, c24), before))
    assert c24.checkpoint_read(str(tmp_path / "missing"))[0] is False
    c24.close()
    ctx.close()


def test_checkpoint_and_vtk_bytes_and_cpu_gpu_restart(tmp_path):
    g, ph, ctx = make()
    start(g, ctx)
    opts = abi.step_opts(**OKW)
    for s in (1, 2):
        ctx.step(DT, s, opts)
    gpu_prefix, cpu_prefix = str(tmp_path / "gpu"), str(tmp_path / "cpu")
    ctx.checkpoint_write(gpu_prefix, 2, 2 * DT)
    S = state(g, ctx)
    IO.write_checkpoint(cpu_prefix, 2, 2 * DT, *S, g.nx, g.ny)
    assert open(IO.checkpoint_name(gpu_prefix), "rb").read() == open(IO.checkpoint_name(cpu_prefix), "rb").read()
    # VTK of the resident state against the restatement on the downloaded fields
    ctx.write_vtk(str(tmp_path / "gpu.vtk"))
    f = {k: ctx.field(k, np.zeros(g.shape(kk))) for k, kk in
         (("p", "p"), ("c", "p"), ("rho_c", "p"), ("kappa_curv", "p"), ("u", "u"), ("v", "v"))}
    IO.write_vtk(str(tmp_path / "cpu.vtk"), f["p"], f["c"], f["rho_c"], f["kappa_curv"], f["u"], f["v"],
                 g.nx, g.ny, g.dx, g.dy)
    a, b = open(tmp_path / "gpu.vtk", "rb").read(), open(tmp_path / "cpu.vtk", "rb").read()
    assert a == b, (len(a), len(b))
    # the oracle continues from the GPU's file, a fresh GPU context from the oracle's
    O.set_sum_mode(0)
    ok, step, t, arrs = IO.read_checkpoint(gpu_prefix, g.nx, g.ny)
    assert ok and step == 2
    st = O.State(g)
    st.u[...], st.v[...], st.u_prev[...], st.v_prev[...], st.p[...], st.c[...] = arrs
    c2 = make()[2]
    ok2, step2, _ = c2.checkpoint_read(cpu_prefix)
    assert ok2 and step2 == 2
    for s in (3, 4):
        O.advance_step(st, ph, DT, s, **OKW)
        c2.step(DT, s, opts)
    for a, b, name in zip(state(g, c2), (st.u, st.v, st.u_prev, st.v_prev, st.p, st.c),
                          ("u", "v", "u_prev", "v_prev", "p", "c")):
        assert np.array_equal(a, b), name
    c2.close()
    ctx.close()
