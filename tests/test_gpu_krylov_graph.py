"""The reference Krylov solvers' iterations replayed from a HIP graph are bit-identical to the eager
launches: BiCGStab (src/mod_bicgstab2.f90:77-264) and BiCGStab(2) (:270-570), as the reference has
it and with consistent boundary conditions, on the coupled step (src/mod_solver_coupled.f90:137-380)
with walls and a lid and on a periodic droplet, in both summation orders.  The graphs are opt-in:
FUGU_KRYLOV_GRAPHS=1 at context creation turns them on, 0 (the default) gives the eager launches.  Every state field,
ghosts included, the iteration counts and the residuals agree exactly over several steps, so a
second capture (a new dt changes the key) is covered too."""
import os

import numpy as np
import pytest

import pyoracle as O
from cases import droplet_c, phys

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")

STATE = ("u", "v", "p", "c", "u_prev", "v_prev", "kappa_curv")


def make_ctx(g, ph, sum_order, graphs):
    old = os.environ.get("FUGU_KRYLOV_GRAPHS")
    os.environ["FUGU_KRYLOV_GRAPHS"] = "1" if graphs else "0"
    try:
        return abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y,
                           params=abi.FuguParams(ph.rho1, ph.rho2, ph.mu1, ph.mu2, ph.sigma, ph.gx, ph.gy,
                                                 ph.time_order, ph.alpha_penalty, ph.cfl_target),
                           sum_order=sum_order)
    finally:
        if old is None:
            del os.environ["FUGU_KRYLOV_GRAPHS"]
        else:
            os.environ["FUGU_KRYLOV_GRAPHS"] = old


def run(g, ph, init, dts, opts, sum_order, graphs):
    st = O.State(g)
    init(st)
    O.halo(g, st.c, "p")
    st.init_properties(ph)
    ctx = make_ctx(g, ph, sum_order, graphs)
    out = []
    for s, dt in enumerate(dts, 1):
        it, res, rc = ctx.advance_step(st, dt, s, opts)
        out.append((it, res, rc))
    ctx.close()
    return st, out


@pytest.mark.parametrize("case", ["cavity_b1", "cavity_b2", "droplet_b2", "droplet_b2_consistent"])
def test_krylov_graph_matches_eager(case, sum_order):
    if case.startswith("cavity"):
        g = O.Grid(32, 32)
        ph = phys(rho1=1.0, rho2=1.0, mu1=0.01, mu2=0.01, time_order=1)

        def init(st):
            st.c[...] = 1.0
            O.velocity_bcs(g, st.u, st.v, 1.0)
        opts = abi.step_opts(use_bicgstab2=case.endswith("b2"), u_lid=1.0, tol=1e-6, max_iter=40)
        dts = [0.01, 0.01, 0.005, 0.005]
    else:
        g = O.Grid(32, 32, per_x=True, per_y=True)
        ph = phys(rho1=1000.0, rho2=1.0, mu1=1.0, mu2=0.01, sigma=1.0, time_order=2)

        def init(st):
            st.c[...] = droplet_c(g, 0.25)
        flags = abi.B2_CONSISTENT_BCS if case.endswith("consistent") else 0
        opts = abi.step_opts(use_bicgstab2=True, tol=1e-8, max_iter=30, solver_flags=flags)
        dts = [1e-4, 1e-4, 5e-5]
    st_e, out_e = run(g, ph, init, dts, opts, sum_order, graphs=False)
    st_g, out_g = run(g, ph, init, dts, opts, sum_order, graphs=True)
    assert out_g == out_e, (out_g, out_e)
    for name in STATE:
        assert np.array_equal(getattr(st_g, name), getattr(st_e, name)), name
