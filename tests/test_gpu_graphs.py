"""GPU tests of the FGMRES steps replayed as HIP graphs (FUGU_GRAPHS, default on): a solve whose
steps are captured once and replayed -- across restarts, across the time steps that follow, and
after a change of dt that forces a new capture -- gives bit for bit the fields and the iteration
counts of the same solve launched kernel by kernel (FUGU_GRAPHS=0).  Every replay is checked: the
fields are compared after each time step, with the library's own eager work (VOF, properties,
restarts, host reads) between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")


def run_steps(monkeypatch, graphs, nx, ny, per_x, mixed, mg_cycles, dts):
    import bench
    monkeypatch.setenv("FUGU_GRAPHS", "1" if graphs else "0")
    monkeypatch.setenv("FUGU_VMG_CYCLES", str(mg_cycles))   # two V-cycles: a copy node in the graph
    p = abi.FuguParams(1000.0, 1.0, 10.0, 0.1, 2.0, 0.0, -1.0, 2, 1e12, 0.5)
    ctx = abi.Context(nx, ny, 1.0, 2.0, per_x, False, params=p)
    nh = 2
    u = np.zeros((ny + 2 * nh, nx + 1 + 2 * nh)); v = np.zeros((ny + 1 + 2 * nh, nx + 2 * nh))
    pp = np.zeros((ny + 2 * nh, nx + 2 * nh))
    ctx.upload(u, v, u.copy(), v.copy(), pp, bench.bubble_c(nx, ny, 1.0, 2.0))
    ctx.init_properties()
    flags = (abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG |
             (abi.SOLVE_MIXED if mixed else 0))
    # restart 8: several cycles per solve, so each captured step is replayed within a solve too
    opts = abi.step_opts(tol=1e-8, max_iter=300, solver_flags=flags, inner_max_iter=-1, restart=8,
                         mg_cycles=mg_cycles)
    out = []
    for s, dt in enumerate(dts, start=1):
        st, rc = ctx.step(dt, s, opts)
        assert rc == 0 and st.res <= 1e-8, (s, rc, st.res)
        f = [ctx.field(k, np.zeros(shape)) for k, shape in (("u", u.shape), ("v", v.shape), ("p", pp.shape))]
        out.append((st.outer, st.pcg_calls, f))
    ctx.close()
    return out


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("per_x,mg_cycles", [(False, 1), (True, 1), (False, 2)])
def test_graph_replay_is_bitwise_the_eager_solve(monkeypatch, mixed, per_x, mg_cycles):
    nx, ny = 64, 128
    dts = [2e-4, 2e-4, 2e-4, 1e-4, 1e-4]   # step 2 changes the BDF2 operator, step 4 the dt
    eager = run_steps(monkeypatch, False, nx, ny, per_x, mixed, mg_cycles, dts)
    graph = run_steps(monkeypatch, True, nx, ny, per_x, mixed, mg_cycles, dts)
    for s, (a, b) in enumerate(zip(eager, graph), start=1):
        assert a[0] == b[0], ("FGMRES steps", s, a[0], b[0])
        assert a[1] == b[1], ("pressure-block calls", s, a[1], b[1])
        for fa, fb, name in zip(a[2], b[2], "uvp"):
            assert np.array_equal(fa, fb), (name, s, np.abs(fa - fb).max())
