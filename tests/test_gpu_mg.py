"""GPU tests of the multigrid pressure block (FUGU_PRECOND_MG, no reference counterpart):
V-cycles converge to the dense solution of the reference's pressure operator, and one
V-cycle is a fixed linear map (what the outer Krylov method needs)."""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, two_phase_case
from test_oracle_solvers import neumann_pressure_matrix

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")


@pytest.mark.parametrize("per_x", [False, True])
def test_mg_converges_to_dense_solution(per_x):
    g = O.Grid(32, 24, 1.0, 0.75, per_x=per_x)
    rng = np.random.default_rng(200)
    c, pr, _, _ = two_phase_case(g, rng, rho=(100.0, 1.0))
    if per_x:
        O.halo(g, pr["rho_u"], "u"); O.halo(g, pr["rho_v"], "v")
    rhs = rand(g, "p", rng)
    sl = interior(g, "p")
    rhs[sl] -= rhs[sl].mean()
    A = neumann_pressure_matrix(g, pr["rho_u"], pr["rho_v"])
    x = np.linalg.lstsq(A, rhs[sl].ravel(), rcond=None)[0]
    x -= x.mean()
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly, per_x=per_x)
    errs = []
    for n in (1, 2, 4, 40):
        phi = g.zeros("p")
        ctx.pressure_mg(phi, rhs, pr["rho_u"], pr["rho_v"], n_cycles=n)
        got = phi[sl].ravel() - phi[sl].mean()
        errs.append(np.abs(got - x).max() / np.abs(x).max())
    assert errs[-1] < 1e-9, errs
    assert errs[1] < errs[0] and errs[2] < errs[1], errs
    assert errs[0] < 0.5, errs


def test_mg_vcycle_is_linear():
    g = O.Grid(64, 32, 2.0, 1.0)
    rng = np.random.default_rng(201)
    c, pr, _, _ = two_phase_case(g, rng, rho=(1000.0, 1.0))
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly)
    b1, b2 = rand(g, "p", rng), rand(g, "p", rng)
    out = []
    for b in (b1, b2, 3.0 * b1 - 2.0 * b2):
        phi = g.zeros("p")
        ctx.pressure_mg(phi, b, pr["rho_u"], pr["rho_v"], n_cycles=1)
        out.append(phi)
    lin = 3.0 * out[0] - 2.0 * out[1]
    assert np.abs(out[2] - lin).max() <= 1e-12 * np.abs(lin).max()


@pytest.mark.parametrize("nx,ny,per_x,per_y", [(32, 24, False, False), (48, 40, True, False), (40, 64, False, True),
                                               (96, 80, True, True), (130, 70, False, False), (64, 128, False, False)])
def test_coarse_tail_is_bitwise_the_multilaunch_cycle(monkeypatch, nx, ny, per_x, per_y):
    """The single-launch coarse tail (register-resident quads, LDS-staged or global) and the fused
    (temporal-blocking) level kernels perform the one-launch-per-half-sweep V-cycle's arithmetic:
    results equal bit for bit (signs of zeros included) with both off, at a 4096-cell threshold,
    with the default (LDS-fit) start and with the whole hierarchy in the tail (odd extents
    included), with the quad form of the tail on and off, with the LDS-tile and the
    register-strip level kernels (large and small strip shapes)."""
    g = O.Grid(nx, ny, 1.0, ny / nx, per_x=per_x, per_y=per_y)
    rng = np.random.default_rng(202)
    c, pr, _, _ = two_phase_case(g, rng, rho=(100.0, 1.0))
    O.halo(g, pr["rho_u"], "u"); O.halo(g, pr["rho_v"], "v")
    rhs = rand(g, "p", rng)
    out = {}
    for cells, fused, strip, quad in (("0", "0", "0", "1"), ("0", "1", "0", "1"), ("4096", "1", "0", "1"),
                                      ("1000000", "1", "0", "1"), ("4096", "0", "0", "1"), ("-1", "1", "0", "1"),
                                      ("-1", "1", "0", "0"), ("-1", "0", "0", "0"), ("0", "1", "2", "1"),
                                      ("-1", "1", "2", "1"), ("0", "1", "3", "1"), ("-1", "1", "3", "1")):
        monkeypatch.setenv("FUGU_MG_TAIL_CELLS", cells)
        monkeypatch.setenv("FUGU_MG_FUSED", fused)
        monkeypatch.setenv("FUGU_MG_STRIP", strip)
        monkeypatch.setenv("FUGU_MG_TAIL_QUAD", quad)
        ctx = abi.Context(g.nx, g.ny, g.lx, g.ly, per_x=per_x, per_y=per_y)
        res = []
        for n in (1, 3):
            phi = g.zeros("p")
            ctx.pressure_mg(phi, rhs, pr["rho_u"], pr["rho_v"], n_cycles=n)
            res.append(phi)
        ctx.close()
        out[(cells, fused, strip, quad)] = res
    for key in out:
        for a, b in zip(out[key], out[("0", "0", "0", "1")]):
            assert np.array_equal(a, b) and np.array_equal(np.signbit(a), np.signbit(b)), key


@pytest.mark.parametrize("theta,coarse_nu,cycles", [("0", "2", "1"), ("3", "2", "1"), ("10", "4", "0"), ("10", "3", "2"),
                                                   ("1e9", "4", "2")])
@pytest.mark.parametrize("mixed", [False, True])
def test_coarse_tail_bitwise_in_the_step_solve(monkeypatch, mixed, theta, coarse_nu, cycles):
    """A refined step solve (pressure and velocity hierarchies: mass terms, Dirichlet faces, FP64
    or FP32) is bit-identical with the coarse tail and fused level kernels on and off -- with the
    velocity hierarchies at full depth (theta 0), cut by their diffusion-to-mass ratio (3, 10; the
    cut coarsest level smoothed by 2, 3 or 4 sweeps) and cut to the finest level alone (1e9,
    whose sweeps are then the whole velocity block), with one, two or the ratio-chosen number of
    V-cycles per velocity block; the fused levels by the LDS-tile or (on every level) the
    register-strip kernels of either shape."""
    import bench
    nx, ny = 96, 192
    outs = []
    monkeypatch.setenv("FUGU_VMG_THETA", theta)
    monkeypatch.setenv("FUGU_VMG_THETA2", theta)
    monkeypatch.setenv("FUGU_VMG_COARSE_NU", coarse_nu)
    monkeypatch.setenv("FUGU_VMG_CYCLES", cycles)
    for cells, fused, strip, quad in (("0", "0", "0", "1"), ("-1", "1", "0", "1"), ("-1", "1", "0", "0"),
                                      ("-1", "1", "2", "1"), ("-1", "1", "3", "1")):
        monkeypatch.setenv("FUGU_MG_TAIL_CELLS", cells)
        monkeypatch.setenv("FUGU_MG_FUSED", fused)
        monkeypatch.setenv("FUGU_MG_STRIP", strip)
        monkeypatch.setenv("FUGU_MG_TAIL_QUAD", quad)
        p = abi.FuguParams(1000.0, 1.0, 10.0, 0.1, 2.0, 0.0, -1.0, 2, 1e12, 0.5)
        ctx = abi.Context(nx, ny, 1.0, 2.0, False, False, params=p)
        nh = 2
        u = np.zeros((ny + 2 * nh, nx + 1 + 2 * nh)); v = np.zeros((ny + 1 + 2 * nh, nx + 2 * nh))
        pp = np.zeros((ny + 2 * nh, nx + 2 * nh))
        ctx.upload(u, v, u.copy(), v.copy(), pp, bench.bubble_c(nx, ny, 1.0, 2.0))
        ctx.init_properties()
        flags = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG | (abi.SOLVE_MIXED if mixed else 0)
        opts = abi.step_opts(tol=1e-8, max_iter=200, solver_flags=flags, inner_max_iter=-1, restart=30)
        its = []
        for s in (1, 2):
            st, rc = ctx.step(2e-4, s, opts)
            assert rc == 0 and st.res <= 1e-8
            its.append(st.outer)
        f = [ctx.field(k, np.zeros(shape)) for k, shape in (("u", u.shape), ("v", v.shape), ("p", pp.shape))]
        ctx.close()
        outs.append((its, f))
    for o in outs[1:]:
        assert outs[0][0] == o[0]
        for a, b in zip(outs[0][1], o[1]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("nu_p,nu_v", [("1", "2"), ("3", "4"), ("2", "3"), ("4", "1")])
def test_sweep_counts_fused_is_bitwise_the_multilaunch_cycle(monkeypatch, nu_p, nu_v):
    """Other red-black sweep counts per V-cycle half (pressure and velocity hierarchies set apart):
    the fused level kernels (LDS tiles, and register strips of either shape on every level) and the coarse tail
    still perform the multi-launch cycle's arithmetic, bit for bit, in a refined step solve."""
    import bench
    nx, ny = 96, 192
    monkeypatch.setenv("FUGU_MG_NU_P", nu_p)
    monkeypatch.setenv("FUGU_MG_NU_V", nu_v)
    monkeypatch.setenv("FUGU_VMG_THETA", "0")     # full-depth velocity hierarchies: their tail runs too
    monkeypatch.setenv("FUGU_VMG_THETA2", "0")
    outs = []
    for cells, fused, strip, quad in (("0", "0", "0", "1"), ("-1", "1", "0", "1"), ("-1", "1", "0", "0"),
                                      ("-1", "1", "2", "1"), ("-1", "1", "3", "1")):
        monkeypatch.setenv("FUGU_MG_TAIL_CELLS", cells)
        monkeypatch.setenv("FUGU_MG_FUSED", fused)
        monkeypatch.setenv("FUGU_MG_STRIP", strip)
        monkeypatch.setenv("FUGU_MG_TAIL_QUAD", quad)
        p = abi.FuguParams(1000.0, 1.0, 10.0, 0.1, 2.0, 0.0, -1.0, 2, 1e12, 0.5)
        ctx = abi.Context(nx, ny, 1.0, 2.0, False, False, params=p)
        nh = 2
        u = np.zeros((ny + 2 * nh, nx + 1 + 2 * nh)); v = np.zeros((ny + 1 + 2 * nh, nx + 2 * nh))
        pp = np.zeros((ny + 2 * nh, nx + 2 * nh))
        ctx.upload(u, v, u.copy(), v.copy(), pp, bench.bubble_c(nx, ny, 1.0, 2.0))
        ctx.init_properties()
        flags = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG
        opts = abi.step_opts(tol=1e-8, max_iter=300, solver_flags=flags, inner_max_iter=-1, restart=30)
        its = []
        for s in (1, 2):
            st, rc = ctx.step(2e-4, s, opts)
            its.append((rc, st.outer, st.res))
        f = [ctx.field(k, np.zeros(shape)) for k, shape in (("u", u.shape), ("v", v.shape), ("p", pp.shape))]
        ctx.close()
        outs.append((its, f))
    assert all(rc == 0 and res <= 1e-8 for rc, _, res in outs[0][0]), outs[0][0]
    for o in outs[1:]:
        assert outs[0][0] == o[0]
        for a, b, name in zip(outs[0][1], o[1], "uvp"):
            assert np.array_equal(a, b), (name, np.abs(a - b).max())


@pytest.mark.parametrize("mixed", [False, True])
def test_big_levels_fused_is_bitwise_the_multilaunch_cycle(monkeypatch, mixed):
    """Levels of 2^21 cells and more take the register-strip level kernels (FP64: 4-row strips for
    the two-sweep pressure levels, 8-row strips for the three-sweep velocity levels; FP32: 8-row
    strips), and with the strips off the LDS-tile kernels (64-row tiles for the FP64 three-sweep
    pre-smoothing): one refined step solve on 1024 x 2048 cells, FP64 or mixed, bit for bit the
    multi-launch cycle's."""
    import bench
    nx, ny = 1024, 2048
    outs = []
    for cells, fused, strip in (("0", "0", "0"), ("-1", "1", "1"), ("-1", "1", "0")):
        monkeypatch.setenv("FUGU_MG_TAIL_CELLS", cells)
        monkeypatch.setenv("FUGU_MG_FUSED", fused)
        monkeypatch.setenv("FUGU_MG_STRIP", strip)
        p = abi.FuguParams(1000.0, 1.0, 10.0, 0.1, 2.0, 0.0, -1.0, 2, 1e12, 0.5)
        ctx = abi.Context(nx, ny, 1.0, 2.0, False, False, params=p)
        nh = 2
        u = np.zeros((ny + 2 * nh, nx + 1 + 2 * nh)); v = np.zeros((ny + 1 + 2 * nh, nx + 2 * nh))
        pp = np.zeros((ny + 2 * nh, nx + 2 * nh))
        ctx.upload(u, v, u.copy(), v.copy(), pp, bench.bubble_c(nx, ny, 1.0, 2.0))
        ctx.init_properties()
        flags = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG | (abi.SOLVE_MIXED if mixed else 0)
        opts = abi.step_opts(tol=1e-8, max_iter=300, solver_flags=flags, inner_max_iter=-1, restart=30)
        st, rc = ctx.step(2e-4, 1, opts)
        f = [ctx.field(k, np.zeros(shape)) for k, shape in (("u", u.shape), ("v", v.shape), ("p", pp.shape))]
        ctx.close()
        outs.append(((rc, st.outer, st.res), f))
        print(f"fused={fused} strip={strip}: rc {rc} outer {st.outer} res {st.res:.3e}", flush=True)
    assert outs[0][0][0] == 0 and outs[0][0][2] <= 1e-8, outs[0][0]
    for o in outs[1:]:
        assert outs[0][0] == o[0]
        for a, b, name in zip(outs[0][1], o[1], "uvp"):
            assert np.array_equal(a, b), (name, np.abs(a - b).max())


def test_odd_periodic_extent_is_refused():
    """Red-black smoothing needs even periodic extents: the multigrid block says so instead of
    running a colouring whose seam neighbours share a colour."""
    g = O.Grid(25, 16, 1.0, 0.64, per_x=True)
    rng = np.random.default_rng(203)
    c, pr, _, _ = two_phase_case(g, rng, rho=(10.0, 1.0))
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly, per_x=True)
    with pytest.raises(Exception, match="even"):
        ctx.pressure_mg(g.zeros("p"), rand(g, "p", rng), pr["rho_u"], pr["rho_v"], n_cycles=1)
    ctx.close()


@pytest.mark.parametrize("per_x,per_y", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("mixed", [False, True])
def test_fused_pcd_glue_is_bitwise(monkeypatch, mixed, per_x, per_y):
    """The PCD's fused glue -- the v-block store, periodic halo and cross term in one launch, and
    the u-block store, periodic halo and the velocity BCs FGMRES applies to z_k in another -- gives
    the fields and step counts of the separate launches, bit for bit (walls, periodic x, periodic y);
    also with the register-strip level kernels (either shape) on every level."""
    import bench
    nx, ny = 96, 192
    outs = []
    for fused, strip in (("0", "0"), ("1", "0"), ("1", "2"), ("1", "3")):
        monkeypatch.setenv("FUGU_PCD_FUSED_GLUE", fused)
        monkeypatch.setenv("FUGU_MG_STRIP", strip)
        p = abi.FuguParams(1000.0, 1.0, 10.0, 0.1, 2.0, 0.0, -1.0, 2, 1e12, 0.5)
        ctx = abi.Context(nx, ny, 1.0, 2.0, per_x, per_y, params=p)
        nh = 2
        u = np.zeros((ny + 2 * nh, nx + 1 + 2 * nh)); v = np.zeros((ny + 1 + 2 * nh, nx + 2 * nh))
        pp = np.zeros((ny + 2 * nh, nx + 2 * nh))
        ctx.upload(u, v, u.copy(), v.copy(), pp, bench.bubble_c(nx, ny, 1.0, 2.0))
        ctx.init_properties()
        flags = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG | (abi.SOLVE_MIXED if mixed else 0)
        opts = abi.step_opts(tol=1e-8, max_iter=200, solver_flags=flags, inner_max_iter=-1, restart=30)
        its = []
        for s in (1, 2, 3):
            st, rc = ctx.step(2e-4, s, opts)
            its.append((rc, st.outer, st.res))
        f = [ctx.field(k, np.zeros(shape)) for k, shape in (("u", u.shape), ("v", v.shape), ("p", pp.shape))]
        ctx.close()
        outs.append((its, f))
    assert all(rc == 0 and res <= 1e-8 for rc, _, res in outs[0][0]), outs[0][0]
    for o in outs[1:]:
        assert outs[0][0] == o[0]
        for a, b, name in zip(outs[0][1], o[1], "uvp"):
            assert np.array_equal(a, b), (name, np.abs(a - b).max())
