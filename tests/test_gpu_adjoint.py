"""GPU adjoint through the C ABI against the CPU adjoint references (oracle/fugu2d_adjoint.c,
oracle/adjoint_ref.py):

- A_eff^T (fugu_matvec_transpose): equal to the oracle's to rounding, and the dot-product
  identity with the oracle's forward A_eff;
- the transposed linearised PCD (fugu_apply_pcd_transpose, tight pressure solve): equal to the
  oracle's; the dot-product identity with the oracle's forward linearised PCD;
- the adjoint solve (fugu_adjoint_solve): against the FP64 sparse transpose solve of the
  assembled operator;
- the per-cell property path: bit for bit;
- the multi-step gradient (fugu_adjoint_gradient) for per-cell mu, rho and sigma: against the
  CPU adjoint reference, and against central finite differences of the GPU's own forward map
  (interface held at the recorded trajectory) at 1e-6 relative.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import pyoracle as O
import adjoint_ref as AR
from cases import rand, two_phase_case, droplet_c

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")

GRIDS = [
    dict(nx=24, ny=20, lx=1.2, ly=1.0, per_x=False, per_y=False),
    dict(nx=16, ny=24, lx=1.0, ly=1.5, per_x=True, per_y=False),
    dict(nx=24, ny=16, lx=1.0, ly=1.0, per_x=False, per_y=True),
    dict(nx=20, ny=16, lx=1.0, ly=1.0, per_x=True, per_y=True),
]
IDS = ["walls", "periodic_x", "periodic_y", "periodic_xy"]


def grid(kw):
    return O.Grid(kw["nx"], kw["ny"], kw["lx"], kw["ly"], per_x=kw["per_x"], per_y=kw["per_y"])


def ctx_for(g, phys=None):
    return abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y,
                       params=abi.FuguParams(*(phys or (1, 1, 1, 1, 0, 0, 0, 1, 1e12, 0.5))))


def rtrip(g, rng):
    return O.mask_owned(g, [rand(g, "u", rng), rand(g, "v", rng), rand(g, "p", rng)])


def zeros3(g):
    return [g.zeros("u"), g.zeros("v"), g.zeros("p")]


def odot(g, a, b):
    return O.coupled_dot(g, a, b) / (g.dx * g.dy)


def rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(max(np.abs(y).max() for y in b), 1e-300)


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
@pytest.mark.parametrize("adv", [True, False])
def test_matvec_transpose(kw, adv):
    g = grid(kw)
    rng = np.random.default_rng(41)
    c, pr, ua, va = two_phase_case(g, rng)
    y = rtrip(g, rng)
    ref = O.matvec_T(g, y, pr, ua, va, 1e-2, adv)
    ctx = ctx_for(g)
    out = zeros3(g)
    ctx.matvec_transpose(y, out, pr, ua, va, 1e-2, adv)
    assert rel(out, ref) < 1e-13, rel(out, ref)
    x = rtrip(g, rng)
    ax = O.matvec_eff(g, x, pr, ua, va, 1e-2, adv)
    lhs, rhs = odot(g, ax, y), odot(g, x, out)
    assert abs(lhs - rhs) < 1e-13 * np.sqrt(odot(g, ax, ax) * odot(g, y, y))
    ctx.close()


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
@pytest.mark.parametrize("visc", [0.0, 1.0])
def test_pcd_transpose(kw, visc):
    g = grid(kw)
    rng = np.random.default_rng(43)
    c, pr, ua, va = two_phase_case(g, rng)
    y = rtrip(g, rng)
    ref = O.pcd_lin_T(g, y, pr, ua, va, 1e-2, ptol=1e-14, pmax=5000, visc=visc)
    ctx = ctx_for(g)
    out = zeros3(g)
    ctx.pcd_transpose(y, out, pr, ua, va, 1e-2, ptol=1e-14, pmax=5000, visc=visc)
    assert rel(out, ref) < 1e-9, rel(out, ref)
    x = rtrip(g, rng)
    px = O.pcd_lin(g, x, pr, ua, va, 1e-2, ptol=1e-14, pmax=5000, visc=visc)
    lhs, rhs = odot(g, px, y), odot(g, x, out)
    assert abs(lhs - rhs) < 1e-9 * np.sqrt(odot(g, px, px) * odot(g, y, y))
    # the multigrid pressure block (the adjoint solve's preconditioner) runs and is finite
    ctx.pcd_transpose(y, out, pr, ua, va, 1e-2, ptol=0.0, pmax=1, visc=visc)
    assert all(np.isfinite(a).all() for a in out)
    ctx.close()


@pytest.mark.parametrize("kw", GRIDS[:2], ids=IDS[:2])
def test_adjoint_solve_against_sparse_transpose(kw):
    g = grid(dict(kw, nx=16, ny=16))
    rng = np.random.default_rng(47)
    c, pr, ua, va = two_phase_case(g, rng, rho=(100.0, 1.0), mu=(0.5, 0.01))
    dt = 1e-3
    lay = AR.Layout(g)
    A = AR.assemble(g, lay, pr, ua, va, dt)
    G = rtrip(g, rng)
    G[2][...] = 0.0
    sol = spla.spsolve(AR.bordered(A, lay).T.tocsc(), np.concatenate([lay.flat(G), [0.0]]))
    ref = lay.trip(sol[:lay.n])
    ctx = ctx_for(g)
    L = zeros3(g)
    it, res, rc = ctx.adjoint_solve(L, G, pr, ua, va, dt, tol=1e-11, max_iter=400)
    assert rc == 0 and res <= 1e-11, (it, res)
    assert rel(L[:2], ref[:2]) < 1e-8, rel(L[:2], ref[:2])
    r = O.matvec_T(g, L, pr, ua, va, dt, True)
    r = [a - b for a, b in zip(r, G)]
    assert np.sqrt(odot(g, r, r) / odot(g, G, G)) < 1e-10
    ctx.close()


def test_cell_properties_bitwise():
    g = O.Grid(20, 16, 1.0, 0.8, per_x=True)
    rng = np.random.default_rng(53)
    mu = rng.uniform(0.01, 2.0, g.shape("p"))
    rho = rng.uniform(1.0, 1000.0, g.shape("p"))
    ref = O.cell_properties(g, mu, rho)
    ctx = ctx_for(g)
    out = {k: np.zeros_like(v) for k, v in ref.items()}
    ctx.cell_properties(mu, rho, out)
    for k in O.PROP_NAMES:
        assert np.array_equal(out[k], ref[k]), k
    ctx.close()


def make_case(per_x=False, time_order=2, nsteps=4):
    g = O.Grid(12, 16, 1.0, 1.25, per_x=per_x)
    rng = np.random.default_rng(29)
    c = np.maximum(droplet_c(g, R=0.22, xc=0.45, yc=0.55), 0.0)
    mu = np.where(c > 0.5, 0.05, 0.5) * np.exp(0.3 * rng.standard_normal(g.shape("p")))
    rho = np.where(c > 0.5, 20.0, 2.0) * np.exp(0.2 * rng.standard_normal(g.shape("p")))
    u0 = 0.05 * rng.standard_normal(g.shape("u"))
    v0 = 0.05 * rng.standard_normal(g.shape("v"))
    return AR.Case(g, mu, rho, 0.3, 0.2, -1.0, 2e-3, nsteps, time_order=time_order, y0_frac=0.6,
                   u0=u0, v0=v0, c0=c)


def gpu_ctx(case):
    g = case.g
    return abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y,
                       params=abi.FuguParams(1.0, 1.0, 1.0, 1.0, case.sigma, case.gx, case.gy, case.time_order,
                                             1e12, 0.5))


def gpu_opts(case, flags=0, tol=1e-12):
    return abi.FuguAdjointOpts(case.nsteps, case.dt, case.y0, tol, 400, 30, 1, flags)


MIXED_ADJ = abi.ADJ_MIXED_ADJOINT


@pytest.mark.parametrize("per_x,time_order,flags", [(False, 2, 0), (True, 2, 0), (False, 1, 0),
                                                    (False, 2, abi.ADJ_PRECOND_VMG), (True, 2, abi.ADJ_PRECOND_VMG),
                                                    (False, 2, MIXED_ADJ), (True, 2, MIXED_ADJ),
                                                    (True, 2, MIXED_ADJ | abi.ADJ_PRECOND_VMG),
                                                    (False, 1, MIXED_ADJ | abi.ADJ_MIXED_FORWARD | abi.ADJ_PRECOND_VMG)],
                         ids=["walls_bdf2", "periodic_x_bdf2", "walls_bdf1", "walls_bdf2_vmg", "periodic_x_bdf2_vmg",
                              "walls_bdf2_mixed_adj", "periodic_x_bdf2_mixed_adj", "periodic_x_bdf2_vmg_mixed_adj",
                              "walls_bdf1_vmg_mixed_both"])
def test_adjoint_gradient_against_cpu_reference(per_x, time_order, flags):
    """The GPU gradient against the CPU adjoint reference: FP64 solves, or (MIXED_ADJ) adjoint
    solves by FP32 Arnoldi cycles of the transposed operator and PCD refined in FP64 to the same
    1e-12 true residual -- with the Jacobi or the multigrid velocity blocks, walls or periodic."""
    case = make_case(per_x, time_order)
    J, g_mu, g_rho, g_sigma, _ = AR.gradient(case)
    ctx = gpu_ctx(case)
    gm, gr = np.zeros_like(case.mu), np.zeros_like(case.rho)
    Jg, gs, st, rc = ctx.adjoint_gradient(gpu_opts(case, flags), case.mu, case.rho, case.sigma, case.c0, case.u0,
                                          case.v0, gm, gr)
    assert rc == 0, (st.max_forward_res, st.max_adjoint_res)
    assert abs(Jg - J) <= 1e-10 * abs(J), (Jg, J)
    assert np.abs(gm - g_mu).max() <= 1e-7 * np.abs(g_mu).max()
    assert np.abs(gr - g_rho).max() <= 1e-7 * np.abs(g_rho).max()
    assert abs(gs - g_sigma) <= 1e-7 * abs(g_sigma)
    ctx.close()


def test_adjoint_gradient_central_differences_on_gpu():
    case = make_case()
    ctx = gpu_ctx(case)
    gm, gr = np.zeros_like(case.mu), np.zeros_like(case.rho)
    J, gs, st, rc = ctx.adjoint_gradient(gpu_opts(case), case.mu, case.rho, case.sigma, case.c0, case.u0,
                                         case.v0, gm, gr)
    assert rc == 0
    rng = np.random.default_rng(59)
    fwd = gpu_opts(case, abi.ADJ_FORWARD_ONLY | abi.ADJ_REUSE_INTERFACE)

    def Jat(mu, rho, sigma):
        return ctx.adjoint_gradient(fwd, mu, rho, sigma, None, case.u0, case.v0)[0]

    eps = 1e-4
    dmu = case.mu * rng.standard_normal(case.mu.shape)
    drho = case.rho * rng.standard_normal(case.rho.shape)
    f = (Jat(case.mu + eps * dmu, case.rho, case.sigma) - Jat(case.mu - eps * dmu, case.rho, case.sigma)) / (2 * eps)
    assert abs(f - np.sum(gm * dmu)) <= 1e-6 * abs(f), (f, np.sum(gm * dmu))
    f = (Jat(case.mu, case.rho + eps * drho, case.sigma) - Jat(case.mu, case.rho - eps * drho, case.sigma)) / (2 * eps)
    assert abs(f - np.sum(gr * drho)) <= 1e-6 * abs(f), (f, np.sum(gr * drho))
    f = (Jat(case.mu, case.rho, case.sigma + eps) - Jat(case.mu, case.rho, case.sigma - eps)) / (2 * eps)
    assert abs(f - gs) <= 1e-6 * abs(f), (f, gs)
    ctx.close()


def test_checkpointed_horizon_matches_resident_horizon():
    """Checkpointing (restart states every 3rd step, segments recomputed in the reverse sweep)
    gives the resident horizon's J bit for bit and its gradients to rounding (the coefficient
    gradients accumulate with atomics, so their order is not fixed)."""
    case = make_case(nsteps=7)
    ctx = gpu_ctx(case)
    out = {}
    for k in (0, 3):
        gm, gr = np.zeros_like(case.mu), np.zeros_like(case.rho)
        o = gpu_opts(case, abi.ADJ_PRECOND_VMG)
        o.checkpoint_every = k
        J, gs, st, rc = ctx.adjoint_gradient(o, case.mu, case.rho, case.sigma, case.c0, case.u0, case.v0, gm, gr)
        assert rc == 0, (k, st.max_forward_res, st.max_adjoint_res)
        out[k] = (J, gs, gm, gr)
    (J0, s0, m0, r0), (J3, s3, m3, r3) = out[0], out[3]
    assert J3 == J0
    assert abs(s3 - s0) <= 1e-12 * abs(s0)
    assert np.abs(m3 - m0).max() <= 1e-12 * np.abs(m0).max()
    assert np.abs(r3 - r0).max() <= 1e-12 * np.abs(r0).max()
    # reusing the interface needs the whole horizon
    fwd = gpu_opts(case, abi.ADJ_FORWARD_ONLY | abi.ADJ_REUSE_INTERFACE)
    fwd.checkpoint_every = 3
    with pytest.raises(Exception, match="whole horizon"):
        ctx.adjoint_gradient(fwd, case.mu, case.rho, case.sigma, None, case.u0, case.v0)
    ctx.close()


def test_mixed_forward_gradient_matches_fp64_forward():
    """Forward solves by the mixed form (FP32 Arnoldi cycles, FP64 restarts) to the same 1e-12
    true residual: J and the gradients agree with the FP64 forward to the solve tolerance."""
    case = make_case(nsteps=4)
    ctx = gpu_ctx(case)
    out = {}
    for fl in (0, abi.ADJ_MIXED_FORWARD):
        gm, gr = np.zeros_like(case.mu), np.zeros_like(case.rho)
        J, gs, st, rc = ctx.adjoint_gradient(gpu_opts(case, fl | abi.ADJ_PRECOND_VMG), case.mu, case.rho, case.sigma,
                                             case.c0, case.u0, case.v0, gm, gr)
        assert rc == 0, (fl, st.max_forward_res, st.max_adjoint_res)
        out[fl] = (J, gs, gm, gr)
    (J0, s0, m0, r0), (J1, s1, m1, r1) = out[0], out[abi.ADJ_MIXED_FORWARD]
    assert abs(J1 - J0) <= 1e-9 * abs(J0), (J1, J0)
    assert abs(s1 - s0) <= 1e-7 * abs(s0)
    assert np.abs(m1 - m0).max() <= 1e-7 * np.abs(m0).max()
    assert np.abs(r1 - r0).max() <= 1e-7 * np.abs(r0).max()
    ctx.close()
