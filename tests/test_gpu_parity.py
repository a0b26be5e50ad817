"""GPU parity of every fine C-ABI entry point against the CPU oracle (same seeded inputs).

Element-wise kernels are compared bit for bit (FP64, no contraction, same operation order);
routines with global reductions to the tolerances stated at each test.  Every call goes
through the C ABI of libfugu2d_hip.so."""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, two_phase_case, droplet_c

pytestmark = pytest.mark.gpu

abi = pytest.importorskip("coupled_codefort_2d_amd.abi")

GRIDS = [
    dict(nx=24, ny=20, lx=1.2, ly=1.0, per_x=False, per_y=False),
    dict(nx=16, ny=24, lx=1.0, ly=1.5, per_x=True, per_y=False),
    dict(nx=20, ny=16, lx=1.0, ly=1.0, per_x=True, per_y=True),
]
IDS = ["walls", "periodic_x", "periodic_xy"]


def ctx_for(g: O.Grid, phys=None, sum_order=None):
    return abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y,
                       params=abi.FuguParams(*(phys or (1, 1, 1, 1, 0, 0, 0, 1, 1e12, 0.5))),
                       sum_order=sum_order)


def same(a, b, what=""):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        k = np.unravel_index(np.argmax(d), d.shape)
        raise AssertionError(f"{what}: not bit-identical, max |diff| {d.max():.3e} at {k} "
                             f"({a[k]!r} vs {b[k]!r})")


def close(a, b, rtol, what=""):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: max rel diff {err:.3e} > {rtol:.1e}"


def match(a, b, sum_order, rtol, what=""):
    """Reference summation order: bit-identical.  Tree order: within rtol relative."""
    if sum_order == abi.SUM_REFERENCE:
        same(a, b, what)
    else:
        close(a, b, rtol, what)


def grid(kw):
    return O.Grid(kw["nx"], kw["ny"], kw["lx"], kw["ly"], per_x=kw["per_x"], per_y=kw["per_y"])


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_halo_and_bcs_bitwise(kw):
    g = grid(kw)
    ctx = ctx_for(g)
    rng = np.random.default_rng(100)
    for kind in "uvp":
        a = rand(g, kind, rng)
        b = a.copy()
        O.halo(g, a, kind)
        ctx.halo(b, kind)
        same(b, a, f"halo {kind}")
    u, v = rand(g, "u", rng), rand(g, "v", rng)
    u2, v2 = u.copy(), v.copy()
    O.velocity_bcs(g, u, v, 1.25)
    ctx.velocity_bcs(u2, v2, 1.25)
    same(u2, u, "bcs u"); same(v2, v, "bcs v")
    O.velocity_bcs(g, u, v, 0.0)
    ctx.velocity_bcs(u2, v2, None)
    same(u2, u, "bcs u hom"); same(v2, v, "bcs v hom")


# the matvec also on grids whose extents are no multiple of its 64 x 8 LDS tiles
MV_GRIDS = GRIDS + [dict(nx=131, ny=67, lx=1.31, ly=0.67, per_x=False, per_y=False),
                    dict(nx=130, ny=70, lx=1.3, ly=0.7, per_x=True, per_y=False)]
MV_IDS = IDS + ["walls_131x67", "periodic_x_130x70"]


@pytest.mark.parametrize("kw", MV_GRIDS, ids=MV_IDS)
@pytest.mark.parametrize("variant", ["plain", "noadv", "penalty", "amp"])
@pytest.mark.parametrize("tile", ["8", "4", "0"])
def test_matvec_bitwise(monkeypatch, kw, variant, tile):
    """Both matvec kernels (operands from LDS tiles of 8 or 4 rows, or straight from the arrays)
    equal the oracle bit for bit, ghosts and wall faces included -- over an output array first
    filled with other values."""
    monkeypatch.setenv("FUGU_MV_TILE", tile)
    g = grid(kw)
    ctx = ctx_for(g)
    # the device output triplet holds other values before the matvec (the dot uploads into it)
    ctx.coupled_dot([np.full(g.shape(k), 7.25) for k in "uvp"], [np.full(g.shape(k), 0.5) for k in "uvp"])
    rng = np.random.default_rng(101)
    c, pr, ua, va = two_phase_case(g, rng)
    X = [rand(g, k, rng) for k in "uvp"]
    pen = (rng.uniform(0, 1e3, g.shape("u")), rng.uniform(0, 1e3, g.shape("v"))) if variant == "penalty" else None
    mdot = (rand(g, "u", rng), rand(g, "v", rng)) if variant == "amp" else None
    adv = variant != "noadv"
    Xo = [a.copy() for a in X]
    Yo = O.matvec(g, Xo, pr, ua, va, 3.7e-3, include_advection=adv, pen=pen, mdot=mdot)
    Xg = [a.copy() for a in X]
    Yg = [g.zeros(k) for k in "uvp"]
    ctx.matvec(Xg, Yg, pr, ua, va, 3.7e-3, include_advection=adv, pen=pen, mdot=mdot)
    for k in range(3):
        same(Yg[k], Yo[k], f"Y[{k}]")
        same(Xg[k], Xo[k], f"X[{k}] (halo in place)")


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_coupled_dot(kw, sum_order):
    """Reference order: the same double as the oracle's sequential loop (and as a plain Python
    loop over the owned points in the reference's order).  Tree order: within 1e-13."""
    g = grid(kw)
    ctx = ctx_for(g, sum_order=sum_order)
    rng = np.random.default_rng(102)
    a = [rand(g, k, rng) for k in "uvp"]
    b = [rand(g, k, rng) for k in "uvp"]
    ref = O.coupled_dot(g, a, b)
    got = ctx.coupled_dot(a, b)
    if sum_order == abi.SUM_REFERENCE:
        assert got == ref, (got, ref)
        # independent restatement: Python floats added one by one, u then v then p, j outer
        vol = g.dx * g.dy
        s = 0.0
        for k, kind in enumerate("uvp"):
            A, Bk = a[k][interior(g, kind)], b[k][interior(g, kind)]
            if kind == "u" and not g.per_x:
                A, Bk = A[:, 1:-1], Bk[:, 1:-1]        # i = 1..nx-1
            elif kind == "u":
                A, Bk = A[:, 1:], Bk[:, 1:]            # i = 1..nx
            if kind == "v" and not g.per_y:
                A, Bk = A[1:-1, :], Bk[1:-1, :]        # j = 1..ny-1
            elif kind == "v":
                A, Bk = A[1:, :], Bk[1:, :]
            for x, y in zip(A.ravel().tolist(), Bk.ravel().tolist()):
                s = s + x * y * vol
        assert got == s, (got, s)
    else:
        assert abs(got - ref) <= 1e-13 * (abs(ref) + np.sqrt(O.coupled_dot(g, a, a) * O.coupled_dot(g, b, b)))


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_mixture_bitwise(kw):
    g = grid(kw)
    ctx = ctx_for(g, (1000.0, 1.0, 1.0, 0.01, 0, 0, 0, 1, 1e12, 0.5))
    rng = np.random.default_rng(103)
    c = rng.uniform(-0.2, 1.2, g.shape("p"))
    ref = O.mixture(g, c, 1000.0, 1.0, 1.0, 0.01)
    out = {k: np.zeros_like(a) for k, a in ref.items()}
    ctx.mixture(c, out["rho_c"], out["rho_u"], out["rho_v"], out["kappa"], out["zeta"], out["eta"])
    for k in ref:
        same(out[k], ref[k], k)


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_jacobi_sweeps_bitwise(kw):
    g = grid(kw)
    ctx = ctx_for(g)
    rng = np.random.default_rng(104)
    c, pr, _, _ = two_phase_case(g, rng)
    for n_sweeps in (1, 3, 5):
        ru, rv = rand(g, "u", rng), rand(g, "v", rng)
        x0u, x0v = rand(g, "u", rng), rand(g, "v", rng)
        pen_u = rng.uniform(0, 10, g.shape("u"))
        a, b = x0u.copy(), x0u.copy()
        O.helmholtz_u(g, a, ru, pr, 2e-3, n_sweeps, pen_u)
        ctx.helmholtz_u(b, ru, pr, 2e-3, n_sweeps, pen_u)
        same(b, a, f"helmholtz_u x{n_sweeps}")
        a, b = x0v.copy(), x0v.copy()
        O.helmholtz_v(g, a, rv, pr, 2e-3, n_sweeps)
        ctx.helmholtz_v(b, rv, pr, 2e-3, n_sweeps)
        same(b, a, f"helmholtz_v x{n_sweeps}")


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_poisson_pcg(kw, sum_order):
    """Truncated 150-iteration PCG: same iteration count; the iterate bit-identical in the
    reference summation order, within 1e-10 relative in the tree order."""
    g = grid(kw)
    ctx = ctx_for(g, sum_order=sum_order)
    rng = np.random.default_rng(105)
    c, pr, _, _ = two_phase_case(g, rng, rho=(100.0, 1.0))
    if g.per_x or g.per_y:
        O.halo(g, pr["rho_u"], "u"); O.halo(g, pr["rho_v"], "v")
    rhs = rand(g, "p", rng)
    if g.per_x or g.per_y:
        rhs[interior(g, "p")] -= rhs[interior(g, "p")].mean()
    for tol, phi0 in ((1e-5, None), (1e-9, "rand")):
        phi_o = g.zeros("p") if phi0 is None else rand(g, "p", rng)
        phi_g = phi_o.copy()
        it_o, res_o = O.poisson_pcg(g, phi_o, rhs, pr["rho_u"], pr["rho_v"], tol=tol, max_iter=150)
        it_g, res_g = ctx.poisson_pcg(phi_g, rhs, pr["rho_u"], pr["rho_v"], tol=tol, max_iter=150)
        assert it_g == it_o
        if sum_order == abi.SUM_REFERENCE:
            assert res_g == res_o, (res_g, res_o)
        assert abs(res_g - res_o) <= 1e-12 * max(res_o, 1.0)
        match(phi_g, phi_o, sum_order, 1e-10, "phi")
    # zero rhs early exit
    phi = np.ones(g.shape("p"))
    it, res = ctx.poisson_pcg(phi, g.zeros("p"), pr["rho_u"], pr["rho_v"])
    assert (it, res) == (0, 0.0) and np.all(phi == 0)


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_pcd_precond(kw, sum_order):
    g = grid(kw)
    ctx = ctx_for(g, sum_order=sum_order)
    rng = np.random.default_rng(106)
    c, pr, ua, va = two_phase_case(g, rng)
    if g.per_x or g.per_y:
        for k, kind in (("rho_u", "u"), ("rho_v", "v"), ("rho_c", "p")):
            O.halo(g, pr[k], kind)
    Z = [rand(g, k, rng) for k in "uvp"]
    if g.per_x or g.per_y:
        Z[2][interior(g, "p")] -= Z[2][interior(g, "p")].mean()
    Ro = O.pcd(g, Z, pr, ua, va, 1e-3)
    Rg = [g.zeros(k) for k in "uvp"]
    ctx.pcd(Z, Rg, pr, ua, va, 1e-3)
    for k in range(3):
        match(Rg[k], Ro[k], sum_order, 1e-10, f"R[{k}]")


@pytest.mark.parametrize("kw", GRIDS[:2], ids=IDS[:2])
@pytest.mark.parametrize("order,consistent", [(1, False), (2, False), (2, True)],
                         ids=["bicgstab", "bicgstab2_ref", "bicgstab2_bc"])
def test_krylov_solve(kw, order, consistent, sum_order):
    """bicgstab_solve / bicgstab2_solve: same iteration count; the solution (u, v, p, ghosts
    included) and the residual bit-identical in the reference summation order, within 1e-10
    relative in the tree order."""
    g = grid(kw)
    ctx = ctx_for(g, sum_order=sum_order)
    rng = np.random.default_rng(107)
    c, pr, ua, va = two_phase_case(g, rng, rho=(10.0, 1.0), mu=(1.0, 0.1))
    B = [rand(g, "u", rng), rand(g, "v", rng), g.zeros("p")]
    X0 = [0.01 * rand(g, k, rng) for k in "uvp"]
    for tol, lid in ((1e-6, None), (1e-9, 1.0)):
        Xo = [a.copy() for a in X0]
        Xg = [a.copy() for a in X0]
        it_o, res_o = O.krylov(g, Xo, B, pr, ua, va, 5e-3, tol=tol, max_iter=40, order=order, u_lid=lid,
                               consistent=consistent)
        it_g, res_g = ctx.krylov(Xg, B, pr, ua, va, 5e-3, tol=tol, max_iter=40, order=order, u_lid=lid,
                                 flags=abi.B2_CONSISTENT_BCS if consistent else 0)
        assert it_g == it_o, (it_g, it_o)
        if sum_order == abi.SUM_REFERENCE:
            assert res_g == res_o, (res_g, res_o)
        for k in range(3):
            match(Xg[k], Xo[k], sum_order, 1e-10, f"X[{k}]")


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
@pytest.mark.parametrize("limiter", [0, 1], ids=["van_leer", "superbee"])
def test_vof_bitwise(kw, limiter):
    g = grid(kw)
    ctx = ctx_for(g)
    ctx.set_vof_limiter(limiter)
    O.set_vof_limiter(limiter)
    try:
        rng = np.random.default_rng(108)
        c = np.clip(droplet_c(g, 0.3) + 0.2 * rng.standard_normal(g.shape("p")), -0.05, 1.05)
        u, v = rand(g, "u", rng), rand(g, "v", rng)
        co, fxo, fyo = O.vof_advection(g, c, u, v, 0.01, want_flux=True)
        cg, fxg, fyg = g.zeros("p"), g.zeros("u"), g.zeros("v")
        ctx.vof_advection(c, cg, u, v, 0.01, fxg, fyg)
        same(cg, co, "C_out"); same(fxg, fxo, "flux_x"); same(fyg, fyo, "flux_y")
    finally:
        O.set_vof_limiter(0)


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_svof_curvature_csf_bitwise(kw):
    g = grid(kw)
    ctx = ctx_for(g, (1, 1, 1, 1, 0.7, 0, 0, 1, 1e12, 0.5))
    c = droplet_c(g, 0.3)
    for it_max, tol in ((150, 1e-7), (7, 1e-12)):
        cs_o, it_o = O.svof(g, c, 1.5 * g.dx, it_max, tol)
        cs_g = g.zeros("p")
        it_g = ctx.svof(c, cs_g, 1.5 * g.dx, it_max, tol)
        assert it_g == it_o
        same(cs_g, cs_o, "C_smooth")
    kc_o = O.curvature(g, cs_o)
    kc_g = g.zeros("p")
    ctx.curvature(cs_o, kc_g)
    same(kc_g, kc_o, "kappa")
    fu_o, fv_o = O.csf(g, c, kc_o, 0.7)
    fu_g, fv_g = g.zeros("u"), g.zeros("v")
    ctx.csf(c, kc_o, fu_g, fv_g)
    same(fu_g, fu_o, "f_u"); same(fv_g, fv_o, "f_v")


SVOF_GRIDS = [
    dict(nx=130, ny=70, lx=1.3, ly=0.7, per_x=False, per_y=False),
    dict(nx=128, ny=96, lx=1.0, ly=0.75, per_x=True, per_y=False),
    dict(nx=96, ny=80, lx=1.2, ly=1.0, per_x=True, per_y=True),
    dict(nx=40, ny=150, lx=0.4, ly=1.5, per_x=False, per_y=True),
]


@pytest.mark.parametrize("kw", SVOF_GRIDS, ids=["walls", "periodic_x", "periodic_xy", "periodic_y"])
def test_svof_blocked_sweeps_bitwise(kw):
    """The SVOF sweeps run several iterations per launch on tiles with halos; their iterate and
    iteration count equal the oracle's one-sweep-at-a-time smoothing bit for bit wherever the
    stopping test fires: inside a block (the block is recomputed up to that iteration), at a
    block's last iteration, or never (iteration caps that are and are not block multiples),
    on multi-tile grids with partial tiles, walls and periodic wrap."""
    g = grid(kw)
    ctx = ctx_for(g)
    rng = np.random.default_rng(110)
    c = np.clip(droplet_c(g, 0.3) + 0.3 * rng.standard_normal(g.shape("p")), -0.1, 1.1)
    seen = set()
    for it_max, tol in ((150, 1e-3), (150, 3e-4), (150, 1e-4), (150, 3e-5), (150, 1e-5), (150, 1e-7),
                        (1, 1e-12), (8, 1e-12), (9, 1e-12), (16, 1e-12), (23, 1e-12)):
        cs_o, it_o = O.svof(g, c, 1.5 * g.dx, it_max, tol)
        cs_g = g.zeros("p")
        it_g = ctx.svof(c, cs_g, 1.5 * g.dx, it_max, tol)
        assert it_g == it_o, (it_max, tol, it_g, it_o)
        same(cs_g, cs_o, f"C_smooth (it_max {it_max}, tol {tol:g}, {it_o} iterations)")
        seen.add(it_o % 8)
    assert len(seen) >= 3, seen   # stops at several positions inside a block of eight
    ctx.close()


def test_adaptive_dt():
    g = O.Grid(32, 24, 1.0, 0.75)
    ph = O.OgPhys(1000.0, 1.0, 1.0, 0.01, 0.5, 0.0, -9.81, 1, 1e12, 0.25)
    ctx = ctx_for(g, (1000.0, 1.0, 1.0, 0.01, 0.5, 0.0, -9.81, 1, 1e12, 0.25))
    rng = np.random.default_rng(109)
    u, v = rand(g, "u", rng, 3.0), rand(g, "v", rng, 2.0)
    assert ctx.adaptive_dt(u, v, 1.0) == O.adaptive_dt(g, ph, u, v, 1.0)
