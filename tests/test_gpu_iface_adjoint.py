"""The moving-interface adjoint on the GPU (no counterpart in the reference: it has no adjoint).

* The GPU's transposes of the interface chain (fugu_vof_advection_transpose, fugu_svof_transpose,
  fugu_curvature_transpose, fugu_csf_transpose) against the oracle's (og_*_T, themselves pinned by
  central differences in tests/test_oracle_interface_adjoint.py) on wall and periodic grids, smooth
  and sharp interfaces (the sharp one exercises the clamps, the flat |den| <= 1e-12 faces and the
  limiter at r <= 0 with the same one-sided conventions) and both limiters: to 1e-12 of each
  array's largest value.
* The gradient with FUGU_ADJ_MOVING_INTERFACE (VOF, SVOF, curvature and CSF transposed in the
  reverse sweep) against central differences of J on the GPU, where J recomputes the interface
  from C_0 at every evaluation (the map the moving-interface gradient differentiates): sigma, a
  random per-cell viscosity direction and a random per-cell density direction, with the forward
  and adjoint solves to 1e-12.  The frozen-interface gradient differentiates another map; its
  distance from the same differences is reported beside it."""
import numpy as np
import pytest

import pyoracle as O
import adjoint_ref as AR

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")


def ctx_for(g):
    return abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y)


def close(a, b, tol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


def states(g, kind, rng):
    X, Y = g.coords("p")
    r = np.hypot(X - 0.5 * g.lx, Y - 0.5 * g.ly)
    if kind == "sharp":
        c = (r <= 0.3 * min(g.lx, g.ly)).astype(np.float64)
    else:
        c = 0.5 * (1.0 - np.tanh((r - 0.3 * min(g.lx, g.ly)) / (1.5 * g.dx)))
    u = 0.5 * rng.standard_normal(g.shape("u"))
    v = 0.5 * rng.standard_normal(g.shape("v"))
    O.velocity_bcs(g, u, v, 0.0)
    O.halo(g, u, "u"); O.halo(g, v, "v"); O.halo(g, c, "p")
    return c, u, v


GRIDS = [(24, 20, False, False), (20, 24, True, True), (24, 16, True, False), (16, 24, False, True)]


@pytest.mark.parametrize("nx,ny,per_x,per_y", GRIDS)
@pytest.mark.parametrize("kind", ["smooth", "sharp"])
@pytest.mark.parametrize("limiter", [0, 1])
def test_interface_transposes_match_oracle(nx, ny, per_x, per_y, kind, limiter):
    g = O.Grid(nx, ny, 1.0, ny / nx, per_x=per_x, per_y=per_y)
    rng = np.random.default_rng(100 + nx + ny + 7 * per_x + 3 * per_y + limiter)
    c, u, v = states(g, kind, rng)
    dt = 0.3 * g.dx
    ctx = ctx_for(g)
    ctx.set_vof_limiter(limiter)
    O.set_vof_limiter(limiter)
    try:
        # VOF
        cb = rng.standard_normal(g.shape("p"))
        ref = O.vof_advection_T(g, c, u, v, dt, cb)
        got = [np.zeros_like(a) for a in ref]
        ctx.vof_advection_T(c, u, v, dt, cb, *got)
        for a, b, name in zip(got, ref, ("cbar", "ubar", "vbar")):
            assert close(a, b), (name, np.abs(a - b).max(), np.abs(b).max())
    finally:
        O.set_vof_limiter(0)
    # SVOF at the forward's sweep count
    delta = 1.5 * min(g.dx, g.dy)
    sb = rng.standard_normal(g.shape("p"))
    got = np.zeros(g.shape("p"))
    it = ctx.svof_T(c, delta, sb, got, 150, 1e-7)
    assert it == O.svof(g, c, delta, 150, 1e-7)[1]
    assert close(got, O.svof_T(g, sb, delta, it))
    # curvature of the filtered field
    cs = O.svof(g, c, delta, 150, 1e-7)[0]
    kb = rng.standard_normal(g.shape("p"))
    got = np.zeros(g.shape("p"))
    ctx.curvature_T(cs, kb, got)
    assert close(got, O.curvature_T(g, cs, kb))
    # CSF
    kc = O.curvature(g, cs)
    fub, fvb = rng.standard_normal(g.shape("u")), rng.standard_normal(g.shape("v"))
    refc, refk = O.csf_T(g, c, kc, 0.7, fub, fvb)
    gc, gk = np.zeros(g.shape("p")), np.zeros(g.shape("p"))
    ctx.csf_T(c, kc, 0.7, fub, fvb, gc, gk)
    assert close(gc, refc) and close(gk, refk)
    ctx.close()


def smooth_case(per_x=False, nsteps=4):
    g = O.Grid(16, 20, 1.0, 1.25, per_x=per_x)
    rng = np.random.default_rng(71)
    X, Y = g.coords("p")
    r = np.hypot(X - 0.45, Y - 0.55)
    c = 0.5 * (1.0 - np.tanh((r - 0.25) / (1.5 * g.dx)))
    mu = (0.05 * c + 0.5 * (1.0 - c)) * np.exp(0.3 * rng.standard_normal(g.shape("p")))
    rho = (20.0 * c + 2.0 * (1.0 - c)) * np.exp(0.2 * rng.standard_normal(g.shape("p")))
    u0 = 0.3 * rng.standard_normal(g.shape("u"))
    v0 = 0.3 * rng.standard_normal(g.shape("v"))
    return AR.Case(g, mu, rho, 0.5, 0.2, -1.0, 4e-3, nsteps, time_order=2, y0_frac=0.6, u0=u0, v0=v0, c0=c)


@pytest.mark.parametrize("per_x", [False, True])
def test_moving_interface_gradient_central_differences(per_x):
    case = smooth_case(per_x)
    g = case.g
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y,
                      params=abi.FuguParams(1.0, 1.0, 1.0, 1.0, case.sigma, case.gx, case.gy, 2, 1e12, 0.5))
    opts = lambda fl: abi.FuguAdjointOpts(case.nsteps, case.dt, case.y0, 1e-12, 400, 30, 1, fl)
    grads = {}
    for name, fl in (("moving", abi.ADJ_MOVING_INTERFACE), ("frozen", 0)):
        gm, gr = np.zeros_like(case.mu), np.zeros_like(case.rho)
        J, gs, st, rc = ctx.adjoint_gradient(opts(fl), case.mu, case.rho, case.sigma, case.c0, case.u0, case.v0,
                                             gm, gr)
        assert rc == 0, (name, st.max_forward_res, st.max_adjoint_res)
        grads[name] = (J, gs, gm, gr)
    assert grads["moving"][0] == grads["frozen"][0]   # the same forward

    def Jat(mu, rho, sigma):   # the interface recomputed from C_0 at every evaluation
        return ctx.adjoint_gradient(opts(abi.ADJ_FORWARD_ONLY), mu, rho, sigma, case.c0, case.u0, case.v0)[0]

    rng = np.random.default_rng(73)
    eps = 1e-4
    dmu = case.mu * rng.standard_normal(case.mu.shape)
    drho = case.rho * rng.standard_normal(case.rho.shape)
    fd = {
        "sigma": (Jat(case.mu, case.rho, case.sigma * (1 + eps)) - Jat(case.mu, case.rho, case.sigma * (1 - eps)))
        / (2 * eps * case.sigma),
        "mu": (Jat(case.mu + eps * dmu, case.rho, case.sigma) - Jat(case.mu - eps * dmu, case.rho, case.sigma))
        / (2 * eps),
        "rho": (Jat(case.mu, case.rho + eps * drho, case.sigma) - Jat(case.mu, case.rho - eps * drho, case.sigma))
        / (2 * eps),
    }
    for name in ("moving", "frozen"):
        _, gs, gm, gr = grads[name]
        ad = {"sigma": gs, "mu": float(np.sum(gm * dmu)), "rho": float(np.sum(gr * drho))}
        err = {k: abs(ad[k] - fd[k]) / abs(fd[k]) for k in fd}
        print(f"{name}: " + ", ".join(f"d{k}: adjoint {ad[k]:.9e} FD {fd[k]:.9e} rel {err[k]:.2e}" for k in fd),
              flush=True)
        if name == "moving":
            for k in fd:
                assert err[k] <= 1e-6, (k, ad[k], fd[k])
    ctx.close()
