"""GPU parity of advance_coupled_step (src/mod_solver_coupled.f90:137-380) over several steps
against the oracle, in both summation orders of the solver reductions: in the reference's
sequential order every state field (ghosts included) is bit-identical after every step; in the
tree order u, v, p agree within 1e-10 relative and C within 1e-12.  Same Krylov iteration
counts in both."""
import numpy as np
import pytest

import pyoracle as O
from cases import droplet_c, phys

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


STATE = ("u", "v", "p", "c", "u_prev", "v_prev", "rho_c", "rho_u", "rho_v", "kappa", "zeta", "eta",
         "kappa_curv", "f_u", "f_v")


def run_pair(g, ph, init, n_steps, dt, opts_kw, phi_s=None, sum_order=0):
    st = O.State(g)
    init(st)
    O.halo(g, st.c, "p")
    st.init_properties(ph)
    sg = st.copy()
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly, g.per_x, g.per_y,
                      params=abi.FuguParams(ph.rho1, ph.rho2, ph.mu1, ph.mu2, ph.sigma, ph.gx, ph.gy,
                                            ph.time_order, ph.alpha_penalty, ph.cfl_target),
                      sum_order=sum_order)
    okw = dict(opts_kw)
    opts = abi.step_opts(**okw)
    for step in range(1, n_steps + 1):
        it_o, res_o = O.advance_step(st, ph, dt, step, phi_s=phi_s, **okw)
        it_g, res_g, rc = ctx.advance_step(sg, dt, step, opts, phi_s=phi_s)
        assert it_g == it_o, (step, it_g, it_o)
        if sum_order == abi.SUM_REFERENCE:
            assert res_g == res_o, (step, res_g, res_o)
            for name in STATE:
                a, b = getattr(sg, name), getattr(st, name)
                assert np.array_equal(a, b), (step, name, rel(a, b))
            continue
        for name, tol in (("u", 1e-10), ("v", 1e-10), ("p", 1e-10), ("c", 1e-12), ("u_prev", 1e-10),
                          ("kappa_curv", 1e-9), ("rho_u", 1e-12)):
            a, b = getattr(sg, name), getattr(st, name)
            if np.abs(b).max() > 0:
                assert rel(a, b) <= tol, (step, name, rel(a, b))
    return st, sg


@pytest.mark.parametrize("order", [1, 2])
def test_step_static_droplet(order, sum_order):
    """Static droplet (rho 1000/1, mu 1/0.01, sigma 1), BDF2, SVOF + CSF active."""
    g = O.Grid(24, 24)
    ph = phys(rho1=1000.0, rho2=1.0, mu1=1.0, mu2=0.01, sigma=1.0, time_order=2)

    def init(st):
        st.c[...] = droplet_c(g, 0.25)
    run_pair(g, ph, init, 3, 1e-4, dict(use_bicgstab2=(order == 2), tol=1e-8, max_iter=60), sum_order=sum_order)


def test_step_lid_cavity(sum_order):
    g = O.Grid(20, 20)
    ph = phys(rho1=1.0, rho2=1.0, mu1=0.01, mu2=0.01, time_order=1)

    def init(st):
        st.c[...] = 1.0
    run_pair(g, ph, init, 3, 1e-2, dict(use_bicgstab2=False, u_lid=1.0, tol=1e-8, max_iter=80),
             sum_order=sum_order)


def test_step_amp_penalty_periodic(sum_order):
    """Periodic-x channel with a Darcy-penalised cylinder and consistent (AMP) mass fluxes."""
    g = O.Grid(24, 16, 1.5, 1.0, per_x=True)
    ph = phys(rho1=10.0, rho2=1.0, mu1=0.1, mu2=0.05, gx=1.0, gy=-1.0, time_order=2, alpha_penalty=1e4)
    Xu, Yu = g.coords("u")
    Xv, Yv = g.coords("v")
    psu = (np.hypot(Xu - 0.75, Yu - 0.5) <= 0.15).astype(np.float64)
    psv = (np.hypot(Xv - 0.75, Yv - 0.5) <= 0.15).astype(np.float64)

    def init(st):
        X, Y = g.coords("p")
        st.c[...] = (Y <= 0.5).astype(np.float64)
    run_pair(g, ph, init, 3, 2e-3, dict(use_bicgstab2=False, tol=1e-8, max_iter=80, use_amp=True),
             phi_s=(psu, psv), sum_order=sum_order)


def test_step_superbee_bitwise():
    """The opt-in Superbee limiter through the whole step (droplet in a shear), bit-identical to the
    oracle's Superbee step in the reference summation order."""
    g = O.Grid(24, 24)
    ph = phys(rho1=1000.0, rho2=1.0, mu1=1.0, mu2=0.01, sigma=1.0, time_order=2)
    O.set_vof_limiter(1)
    O.set_sum_mode(0)
    try:
        st = O.State(g)
        st.c[...] = droplet_c(g, 0.25)
        Xu, Yu = g.coords("u")
        st.u[...] = 0.5 * (Yu - 0.5)
        O.velocity_bcs(g, st.u, st.v)
        st.init_properties(ph)
        sg = st.copy()
        ctx = abi.Context(24, 24, params=abi.FuguParams(1000.0, 1.0, 1.0, 0.01, 1.0, 0.0, 0.0, 2, 1e12, 0.5))
        ctx.set_vof_limiter(abi.LIMITER_SUPERBEE)
        okw = dict(use_bicgstab2=False, tol=1e-8, max_iter=60)
        for step in (1, 2, 3):
            it_o, res_o = O.advance_step(st, ph, 1e-3, step, **okw)
            it_g, res_g, rc = ctx.advance_step(sg, 1e-3, step, abi.step_opts(**okw))
            assert it_g == it_o and res_g == res_o
            for name in STATE:
                assert np.array_equal(getattr(sg, name), getattr(st, name)), (step, name)
        ctx.close()
    finally:
        O.set_vof_limiter(0)


@pytest.mark.parametrize("limiter", [0, 1], ids=["van_leer", "superbee"])
def test_step_volume_budget(limiter):
    """Per step: vol_change = sum (C^(n+1) - C^n) dx dy of the downloaded states (to rounding of the
    sums), and the flux form conserves volume up to the clamp's part: |vol_change - vol_clamp| <=
    1e-13 vol_before (walls carry no flux).  The first step advects with the initial velocity, a
    compression toward the centre (div u < 0, not yet projected), so C = 1 cells overshoot and the
    clamp fires."""
    g = O.Grid(48, 48)
    ph = phys(rho1=1000.0, rho2=1.0, mu1=1.0, mu2=0.01, sigma=1.0, time_order=2)
    st = O.State(g)
    st.c[...] = droplet_c(g, 0.25)
    Xu, Yu = g.coords("u")
    st.u[...] = -2.0 * (Xu - 0.5)
    O.velocity_bcs(g, st.u, st.v)
    st.init_properties(ph)
    ctx = abi.Context(48, 48, params=abi.FuguParams(1000.0, 1.0, 1.0, 0.01, 1.0, 0.0, 0.0, 2, 1e12, 0.5))
    ctx.set_vof_limiter(limiter)
    ctx.upload(st.u, st.v, st.u_prev, st.v_prev, st.p, st.c)
    ctx.init_properties()
    s = (slice(g.nh, g.nh + g.ny), slice(g.nh, g.nh + g.nx))
    vol = g.dx * g.dy
    c_prev = st.c.copy()
    clamp_seen = 0.0
    opts = abi.step_opts(tol=1e-8, max_iter=200, solver_flags=abi.SOLVE_REFINED | abi.PRECOND_MG, inner_max_iter=-1)
    for step in range(1, 6):
        stt, rc = ctx.step(2e-3, step, opts)
        c_now = ctx.field("c", np.zeros(g.shape("p")))
        host = (c_now[s].sum() - c_prev[s].sum()) * vol
        assert abs(stt.vol_before - c_prev[s].sum() * vol) <= 1e-14 * stt.vol_before
        assert abs(stt.vol_change - host) <= 1e-13 * stt.vol_before, (step, stt.vol_change, host)
        assert abs(stt.vol_change - stt.vol_clamp) <= 1e-13 * stt.vol_before, (step, stt.vol_change, stt.vol_clamp)
        clamp_seen = max(clamp_seen, abs(stt.vol_clamp))
        c_prev = c_now
    assert clamp_seen > 0.0
    ctx.close()
