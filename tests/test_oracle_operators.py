"""Pins of the oracle's operator restatements (src/mod_operators.f90, src/mod_mpi.f90,
src/mod_viscosity.f90, src/mod_bicgstab2.f90:32-72) against closed forms, the
reference's own known-answer tests and structural identities -- never against a
retyped formula."""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, phys

TWO_PI = 2.0 * np.pi


def taylor_green(g):
    """Fields of tests/test_operators.f90:94-129 over the full padded arrays."""
    Xu, Yu = g.coords("u")
    Xv, Yv = g.coords("v")
    Xp, Yp = g.coords("p")
    u = np.sin(TWO_PI * Xu) * np.cos(TWO_PI * Yu)
    v = -np.cos(TWO_PI * Xv) * np.sin(TWO_PI * Yv)
    p = 0.25 * (np.cos(4 * np.pi * Xp) + np.cos(4 * np.pi * Yp))
    return u, v, p


def test_halo_exchange_periodic_wrap():
    """tests/test_halo_exchange.f90: periodic ghosts reproduce the analytic field (< 1e-13)."""
    g = O.Grid(64, 64, per_x=True, per_y=True)
    for kind in ("p", "u", "v"):
        X, Y = g.coords(kind)
        exact = np.sin(TWO_PI * X) * np.cos(TWO_PI * Y)
        a = np.full(g.shape(kind), 1.0e30)
        a[interior(g, kind)] = exact[interior(g, kind)]
        O.halo(g, a, kind)
        # filled ranges: x ghosts 1-nh..0 and n+1..n+nh (full j), y ghosts likewise (full i);
        # u(-nh) and v(., -nh) are never written by the exchange (src/mod_mpi.f90:175,235).
        mask = np.ones(g.shape(kind), bool)
        if kind == "u":
            mask[:, 0] = False
        if kind == "v":
            mask[0, :] = False
        assert np.abs(a - exact)[mask].max() < 1e-13


def test_halo_walls_untouched():
    g = O.Grid(8, 6)
    a = np.random.default_rng(0).standard_normal(g.shape("p"))
    b = a.copy()
    O.halo(g, b, "p")
    assert np.array_equal(a, b)


def test_taylor_green_operators():
    """tests/test_operators.f90: A) |div| < 1e-13, B) grad err < 5e-2, C) visc err < 1."""
    g = O.Grid(64, 64, per_x=True, per_y=True)
    u, v, p = taylor_green(g)
    for a, k in ((u, "u"), (v, "v"), (p, "p")):
        O.halo(g, a, k)
    c = np.ones(g.shape("p"))
    pr = O.mixture(g, c, 1.0, 1.0, 1.0, 1.0)
    div = g.zeros("p")
    O.lib().og_divergence(g.c, O.ptr(u), O.ptr(v), O.ptr(div))
    assert np.abs(div[interior(g, "p")]).max() < 1e-13
    gu, gv = g.zeros("u"), g.zeros("v")
    O.lib().og_grad_pressure(g.c, O.ptr(p), O.ptr(gu), O.ptr(gv))
    Xu, Yu = g.coords("u")
    sl = (slice(g.nh, g.nh + g.ny), slice(g.nh + 1, g.nh + g.nx))      # i = 1..nx-1
    assert np.abs(gu - (-np.pi * np.sin(4 * np.pi * Xu)))[sl].max() < 5e-2
    vu, vv = g.zeros("u"), g.zeros("v")
    O.lib().og_viscous_stress(g.c, O.ptr(u), O.ptr(v), O.ptr(pr["kappa"]), O.ptr(pr["zeta"]),
                              O.ptr(pr["eta"]), O.ptr(vu), O.ptr(vv))
    exact = 8 * np.pi ** 2 * np.sin(TWO_PI * Xu) * np.cos(TWO_PI * Yu)
    err = np.abs(vu - exact)[sl].max()
    assert err < 1.0
    # second order: halving h divides the error by ~4
    g2 = O.Grid(128, 128, per_x=True, per_y=True)
    u2, v2, _ = taylor_green(g2)
    O.halo(g2, u2, "u"); O.halo(g2, v2, "v")
    pr2 = O.mixture(g2, np.ones(g2.shape("p")), 1.0, 1.0, 1.0, 1.0)
    vu2, vv2 = g2.zeros("u"), g2.zeros("v")
    O.lib().og_viscous_stress(g2.c, O.ptr(u2), O.ptr(v2), O.ptr(pr2["kappa"]), O.ptr(pr2["zeta"]),
                              O.ptr(pr2["eta"]), O.ptr(vu2), O.ptr(vv2))
    Xu2, Yu2 = g2.coords("u")
    sl2 = (slice(g2.nh, g2.nh + g2.ny), slice(g2.nh + 1, g2.nh + g2.nx))
    err2 = np.abs(vu2 - 8 * np.pi ** 2 * np.sin(TWO_PI * Xu2) * np.cos(TWO_PI * Yu2))[sl2].max()
    assert 3.5 < err / err2 < 4.5


def test_viscous_exact_for_quadratics():
    """Central differences reproduce -div(T) exactly for quadratic fields with constant mu:
    visc_u = -mu (2 u_xx + u_yy + v_xy), visc_v = -mu (u_xy + v_xx + 2 v_yy)."""
    g = O.Grid(12, 10, lx=1.0, ly=0.8)
    mu = 0.7
    Xu, Yu = g.coords("u")
    Xv, Yv = g.coords("v")
    u = Xu ** 2 + Xu * Yu + Yu ** 2            # u_xx=2 u_yy=2 u_xy=1
    v = Xv ** 2 - 3 * Xv * Yv + 2 * Yv ** 2    # v_xx=2 v_yy=4 v_xy=-3
    pr = O.mixture(g, np.ones(g.shape("p")), 1.0, 1.0, mu, mu)
    vu, vv = g.zeros("u"), g.zeros("v")
    O.lib().og_viscous_stress(g.c, O.ptr(u), O.ptr(v), O.ptr(pr["kappa"]), O.ptr(pr["zeta"]),
                              O.ptr(pr["eta"]), O.ptr(vu), O.ptr(vv))
    assert np.allclose(vu[interior(g, "u")], -mu * (4 + 2 - 3), rtol=0, atol=1e-9)
    assert np.allclose(vv[interior(g, "v")], -mu * (1 + 2 + 8), rtol=0, atol=1e-9)


def test_grad_exact_and_grad_div_adjoint():
    """grad of a quadratic p is exact at face midpoints; and <div X, q> = -<X, grad q>
    for X with zero wall-normal velocity (pins grad/div index conventions together)."""
    g = O.Grid(9, 7, lx=0.9, ly=0.7)
    Xp, Yp = g.coords("p")
    p = 3 * Xp ** 2 - Xp * Yp + 0.5 * Yp ** 2
    gu, gv = g.zeros("u"), g.zeros("v")
    O.lib().og_grad_pressure(g.c, O.ptr(p), O.ptr(gu), O.ptr(gv))
    Xu, Yu = g.coords("u")
    Xv, Yv = g.coords("v")
    assert np.allclose(gu[interior(g, "u")], (6 * Xu - Yu)[interior(g, "u")], atol=1e-10)
    assert np.allclose(gv[interior(g, "v")], (-Xv + Yv)[interior(g, "v")], atol=1e-10)
    rng = np.random.default_rng(1)
    u, v, q = rand(g, "u", rng), rand(g, "v", rng), rand(g, "p", rng)
    O.velocity_bcs(g, u, v)
    div = g.zeros("p")
    O.lib().og_divergence(g.c, O.ptr(u), O.ptr(v), O.ptr(div))
    O.lib().og_grad_pressure(g.c, O.ptr(q), O.ptr(gu), O.ptr(gv))
    lhs = (div * q)[interior(g, "p")].sum()
    rhs = -((u * gu)[interior(g, "u")].sum() + (v * gv)[interior(g, "v")].sum())
    assert abs(lhs - rhs) < 1e-10 * (abs(lhs) + 1)


def test_advection_exact_linear():
    """rho = 1, u_adv = (a, b) constant: central flux differences give a u_x + b u_y exactly
    for linear u (the discrete product rule is exact when the carrier is constant)."""
    g = O.Grid(10, 8, per_x=False, per_y=False)
    a_, b_ = 0.7, -0.3
    Xu, Yu = g.coords("u")
    Xv, Yv = g.coords("v")
    u = 2 * Xu - 5 * Yu
    v = -Xv + 4 * Yv
    pr = O.mixture(g, np.ones(g.shape("p")), 1.0, 1.0, 1.0, 1.0)
    ua, va = np.full(g.shape("u"), a_), np.full(g.shape("v"), b_)
    au, av = g.zeros("u"), g.zeros("v")
    O.lib().og_advection(g.c, O.ptr(ua), O.ptr(va), O.ptr(pr["rho_c"]), O.ptr(pr["rho_u"]),
                         O.ptr(pr["rho_v"]), O.ptr(u), O.ptr(v), O.ptr(au), O.ptr(av), None, None)
    s_u = (slice(g.nh, g.nh + g.ny), slice(g.nh + 1, g.nh + g.nx))    # i=1..nx-1
    s_v = (slice(g.nh + 1, g.nh + g.ny), slice(g.nh, g.nh + g.nx))    # j=1..ny-1
    assert np.allclose(au[s_u], a_ * 2 + b_ * (-5), atol=1e-10)
    assert np.allclose(av[s_v], a_ * (-1) + b_ * 4, atol=1e-10)
    # rows outside i_u_max / j_v_max stay zero (src/mod_operators.f90:192-199)
    assert np.all(au[:, g.nh + g.nx] == 0) and np.all(av[g.nh + g.ny, :] == 0)


def test_velocity_bcs_semantics():
    """src/mod_operators.f90:254-333: wall faces 0, antisymmetric ghosts, lid 2U - u,
    x pass before y pass (corners take the y-pass value)."""
    g = O.Grid(6, 5)
    rng = np.random.default_rng(2)
    u, v = rand(g, "u", rng), rand(g, "v", rng)
    u0, v0 = u.copy(), v.copy()
    O.velocity_bcs(g, u, v, 1.0)
    nx, ny = g.nx, g.ny
    for j in range(1, ny + 1):
        assert u[g.iu(0, j)] == 0 and u[g.iu(nx, j)] == 0
        assert u[g.iu(-1, j)] == -u0[g.iu(1, j)] and u[g.iu(nx + 2, j)] == -u0[g.iu(nx - 2, j)]
        if j < ny:   # v(., ny) rows are rewritten to 0 by the later y pass
            assert v[g.iv(0, j)] == -v0[g.iv(1, j)] and v[g.iv(nx + 1, j)] == -v0[g.iv(nx, j)]
    for i in range(1, nx):
        assert v[g.iv(i, 0)] == 0 and v[g.iv(i, ny)] == 0
        assert u[g.iu(i, 0)] == -u0[g.iu(i, 1)]
        assert u[g.iu(i, ny + 1)] == 2.0 - u0[g.iu(i, ny)]
        assert u[g.iu(i, ny + 2)] == 2.0 - u0[g.iu(i, ny - 1)]
    # corner: x pass wrote u(-1,0) = -u(1,0) first, the y pass overwrote it with -u(-1,1)
    assert u[g.iu(-1, 0)] == -(-u0[g.iu(1, 1)])
    # interior untouched
    assert np.array_equal(u[g.iu(1, 1)], u0[g.iu(1, 1)])


def test_mixture_closed_forms():
    """src/mod_viscosity.f90: rho_c linear in clamped C; faces are arithmetic means;
    vertex mu of the mean C equals the harmonic mean of the four cell viscosities."""
    g = O.Grid(7, 6)
    rng = np.random.default_rng(3)
    c = rng.uniform(-0.1, 1.1, g.shape("p"))
    r1, r2, m1, m2 = 1000.0, 1.0, 1.0, 0.01
    pr = O.mixture(g, c, r1, r2, m1, m2)
    cc = np.clip(c, 0, 1)
    assert np.allclose(pr["rho_c"], r1 * cc + r2 * (1 - cc), rtol=1e-15)
    mu_c = 1.0 / (cc / m1 + (1 - cc) / m2)
    assert np.allclose(pr["kappa"], 2 * mu_c, rtol=1e-13)
    # vertex (i,j) at corner of cells (i,j),(i+1,j),(i,j+1),(i+1,j+1)
    i, j = 3, 2
    cells = [c[g.ip(i, j)], c[g.ip(i + 1, j)], c[g.ip(i, j + 1)], c[g.ip(i + 1, j + 1)]]
    cm = np.clip(np.mean(cells), 0, 1)
    assert np.isclose(pr["eta"][g.iw(i, j)], 1.0 / (cm / m1 + (1 - cm) / m2), rtol=1e-13)
    assert np.isclose(pr["zeta"][g.iw(i, j)], 2 * pr["eta"][g.iw(i, j)], rtol=1e-15)
    if 0.0 <= min(cells) and max(cells) <= 1.0:
        hm = 4.0 / sum(1.0 / (1.0 / (x / m1 + (1 - x) / m2)) for x in cells)
        assert np.isclose(pr["eta"][g.iw(i, j)], hm, rtol=1e-12)
    assert np.isclose(pr["rho_u"][g.iu(2, 3)], 0.5 * (pr["rho_c"][g.ip(2, 3)] + pr["rho_c"][g.ip(3, 3)]))
    # edge copies at the extreme padded columns / rows
    assert np.array_equal(pr["rho_u"][:, 0], pr["rho_u"][:, 1])
    assert np.array_equal(pr["rho_v"][0, :], pr["rho_v"][1, :])
    assert np.array_equal(pr["zeta"][:, -1], pr["zeta"][:, -2])
    # equal viscosities short-circuit; denominators at the clamp limits
    assert O.lib().og_harmonic_viscosity(0.3, 2.0, 2.0) == 2.0
    assert O.lib().og_harmonic_viscosity(1.5, m1, m2) == m1


def test_coupled_dot_owned_sets():
    """src/mod_bicgstab2.f90:32-72: sum over owned u (i<=nx-1 with walls), v, p; times vol."""
    for per_x, per_y in ((False, False), (True, False), (True, True)):
        g = O.Grid(9, 7, lx=0.9, ly=1.4, per_x=per_x, per_y=per_y)
        rng = np.random.default_rng(4)
        a = [rand(g, k, rng) for k in "uvp"]
        b = [rand(g, k, rng) for k in "uvp"]
        nh, nx, ny = g.nh, g.nx, g.ny
        iu = slice(nh + 1, nh + 1 + (nx if per_x else nx - 1))
        jv = slice(nh + 1, nh + 1 + (ny if per_y else ny - 1))
        exp = ((a[0] * b[0])[nh:nh + ny, iu].sum() + (a[1] * b[1])[jv, nh:nh + nx].sum()
               + (a[2] * b[2])[nh:nh + ny, nh:nh + nx].sum()) * g.dx * g.dy
        assert np.isclose(O.coupled_dot(g, a, b), exp, rtol=1e-13)


def test_viscous_block_symmetric_with_walls():
    """With walls, no advection and X_p = 0, the velocity block of the coupled operator acting on
    owned faces (ghosts by antisymmetric reflection) is symmetric: A == A^T.  A transposed or
    shifted index in the stress stencil breaks this."""
    g = O.Grid(6, 5, lx=1.2, ly=1.0)
    rng = np.random.default_rng(5)
    c = rng.uniform(0, 1, g.shape("p"))
    pr = O.mixture(g, c, 3.0, 1.0, 2.0, 0.05)
    dofs = [("u", i, j) for j in range(1, g.ny + 1) for i in range(1, g.nx)] + \
           [("v", i, j) for j in range(1, g.ny) for i in range(1, g.nx + 1)]
    n = len(dofs)
    A = np.zeros((n, n))
    zu, zv = g.zeros("u"), g.zeros("v")

    def idx(k, i, j):
        return g.iu(i, j) if k == "u" else g.iv(i, j)
    for col, (k, i, j) in enumerate(dofs):
        Xu, Xv, Xp = g.zeros("u"), g.zeros("v"), g.zeros("p")
        (Xu if k == "u" else Xv)[idx(k, i, j)] = 1.0
        O.velocity_bcs(g, Xu, Xv)
        Y = O.matvec(g, (Xu, Xv, Xp), pr, zu, zv, 1.0, include_advection=False)
        for row, (k2, i2, j2) in enumerate(dofs):
            A[row, col] = (Y[0] if k2 == "u" else Y[1])[idx(k2, i2, j2)]
    assert np.abs(A - A.T).max() < 1e-10 * np.abs(A).max()
    # positive definite (mass + viscous dissipation)
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


def test_matvec_structure():
    """Y = [F G; D 0] X with identity rows on wall faces and zero ghosts: the pressure column
    is the gradient, the pressure row is the divergence, the mass term is rho/dt."""
    g = O.Grid(8, 6)
    rng = np.random.default_rng(6)
    c = rng.uniform(0, 1, g.shape("p"))
    pr = O.mixture(g, c, 10.0, 1.0, 1.0, 0.1)
    ua, va = rand(g, "u", rng), rand(g, "v", rng)
    Xu, Xv, Xp = rand(g, "u", rng), rand(g, "v", rng), rand(g, "p", rng)
    Y = O.matvec(g, (Xu.copy(), Xv.copy(), Xp.copy()), pr, ua, va, 0.01)
    # wall rows are identity, ghosts zero
    nh, nx, ny = g.nh, g.nx, g.ny
    assert np.array_equal(Y[0][nh:nh + ny, nh], Xu[nh:nh + ny, nh])
    assert np.array_equal(Y[0][nh:nh + ny, nh + nx], Xu[nh:nh + ny, nh + nx])
    assert np.all(Y[0][:nh] == 0) and np.all(Y[2][:, :nh] == 0)
    # pressure-only input: u-rows equal grad p, p-row zero
    zu, zv = g.zeros("u"), g.zeros("v")
    Yp = O.matvec(g, (zu.copy(), zv.copy(), Xp.copy()), pr, ua, va, 0.01)
    gu, gv = g.zeros("u"), g.zeros("v")
    O.lib().og_grad_pressure(g.c, O.ptr(Xp), O.ptr(gu), O.ptr(gv))
    s = (slice(nh, nh + ny), slice(nh + 1, nh + nx))
    assert np.allclose(Yp[0][s], gu[s], rtol=1e-14, atol=0)
    assert np.all(Yp[2] == 0)
    # linearity in X
    Y2 = O.matvec(g, (2 * Xu, 2 * Xv, 2 * Xp), pr, ua, va, 0.01)
    assert np.allclose(Y2[0], 2 * Y[0], rtol=1e-13, atol=1e-9)
