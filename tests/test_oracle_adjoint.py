"""Pins of the oracle's adjoint building blocks (oracle/fugu2d_adjoint.c):

- transposed halo and velocity BCs: the dot-product identity on the full padded arrays;
- A_eff^T: the dot-product identity <A x, y> = <x, A^T y> on the owned degrees of freedom and,
  at small size, the assembled A_eff^T against the transpose of the assembled A_eff;
- the linearised PCD transpose: dot-product identity with a tight pressure solve;
- coefficient gradients of lam^T A(coef) X: exact, because A is linear in every coefficient
  field and in the advecting velocity;
- the per-cell property path: equals the mixture law for mixture-law inputs, its transpose
  passes the dot-product identity (rho, linear) and central differences (mu, harmonic means).
"""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, two_phase_case

GRIDS = [
    dict(nx=12, ny=10, lx=1.2, ly=1.0, per_x=False, per_y=False),
    dict(nx=10, ny=12, lx=1.0, ly=1.5, per_x=True, per_y=False),
    dict(nx=12, ny=8, lx=1.0, ly=1.0, per_x=False, per_y=True),
    dict(nx=10, ny=10, lx=1.0, ly=1.0, per_x=True, per_y=True),
]
IDS = ["walls", "periodic_x", "periodic_y", "periodic_xy"]


def grid(kw):
    return O.Grid(kw["nx"], kw["ny"], kw["lx"], kw["ly"], per_x=kw["per_x"], per_y=kw["per_y"])


def rtrip(g, rng):
    return [rand(g, "u", rng), rand(g, "v", rng), rand(g, "p", rng)]


def odot(g, a, b):
    """Plain owned-set inner product (coupled_dot without the constant vol)."""
    return O.coupled_dot(g, a, b) / (g.dx * g.dy)


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_halo_and_bcs_transpose(kw):
    g = grid(kw)
    rng = np.random.default_rng(3)
    for kind in ("u", "v", "p"):
        x, y = rand(g, kind, rng), rand(g, kind, rng)
        hx = x.copy(); O.halo(g, hx, kind)
        hy = y.copy(); O.halo_T(g, hy, kind)
        assert abs(np.sum(hx * y) - np.sum(x * hy)) < 1e-12 * np.sum(np.abs(hx * y))
    xu, xv, yu, yv = rand(g, "u", rng), rand(g, "v", rng), rand(g, "u", rng), rand(g, "v", rng)
    bu, bv = xu.copy(), xv.copy(); O.velocity_bcs(g, bu, bv, 0.0)
    tu, tv = yu.copy(), yv.copy(); O.velocity_bcs_T(g, tu, tv)
    lhs = np.sum(bu * yu) + np.sum(bv * yv)
    rhs = np.sum(xu * tu) + np.sum(xv * tv)
    assert abs(lhs - rhs) < 1e-12 * (np.sum(np.abs(bu * yu)) + np.sum(np.abs(bv * yv)))


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
@pytest.mark.parametrize("adv", [True, False])
def test_matvec_transpose_dot_product(kw, adv):
    g = grid(kw)
    rng = np.random.default_rng(5)
    c, pr, ua, va = two_phase_case(g, rng)
    x, y = O.mask_owned(g, rtrip(g, rng)), O.mask_owned(g, rtrip(g, rng))
    ax = O.matvec_eff(g, x, pr, ua, va, 1e-2, adv)
    aty = O.matvec_T(g, y, pr, ua, va, 1e-2, adv)
    lhs, rhs = odot(g, ax, y), odot(g, x, aty)
    scale = np.sqrt(odot(g, ax, ax) * odot(g, y, y))
    assert abs(lhs - rhs) < 1e-13 * scale, (lhs, rhs)
    # outputs live on the owned set only
    for a, b in zip(aty, O.mask_owned(g, [t.copy() for t in aty])):
        assert np.array_equal(a, b)


def _owned_index(g):
    """(kind, flat index) of every owned degree of freedom, in u, v, p order."""
    probe = O.mask_owned(g, [np.ones(g.shape(k)) for k in "uvp"])
    return [(k, i) for k, a in enumerate(probe) for i in np.flatnonzero(a.ravel())]


def _assemble(g, op):
    idx = _owned_index(g)
    n = len(idx)
    A = np.zeros((n, n))
    for col, (k, i) in enumerate(idx):
        X = [g.zeros(t) for t in "uvp"]
        X[k].ravel()[i] = 1.0
        Y = op(X)
        A[:, col] = [Y[kk].ravel()[ii] for kk, ii in idx]
    return A


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_matvec_transpose_dense(kw):
    g = grid(dict(kw, nx=6, ny=6))
    rng = np.random.default_rng(7)
    c, pr, ua, va = two_phase_case(g, rng)
    A = _assemble(g, lambda X: O.matvec_eff(g, X, pr, ua, va, 1e-2, True))
    AT = _assemble(g, lambda X: O.matvec_T(g, X, pr, ua, va, 1e-2, True))
    assert np.abs(AT - A.T).max() < 1e-12 * np.abs(A).max()
    # the effective operator is not symmetric (advection, the divergence/gradient pair)
    assert np.abs(A - A.T).max() > 1e-3 * np.abs(A).max()


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
@pytest.mark.parametrize("visc", [0.0, 1.0])
def test_pcd_lin_transpose_dot_product(kw, visc):
    g = grid(kw)
    rng = np.random.default_rng(11)
    c, pr, ua, va = two_phase_case(g, rng)
    x, y = O.mask_owned(g, rtrip(g, rng)), O.mask_owned(g, rtrip(g, rng))
    px = O.pcd_lin(g, x, pr, ua, va, 1e-2, visc=visc)
    pty = O.pcd_lin_T(g, y, pr, ua, va, 1e-2, visc=visc)
    lhs, rhs = odot(g, px, y), odot(g, x, pty)
    scale = np.sqrt(odot(g, px, px) * odot(g, y, y))
    assert abs(lhs - rhs) < 1e-9 * scale, (lhs, rhs)


def test_pcd_lin_is_linear_and_matches_reference_blocks():
    """With the pressure solve converged, the linearised PCD is linear; its velocity part with
    z_p = 0 equals the reference PCD's (the pressure solve of a zero rhs is zero in both)."""
    g = O.Grid(12, 10, 1.2, 1.0)
    rng = np.random.default_rng(13)
    c, pr, ua, va = two_phase_case(g, rng)
    a, b = O.mask_owned(g, rtrip(g, rng)), O.mask_owned(g, rtrip(g, rng))
    pa, pb = O.pcd_lin(g, a, pr, ua, va, 1e-2), O.pcd_lin(g, b, pr, ua, va, 1e-2)
    pab = O.pcd_lin(g, [2.0 * x - 3.0 * y for x, y in zip(a, b)], pr, ua, va, 1e-2)
    for k in range(3):
        assert np.abs(pab[k] - (2.0 * pa[k] - 3.0 * pb[k])).max() < 1e-9 * np.abs(pab[k]).max()
    z = O.mask_owned(g, [a[0], a[1], g.zeros("p")])
    lin = O.pcd_lin(g, z, pr, ua, va, 1e-2)
    ref = O.mask_owned(g, list(O.pcd(g, z, pr, ua, va, 1e-2)))
    for k in range(3):
        assert np.array_equal(lin[k], ref[k])


@pytest.mark.parametrize("kw", GRIDS, ids=IDS)
def test_operator_coefficient_gradients_exact(kw):
    g = grid(kw)
    rng = np.random.default_rng(17)
    c, pr, ua, va = two_phase_case(g, rng)
    dt = 1e-2
    X = O.mask_owned(g, rtrip(g, rng))
    L = O.mask_owned(g, rtrip(g, rng))
    Xe = O.extend(g, X)
    grads = O.operator_coef_grad(g, L, Xe, pr, ua, va, dt, True, 1.0)

    def form(props, u_adv, v_adv):
        Y = [g.zeros(k) for k in "uvp"]
        Xc = [a.copy() for a in Xe]
        O.lib().og_matvec_coupled(O.C.byref(g.c), *[O.ptr(a) for a in Xc], *[O.ptr(a) for a in Y], O.ptr(u_adv),
                                  O.ptr(v_adv), *[O.ptr(props[k]) for k in O.PROP_NAMES], dt, 1,
                                  None, None, None, None)
        O.mask_owned(g, Y)
        return odot(g, L, Y)

    base = form(pr, ua, va)
    kinds = dict(rho_c="p", rho_u="u", rho_v="v", kappa="p", zeta="w", eta="w")
    for name, kind in kinds.items():
        d = rand(g, kind, rng)
        p2 = dict(pr); p2[name] = pr[name] + d
        lhs = form(p2, ua, va) - base
        rhs = np.sum(grads[name] * d)
        assert abs(lhs - rhs) < 1e-9 * max(abs(lhs), 1.0), (name, lhs, rhs)
    du, dv = rand(g, "u", rng), rand(g, "v", rng)
    lhs = form(pr, ua + du, va) - base
    assert abs(lhs - np.sum(grads["u_adv"] * du)) < 1e-9 * max(abs(lhs), 1.0)
    lhs = form(pr, ua, va + dv) - base
    assert abs(lhs - np.sum(grads["v_adv"] * dv)) < 1e-9 * max(abs(lhs), 1.0)


def test_cell_properties_match_mixture_law():
    g = O.Grid(16, 12, 1.0, 0.75)
    rng = np.random.default_rng(19)
    c = np.clip(rng.uniform(-0.2, 1.2, g.shape("p")), 0.0, 1.0)
    rho1, rho2, mu1, mu2 = 1000.0, 1.0, 2.0, 0.01
    mix = O.mixture(g, c, rho1, rho2, mu1, mu2)
    mu_cell = np.vectorize(lambda x: O.lib().og_harmonic_viscosity(x, mu1, mu2))(c)
    rho_cell = rho1 * c + rho2 * (1.0 - c)
    cp = O.cell_properties(g, mu_cell, rho_cell)
    for k in O.PROP_NAMES:
        assert np.abs(cp[k] - mix[k]).max() <= 1e-13 * np.abs(mix[k]).max(), k


def test_cell_properties_transpose():
    g = O.Grid(10, 8, per_x=True)
    rng = np.random.default_rng(23)
    mu = rng.uniform(0.01, 2.0, g.shape("p"))
    rho = rng.uniform(1.0, 1000.0, g.shape("p"))
    gc = dict(rho_c=rand(g, "p", rng), rho_u=rand(g, "u", rng), rho_v=rand(g, "v", rng),
              kappa=rand(g, "p", rng), zeta=rand(g, "w", rng), eta=rand(g, "w", rng))

    def F(m, r):
        cp = O.cell_properties(g, m, r)
        return sum(np.sum(cp[k] * gc[k]) for k in O.PROP_NAMES)

    g_mu, g_rho = O.cell_properties_T(g, mu, gc)
    d = rand(g, "p", rng)
    lhs = F(mu, rho + d) - F(mu, rho)
    assert abs(lhs - np.sum(g_rho * d)) < 1e-9 * abs(lhs)
    eps = 1e-3      # central differences: truncation ~eps^2 and rounding ~1e-16/eps balance here
    dm = d * mu * 0.1
    fd = (F(mu + eps * dm, rho) - F(mu - eps * dm, rho)) / (2 * eps)
    assert abs(fd - np.sum(g_mu * dm)) < 1e-7 * abs(fd)
