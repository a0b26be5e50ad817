"""Pins of the oracle's Krylov restatements (src/mod_bicgstab2.f90:77-264 BiCGStab,
:270-570 BiCGStab(2)) by closed forms and dense solves.

The PCD's inner PCG normally stops at 1e-5 relative (src/mod_pcd_precond.f90:50), which makes the
preconditioner P^-1 depend slightly on its input; here it runs to 1e-13 instead (og_set_pcd_pcg),
so P^-1 is a fixed linear map to rounding and the left-preconditioned operator M = P^-1 A on the
owned unknowns can be assembled column by column (the same BCs, halos and ghost conventions the
solvers apply).  Then:

* dense solves: both solvers, the reference's BiCGStab(2) included (on periodic grids, where the
  velocity BCs it omits on r and s, :430 and :489, do nothing), reproduce the dense solution of
  A x = b to 1e-8 at tol 1e-12 -- on the problem scaled by 2^40.  The reference's breakdown guards
  are absolute (:510 `lambda * tau > 1e-30`, :518, :522, :181 ...): at the unit scale of these
  cases lambda * tau = |s|^2 |t|^2 falls below 1e-30 once the residual is near 1e-7, the GCR(2)
  step then drops its second direction (omega2 = 0) and the iteration stagnates.  A power-of-two
  scale changes no rounding, so the scaled run repeats the unscaled one exactly (scaled) until a
  guard decides differently, and then converges; test_reference_guards_are_absolute pins both;
* the BiCGStab(2) iterates of the first five iterations against an independent textbook
  BiCGStab(l = 2) (Sleijpen and Fokkema, ETNA 1 (1993) 11-32; minimal-residual part by least
  squares), up to the pressure constant -- the carries between iterations (xi0 <- -omega2 xi0,
  the u and x updates);
* closed forms of the first iteration (max_iter = 1), independent of the recurrences:
    BiCGStab:    r1 = (I - w M)(I - a M) r0,  a = <r0, r0>/<r0, M r0>,  w = <t, s>/<t, t>;
    BiCGStab(2): r^ = T(M) r0 with T(0) = 1, deg T = 2 and <T(M) r0, r0> = <M T(M) r0, r0> = 0
                 (the two BiCG steps, shadow residual r0), then the GCR(2) minimal-residual step
                 r1 = r^ - w1 M r^ - w2 M^2 r^ with (w1, w2) minimising ||r1|| (:504-530);
  x1 = x0 + q(M) r0 where r0 - r1 = M q(M) r0.  The iterate and the reported residual
  ||r1|| / ||P^-1 b|| must agree; the same closed forms with a coefficient slipped (the GCR(2)
  weights swapped, the BiCG step's polynomial dropped, the BiCGStab weight's sign flipped) are far
  from the oracle's iterate, so the comparison discriminates.
All inner products are coupled_dot (vol-weighted sums over the owned unknowns)."""
import numpy as np
import pytest

import pyoracle as O
from cases import rand, two_phase_case


@pytest.fixture
def linear_pcd():
    O.set_pcd_pcg(1e-13, 1000)
    yield
    O.set_pcd_pcg(1e-5, 150)


def dofs(g):
    iu = g.nx if g.per_x else g.nx - 1
    jv = g.ny if g.per_y else g.ny - 1
    return ([("u", i, j) for j in range(1, g.ny + 1) for i in range(1, iu + 1)] +
            [("v", i, j) for j in range(1, jv + 1) for i in range(1, g.nx + 1)] +
            [("p", i, j) for j in range(1, g.ny + 1) for i in range(1, g.nx + 1)])


IX = {"u": "iu", "v": "iv", "p": "ip"}


def to_vec(g, X, D):
    return np.array([X["uvp".index(k)][getattr(g, IX[k])(i, j)] for (k, i, j) in D])


def from_vec(g, x, D):
    X = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    for val, (k, i, j) in zip(x, D):
        X["uvp".index(k)][getattr(g, IX[k])(i, j)] = val
    return X


def operators(g, pr, ua, va, dt):
    """A (velocity BCs, then matvec_coupled) and M = P^-1 A on the owned unknowns."""
    D = dofs(g)
    n = len(D)
    A, M = np.zeros((n, n)), np.zeros((n, n))
    for col in range(n):
        e = np.zeros(n)
        e[col] = 1.0
        X = from_vec(g, e, D)
        O.velocity_bcs(g, X[0], X[1])
        Y = O.matvec(g, X, pr, ua, va, dt)
        A[:, col] = to_vec(g, Y, D)
        M[:, col] = to_vec(g, O.pcd(g, Y, pr, ua, va, dt), D)
    return D, A, M


def case(shape, per, seed=14):
    nx, ny = shape
    g = O.Grid(nx, ny, 1.0, ny / nx, per_x=per, per_y=per)
    rng = np.random.default_rng(seed)
    c, pr, ua, va = two_phase_case(g, rng, rho=(10.0, 1.0), mu=(1.0, 0.1))
    if per:
        for k, kind in (("rho_u", "u"), ("rho_v", "v"), ("rho_c", "p")):
            O.halo(g, pr[k], kind)
        O.halo(g, ua, "u"); O.halo(g, va, "v")
    D = dofs(g)
    b = np.zeros(len(D))
    nvel = sum(1 for d in D if d[0] != "p")
    b[:nvel] = rng.standard_normal(nvel)          # the pressure rows of B are zero
    x0 = 0.01 * rng.standard_normal(len(D))
    return g, pr, ua, va, D, b, x0, nvel


SOLVERS = [(1, False), (2, False), (2, True)]
SCALE = 2.0 ** 40   # keeps the reference's absolute breakdown guards from firing


@pytest.mark.parametrize("per", [False, True], ids=["walls", "periodic_xy"])
@pytest.mark.parametrize("order,consistent", SOLVERS, ids=["bicgstab", "bicgstab2_ref", "bicgstab2_bc"])
def test_krylov_matches_dense_solve(linear_pcd, per, order, consistent):
    if not per and order == 2 and not consistent:
        pytest.skip("the reference BiCGStab(2) stalls on walled domains (test_oracle_solvers)")
    g, pr, ua, va, D, b, x0, nvel = case((8, 8), per)
    dt = 0.01
    b, x0 = b * SCALE, x0 * SCALE
    B = from_vec(g, b, D)
    X = from_vec(g, x0, D)
    it, res = O.krylov(g, X, B, pr, ua, va, dt, tol=1e-12, max_iter=300, order=order, consistent=consistent)
    assert res < 1e-12, (it, res)
    _, A, _ = operators(g, pr, ua, va, dt)
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    got = to_vec(g, X, D)
    assert np.abs(got[:nvel] - x[:nvel]).max() <= 1e-8 * np.abs(x[:nvel]).max()
    pg, px = got[nvel:] - got[nvel:].mean(), x[nvel:] - x[nvel:].mean()
    assert np.abs(pg - px).max() <= 1e-8 * np.abs(px).max()


def first_iterate_bicgstab(M, r0, x0, slip=False):
    a = (r0 @ r0) / (r0 @ (M @ r0))
    s = r0 - a * (M @ r0)
    t = M @ s
    w = (t @ s) / (t @ t)
    if slip:
        w = -w
    return x0 + a * r0 + w * s, s - w * t


def first_iterate_bicgstab2(M, r0, x0, slip=None):
    K = [r0]
    for _ in range(4):
        K.append(M @ K[-1])
    # T(z) = 1 + c1 z + c2 z^2 with <T(M) r0, r0> = <M T(M) r0, r0> = 0
    G = np.array([[K[1] @ r0, K[2] @ r0], [K[2] @ r0, K[3] @ r0]])
    c = np.linalg.solve(G, -np.array([K[0] @ r0, K[1] @ r0]))
    if slip == "no_bicg":
        c = np.zeros(2)
    rh = K[0] + c[0] * K[1] + c[1] * K[2]
    s, t = M @ rh, M @ (M @ rh)
    w = np.linalg.solve(np.array([[s @ s, s @ t], [t @ s, t @ t]]), np.array([rh @ s, rh @ t]))
    if slip == "swap":
        w = w[::-1]
    # r1 = (1 - w1 z - w2 z^2)(1 + c1 z + c2 z^2) r0 = sum_k d_k z^k r0
    d = np.zeros(5)
    prod = np.polynomial.polynomial.polymul([1.0, -w[0], -w[1]], [1.0, c[0], c[1]])
    d[:len(prod)] = prod
    r1 = sum(d[k] * K[k] for k in range(5))
    x1 = x0 - sum(d[k] * K[k - 1] for k in range(1, 5))
    return x1, r1


@pytest.mark.parametrize("shape,per,order,consistent",
                         [((4, 4), True, 1, False), ((6, 5), False, 1, False),
                          ((4, 4), True, 2, False), ((6, 5), False, 2, True)],
                         ids=["bicgstab_periodic", "bicgstab_walls", "bicgstab2_ref_periodic", "bicgstab2_bc_walls"])
def test_first_iterate_closed_form(linear_pcd, shape, per, order, consistent):
    g, pr, ua, va, D, b, x0, nvel = case(shape, per, seed=21)
    dt = 0.01
    _, A, M = operators(g, pr, ua, va, dt)
    B = from_vec(g, b, D)
    Pb = to_vec(g, O.pcd(g, B, pr, ua, va, dt), D)
    r0 = Pb - M @ x0
    X = from_vec(g, x0, D)
    it, res = O.krylov(g, X, B, pr, ua, va, dt, tol=1e-300, max_iter=1, order=order, consistent=consistent)
    assert it == 1
    got = to_vec(g, X, D)
    if order == 1:
        x1, r1 = first_iterate_bicgstab(M, r0, x0)
        xs, _ = first_iterate_bicgstab(M, r0, x0, slip=True)
        slips = [xs]
    else:
        x1, r1 = first_iterate_bicgstab2(M, r0, x0)
        slips = [first_iterate_bicgstab2(M, r0, x0, slip=s)[0] for s in ("swap", "no_bicg")]
    scale = np.abs(x1).max()
    err = np.abs(got - x1).max() / scale
    assert err <= 1e-9, err
    assert abs(res - np.linalg.norm(r1) / np.linalg.norm(Pb)) <= 1e-9 * res
    for xs in slips:   # the check tells a slipped coefficient apart from the restatement
        assert np.abs(got - xs).max() / scale > 1e3 * max(err, 1e-12)


def bicgstab_l(M, r0, x0, ell, iters):
    """Textbook BiCGStab(l) (Sleijpen and Fokkema 1993) for M x = b with x0 and r0 = b - M x0,
    shadow residual r0; the minimal-residual part by least squares.  Iterates after each sweep."""
    x, rt = x0.copy(), r0.copy()
    r = [r0.copy()] + [None] * ell
    uh = [np.zeros_like(r0)] + [None] * ell
    rho0, alpha, omega = 1.0, 0.0, 1.0
    out = []
    for _ in range(iters):
        rho0 = -omega * rho0
        for j in range(ell):
            rho1 = r[j] @ rt
            beta = alpha * rho1 / rho0
            rho0 = rho1
            for i in range(j + 1):
                uh[i] = r[i] - beta * uh[i]
            uh[j + 1] = M @ uh[j]
            alpha = rho0 / (uh[j + 1] @ rt)
            for i in range(j + 1):
                r[i] = r[i] - alpha * uh[i + 1]
            r[j + 1] = M @ r[j]
            x = x + alpha * uh[0]
        gam = np.linalg.lstsq(np.array(r[1:]).T, r[0], rcond=None)[0]
        x = x + sum(gam[j] * r[j] for j in range(ell))
        r = [r[0] - sum(gam[j] * r[j + 1] for j in range(ell))] + [None] * ell
        uh = [uh[0] - sum(gam[j] * uh[j + 1] for j in range(ell))] + [None] * ell
        omega = gam[-1]
        out.append(x.copy())
    return out


def mod_pressure_constant(x, nvel):
    y = x.copy()
    y[nvel:] -= y[nvel:].mean()
    return y


@pytest.mark.parametrize("per", [False, True], ids=["walls", "periodic_xy"])
def test_bicgstab2_iterates_match_textbook(linear_pcd, per):
    g, pr, ua, va, D, b, x0, nvel = case((8, 8), per)
    dt = 0.01
    _, A, M = operators(g, pr, ua, va, dt)
    B = from_vec(g, b, D)
    Pb = to_vec(g, O.pcd(g, B, pr, ua, va, dt), D)
    ref = bicgstab_l(M, Pb - M @ x0, x0, 2, 5)
    for k in range(1, 6):
        X = from_vec(g, x0, D)
        O.krylov(g, X, B, pr, ua, va, dt, tol=1e-300, max_iter=k, order=2, consistent=not per)
        got = mod_pressure_constant(to_vec(g, X, D), nvel)
        want = mod_pressure_constant(ref[k - 1], nvel)
        assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max(), k


def test_reference_guards_are_absolute(linear_pcd):
    """The unscaled BiCGStab(2) run and the run scaled by 2^40 agree exactly (scaled) for the first
    two iterations (in the third the PCD's inner PCG, run to 1e-13 here, meets its own absolute
    `pAp < 1e-30` guard at the unit scale); later the unscaled run's GCR(2) guard drops omega2 and
    it stagnates above 1e-9 while the scaled one converges below 1e-12 (the same algorithm with
    guards relative to ||P^-1 b||^2 would converge at both scales)."""
    g, pr, ua, va, D, b, x0, nvel = case((8, 8), True)
    dt = 0.01
    out = {}
    for sc in (1.0, SCALE):
        for k in (2, 60):
            X = from_vec(g, x0 * sc, D)
            it, res = O.krylov(g, X, from_vec(g, b * sc, D), pr, ua, va, dt, tol=1e-14, max_iter=k, order=2)
            out[sc, k] = (to_vec(g, X, D), res)
    assert np.array_equal(out[SCALE, 2][0], SCALE * out[1.0, 2][0]) and out[SCALE, 2][1] == out[1.0, 2][1]
    assert out[1.0, 60][1] > 1e-9 and out[SCALE, 60][1] < 1e-12, (out[1.0, 60][1], out[SCALE, 60][1])
