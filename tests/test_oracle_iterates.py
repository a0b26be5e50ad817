"""Iterates of the oracle's truncated inner solvers pinned by independent restatements.

* SVOF filter (src/mod_vof.f90:145-237): k Jacobi sweeps of (1 - delta^2 Lap) C_s = C are an explicit
  affine map; a vectorised numpy restatement (halo exchange or Neumann mirror, then the sweep) gives
  the oracle's iterate bit for bit at every sweep count, and the same stopping sweep for a tolerance.
* Jacobi-PCG (src/mod_subsolvers.f90:32-246): a textbook preconditioned CG with D = diag(A) on the
  Neumann pressure matrix assembled from the physics (mean projection of the rhs with walls, phi_0 =
  0, the final mean removal) gives the oracle's truncated iterate after k iterations to rounding."""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, two_phase_case
from test_oracle_solvers import neumann_pressure_matrix


def mirror(g, a):
    nh, nx, ny = g.nh, g.nx, g.ny
    if not g.per_x:
        for i in range(1, nh + 1):
            a[:, nh - i] = a[:, nh + i - 1]                  # phi(1-i, j) = phi(i, j)
            a[:, nh + nx + i - 1] = a[:, nh + nx - i]        # phi(nx+i, j) = phi(nx-i+1, j)
    if not g.per_y:
        for j in range(1, nh + 1):
            a[nh - j, :] = a[nh + j - 1, :]
            a[nh + ny + j - 1, :] = a[nh + ny - j, :]


def svof_numpy(g, craw, delta, k, tol=None):
    cx, cy = (delta / g.dx) ** 2, (delta / g.dy) ** 2
    diag = 1.0 + 2.0 * (cx + cy)
    cs = craw.copy()
    nh, nx, ny = g.nh, g.nx, g.ny
    c, e, w = (slice(nh, nh + ny), slice(nh, nh + nx)), slice(nh + 1, nh + nx + 1), slice(nh - 1, nh + nx - 1)
    n, s = slice(nh + 1, nh + ny + 1), slice(nh - 1, nh + ny - 1)
    for it in range(1, k + 1):
        O.halo(g, cs, "p")
        mirror(g, cs)
        old = cs.copy()
        cs[c] = (craw[c] + cx * (old[c[0], e] + old[c[0], w]) + cy * (old[n, c[1]] + old[s, c[1]])) / diag
        if tol is not None and np.abs(cs[c] - old[c]).max() < tol:
            break
    O.halo(g, cs, "p")
    mirror(g, cs)
    return cs, it


@pytest.mark.parametrize("per", [False, True], ids=["walls", "periodic"])
def test_svof_iterates_bitwise(per):
    g = O.Grid(20, 14, 1.0, 0.7, per_x=per, per_y=per)
    rng = np.random.default_rng(31)
    craw = np.clip(rng.uniform(-0.2, 1.2, g.shape("p")), 0, 1)
    delta = 1.5 * min(g.dx, g.dy)
    for k in (1, 2, 7, 40):
        want, _ = svof_numpy(g, craw, delta, k)
        got, it = O.svof(g, craw, delta, max_iter=k, tol=0.0)
        assert it == k and np.array_equal(got, want), k
    want, kw = svof_numpy(g, craw, delta, 500, tol=1e-5)
    got, it = O.svof(g, craw, delta, max_iter=500, tol=1e-5)
    assert it == kw and np.array_equal(got, want)


def jacobi_pcg(A, b, k, walls):
    if walls:
        b = b - b.mean()
    d = np.diag(A).copy()
    dinv = np.where(d > 1e-14, 1.0 / np.where(d > 1e-14, d, 1.0), 1.0)
    x = np.zeros_like(b)
    r = b.copy()
    z = dinv * r
    p = z.copy()
    rz = r @ z
    for _ in range(k):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x - x.mean() if walls else x


@pytest.mark.parametrize("per_x", [False, True], ids=["walls", "periodic_x"])
def test_pcg_truncated_iterates(per_x):
    g = O.Grid(10, 8, lx=1.0, ly=0.8, per_x=per_x)
    rng = np.random.default_rng(12)
    c, pr, _, _ = two_phase_case(g, rng, rho=(100.0, 1.0))
    if per_x:
        O.halo(g, pr["rho_u"], "u"); O.halo(g, pr["rho_v"], "v")
    rhs = rand(g, "p", rng)
    sl = interior(g, "p")
    if per_x:
        rhs[sl] -= rhs[sl].mean()
    A = neumann_pressure_matrix(g, pr["rho_u"], pr["rho_v"])
    for k in (1, 2, 5, 12, 25):
        phi = g.zeros("p")
        it, _ = O.poisson_pcg(g, phi, rhs, pr["rho_u"], pr["rho_v"], tol=1e-300, max_iter=k)
        assert it == k
        want = jacobi_pcg(A, rhs[sl].ravel(), k, walls=not per_x)
        got = phi[sl].ravel()
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), k
