"""Pins of the oracle's solver restatements (src/mod_subsolvers.f90, src/mod_pcd_precond.f90,
src/mod_bicgstab2.f90) against dense linear algebra on operators assembled from their physical
definitions, and against the reference's own integration tests."""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, phys, two_phase_case, droplet_c


def neumann_pressure_matrix(g, rho_u, rho_v):
    """-div(rho^-1 grad phi) with zero-flux walls / periodic wrap, finite-volume fluxes
    (phi_i - phi_nb) / (rho_face h^2) -- assembled from the physics."""
    nx, ny = g.nx, g.ny
    n = nx * ny
    A = np.zeros((n, n))
    k = lambda i, j: (i - 1) + (j - 1) * nx
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            row = k(i, j)
            nbrs = []
            if i < nx or g.per_x:
                nbrs.append((i % nx + 1, j, 1.0 / (rho_u[g.iu(i, j)] * g.dx ** 2)))
            if i > 1 or g.per_x:
                nbrs.append(((i - 2) % nx + 1, j, 1.0 / (rho_u[g.iu(i - 1, j)] * g.dx ** 2)))
            if j < ny or g.per_y:
                nbrs.append((i, j % ny + 1, 1.0 / (rho_v[g.iv(i, j)] * g.dy ** 2)))
            if j > 1 or g.per_y:
                nbrs.append((i, (j - 2) % ny + 1, 1.0 / (rho_v[g.iv(i, j - 1)] * g.dy ** 2)))
            for (i2, j2, cf) in nbrs:
                A[row, row] += cf
                A[row, k(i2, j2)] -= cf
    return A


@pytest.mark.parametrize("per_x", [False, True])
def test_pcg_converges_to_dense_solution(per_x):
    g = O.Grid(10, 8, lx=1.0, ly=0.8, per_x=per_x)
    rng = np.random.default_rng(10)
    c, pr, _, _ = two_phase_case(g, rng, rho=(100.0, 1.0))
    if per_x:
        O.halo(g, pr["rho_u"], "u"); O.halo(g, pr["rho_v"], "v")
    rhs = rand(g, "p", rng)
    sl = interior(g, "p")
    if per_x:   # no mean projection with a periodic direction: feed a compatible rhs
        rhs[sl] -= rhs[sl].mean()
    phi = g.zeros("p")
    it, res = O.poisson_pcg(g, phi, rhs, pr["rho_u"], pr["rho_v"], tol=1e-14, max_iter=5000)
    assert res < 1e-13
    A = neumann_pressure_matrix(g, pr["rho_u"], pr["rho_v"])
    b = rhs[sl].ravel() - (0 if per_x else rhs[sl].mean())
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    x -= x.mean()
    got = phi[sl].ravel() - phi[sl].mean()
    assert np.abs(got - x).max() < 1e-9 * np.abs(x).max()
    if not per_x:
        assert abs(phi[sl].mean()) < 1e-13 * np.abs(x).max()     # mean removal (mod_subsolvers:227-241)


def test_pcg_truncated_reports_true_residual_and_cap():
    """The 150-iteration / 1e-5 PCG of the preconditioner: iteration count is capped and the
    reported relative residual is the true one of the returned iterate."""
    g = O.Grid(48, 48)
    rng = np.random.default_rng(11)
    c, pr, _, _ = two_phase_case(g, rng, rho=(1000.0, 1.0))
    rhs = rand(g, "p", rng)
    phi = g.zeros("p")
    it, res = O.poisson_pcg(g, phi, rhs, pr["rho_u"], pr["rho_v"])
    assert it == 150 and res > 1e-5
    A = neumann_pressure_matrix(g, pr["rho_u"], pr["rho_v"])
    sl = interior(g, "p")
    b = rhs[sl].ravel() - rhs[sl].mean()
    true_res = np.linalg.norm(b - A @ phi[sl].ravel()) / np.linalg.norm(b)
    assert abs(true_res - res) < 1e-8 * res
    # a short run converges and stops at the first iterate below tol
    g2 = O.Grid(8, 8)
    c2, pr2, _, _ = two_phase_case(g2, rng, rho=(2.0, 1.0))
    rhs2 = rand(g2, "p", rng)
    phi2 = g2.zeros("p")
    it2, res2 = O.poisson_pcg(g2, phi2, rhs2, pr2["rho_u"], pr2["rho_v"])
    assert it2 < 150 and res2 < 1e-5
    # zero rhs: early exit, phi zeroed (mod_subsolvers.f90:115-120)
    phi3 = np.ones(g2.shape("p"))
    assert O.poisson_pcg(g2, phi3, g2.zeros("p"), pr2["rho_u"], pr2["rho_v"]) == (0, 0.0)
    assert np.all(phi3 == 0)


def helmholtz_u_matrix(g, pr, dt):
    """Diagonal block rho/dt - d/dx(kappa d/dx) - d/dy((zeta-eta) d/dy) on u faces i=1..nx-1,
    with the Jacobi smoother's zero values outside the swept rows (wall faces held at rhs)."""
    nx, ny = g.nx, g.ny
    dofs = [(i, j) for j in range(1, ny + 1) for i in range(1, nx)]
    pos = {d: n for n, d in enumerate(dofs)}
    A = np.zeros((len(dofs), len(dofs)))
    for (i, j), r in pos.items():
        cE = pr["kappa"][g.ip(i + 1, j)] / g.dx ** 2
        cW = pr["kappa"][g.ip(i, j)] / g.dx ** 2
        cN = (pr["zeta"][g.iw(i, j)] - pr["eta"][g.iw(i, j)]) / g.dy ** 2
        cS = (pr["zeta"][g.iw(i, j - 1)] - pr["eta"][g.iw(i, j - 1)]) / g.dy ** 2
        A[r, r] = pr["rho_u"][g.iu(i, j)] / dt + cE + cW + cN + cS
        for (i2, j2), cf in (((i + 1, j), cE), ((i - 1, j), cW), ((i, j + 1), cN), ((i, j - 1), cS)):
            if (i2, j2) in pos:
                A[r, pos[(i2, j2)]] -= cf
    return dofs, A


def test_jacobi_u_converges_to_block_solve():
    """Many damped-Jacobi sweeps (omega = 0.8) converge to the diagonal-block solve; the wall
    faces carry the rhs (mod_subsolvers.f90:302-305)."""
    g = O.Grid(8, 6)
    rng = np.random.default_rng(12)
    c, pr, _, _ = two_phase_case(g, rng, rho=(10.0, 1.0), mu=(1.0, 0.1))
    rhs = rand(g, "u", rng)
    # zero the wall-face rhs so the interior block decouples from the identity rows
    nh = g.nh
    rhs[:, nh] = 0.0
    rhs[:, nh + g.nx] = 0.0
    dt = 0.05
    x = g.zeros("u")
    O.helmholtz_u(g, x, rhs, pr, dt, n_sweeps=3000)
    dofs, A = helmholtz_u_matrix(g, pr, dt)
    b = np.array([rhs[g.iu(i, j)] for (i, j) in dofs])
    sol = np.linalg.solve(A, b)
    got = np.array([x[g.iu(i, j)] for (i, j) in dofs])
    assert np.abs(got - sol).max() < 1e-11 * np.abs(sol).max()
    # three sweeps from zero: first sweep is 0.8 * rhs / diag exactly
    x1 = g.zeros("u")
    O.helmholtz_u(g, x1, rhs, pr, dt, n_sweeps=1)
    assert np.allclose([x1[g.iu(i, j)] for (i, j) in dofs], 0.8 * b / np.diag(A), rtol=1e-14)


def test_pcd_blocks():
    """apply_pcd_precond: zero ghosts at walls on all outputs; a pure pressure input produces
    r_p = -(phi/dt + conv) with phi the PCG solution (mean zero)."""
    g = O.Grid(12, 10)
    rng = np.random.default_rng(13)
    c, pr, ua, va = two_phase_case(g, rng)
    Z = (rand(g, "u", rng), rand(g, "v", rng), rand(g, "p", rng))
    R = O.pcd(g, Z, pr, ua, va, 0.01)
    nh = g.nh
    assert np.all(R[2][:nh] == 0) and np.all(R[2][:, :nh] == 0) and np.all(R[2][-nh:] == 0)
    assert np.all(R[0][:nh] == 0) and np.all(R[0][:, :nh] == 0)
    assert np.all(R[1][:nh] == 0) and np.all(R[1][:, -nh:] == 0)
    # wall faces of r_u carry z_u (Jacobi boundary rows, mod_subsolvers.f90:302-305)
    assert np.array_equal(R[0][nh:nh + g.ny, nh], Z[0][nh:nh + g.ny, nh])
    # linear in z when the inner PCG is not truncated: tiny grid converges before the cap
    g2 = O.Grid(6, 6)
    c2, pr2, ua2, va2 = two_phase_case(g2, rng, rho=(2.0, 1.0), mu=(1.0, 0.5))
    Z2 = (rand(g2, "u", rng), rand(g2, "v", rng), rand(g2, "p", rng))
    R1 = O.pcd(g2, Z2, pr2, ua2, va2, 0.01)
    R2 = O.pcd(g2, tuple(2 * z for z in Z2), pr2, ua2, va2, 0.01)
    for a, b in zip(R1, R2):
        assert np.allclose(2 * a, b, rtol=1e-4, atol=1e-6 * np.abs(b).max())


def coupled_dense_operator(g, pr, ua, va, dt):
    """The coupled operator on owned unknowns (u i=1..nx-1, v j=1..ny-1, all p) with ghosts by
    velocity-BC reflection -- the map the Krylov iterates see after apply_velocity_bcs."""
    dofs = [("u", i, j) for j in range(1, g.ny + 1) for i in range(1, g.nx)] + \
           [("v", i, j) for j in range(1, g.ny) for i in range(1, g.nx + 1)] + \
           [("p", i, j) for j in range(1, g.ny + 1) for i in range(1, g.nx + 1)]
    ix = {"u": g.iu, "v": g.iv, "p": g.ip}
    n = len(dofs)
    A = np.zeros((n, n))
    for col, (k, i, j) in enumerate(dofs):
        X = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
        X["uvp".index(k)][ix[k](i, j)] = 1.0
        O.velocity_bcs(g, X[0], X[1])
        Y = O.matvec(g, X, pr, ua, va, dt)
        A[:, col] = [Y["uvp".index(k2)][ix[k2](i2, j2)] for (k2, i2, j2) in dofs]
    return dofs, A, ix


@pytest.mark.parametrize("order,consistent", [(1, False), (2, True)])
def test_krylov_converges_to_dense_solution(order, consistent):
    """BiCGStab, and BiCGStab(2) with the velocity BCs on r and s, + PCD to a tight tolerance
    reproduce the dense solve of the coupled saddle system (velocity, and pressure up to its
    constant)."""
    g = O.Grid(8, 8)
    rng = np.random.default_rng(14)
    c, pr, ua, va = two_phase_case(g, rng, rho=(10.0, 1.0), mu=(1.0, 0.1))
    dt = 0.01
    B = [rand(g, "u", rng), rand(g, "v", rng), g.zeros("p")]
    # rhs only on owned rows (the reference assembles B_u for i<=i_u_max, B_v for j<=j_v_max)
    mask_u = np.zeros(g.shape("u"), bool); mask_u[g.nh:g.nh + g.ny, g.nh + 1:g.nh + g.nx] = True
    mask_v = np.zeros(g.shape("v"), bool); mask_v[g.nh + 1:g.nh + g.ny, g.nh:g.nh + g.nx] = True
    B[0][~mask_u] = 0; B[1][~mask_v] = 0
    X = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    it, res = O.krylov(g, X, B, pr, ua, va, dt, tol=1e-11, max_iter=400, order=order, consistent=consistent)
    assert res < 1e-11 or (order == 2 and res < 1e-5)
    dofs, A, ix = coupled_dense_operator(g, pr, ua, va, dt)
    b = np.array([B["uvp".index(k)][ix[k](i, j)] for (k, i, j) in dofs])
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    got = np.array([X["uvp".index(k)][ix[k](i, j)] for (k, i, j) in dofs])
    nvel = sum(1 for d in dofs if d[0] != "p")
    # With the reference PCD (inner PCG stopped at 1e-5) the preconditioner is not a fixed linear
    # map, so the recurrence residual drifts from the true one and 1e-4 is what this form allows.
    # The tight pins (1e-8 against dense solves, closed-form first iterates, textbook iterates)
    # run the PCD's PCG to 1e-13 instead: tests/test_oracle_krylov_pins.py.
    assert np.abs(got[:nvel] - x[:nvel]).max() < 1e-4 * np.abs(x[:nvel]).max()
    pg, px = got[nvel:] - got[nvel:].mean(), x[nvel:] - x[nvel:].mean()
    assert np.abs(pg - px).max() < 1e-4 * np.abs(px).max()


def test_krylov_iteration_bookkeeping():
    """iter_out = min(k, max_iter) with the Fortran DO-exit convention; a converged initial
    guess exits at k = 1 without touching X (besides BCs)."""
    g = O.Grid(8, 8)
    rng = np.random.default_rng(15)
    c, pr, ua, va = two_phase_case(g, rng, rho=(10.0, 1.0), mu=(1.0, 0.1))
    B = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    X = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    it, res = O.krylov(g, X, B, pr, ua, va, 0.01, tol=1e-8, max_iter=5, order=2)
    assert it == 1 and res == 0.0
    B = [rand(g, "u", rng), rand(g, "v", rng), g.zeros("p")]
    it, res = O.krylov(g, X, B, pr, ua, va, 0.01, tol=1e-30, max_iter=3, order=2)
    assert it == 3


def test_reference_bicgstab2_stalls_on_walls():
    """src/mod_bicgstab2.f90:430,489 pass r and s to matvec_coupled without velocity BCs;
    on a walled domain the reference's BiCGStab(2) therefore does not converge.  The
    restatement keeps that behaviour; adding the two BC calls restores convergence."""
    g = O.Grid(8, 8)
    rng = np.random.default_rng(14)
    c, pr, ua, va = two_phase_case(g, rng, rho=(10.0, 1.0), mu=(1.0, 0.1))
    B = [rand(g, "u", rng), rand(g, "v", rng), g.zeros("p")]
    B[0][:, :g.nh + 1] = 0; B[0][:, g.nh + g.nx:] = 0
    res = {}
    for consistent in (False, True):
        X = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
        res[consistent] = O.krylov(g, X, B, pr, ua, va, 0.01, tol=1e-12, max_iter=20, order=2,
                                   consistent=consistent)[1]
    assert res[False] > 1e-2 and res[True] < 1e-4
