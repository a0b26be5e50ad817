"""CPU reference of the multi-step discrete adjoint (TEST INFRASTRUCTURE ONLY: the checker of
the GPU adjoint, never the product).  No counterpart in the reference, which has no adjoint.

Forward map (the frozen-interface BDF2 horizon the GPU adjoint differentiates):
  parameters  theta = (mu_cell, rho_cell) over the padded cells, and sigma;
  properties  fixed per-cell fields -> rho_c, rho_u, rho_v, kappa, zeta, eta (og_cell_properties,
              the mixture-law edge conventions of src/mod_viscosity.f90:54-151);
  step n      u_adv = u^{n-1} (n = 1) or 2u^{n-1} - u^{n-2} (src/mod_solver_coupled.f90:209-222);
              C^n = VOF(C^{n-1}, u^{n-1}) and f_sigma^n = sigma * CSF(C^n, kappa(SVOF(C^n)); 1)
              (:224-252), recorded and held fixed by the adjoint (frozen interface);
              B_u = (rho_u/dt)(2u^{n-1} - u^{n-2}/2) (BDF2, n >= 2 and time_order 2) or
              (rho_u/dt) u^{n-1}, + rho_u g_x + f_u; B_v alike; B_p = 0 (:261-303);
              A_eff(dt_eff) X^n = B, solved exactly (sparse LU of the assembled operator, bordered
              by the zero-mean pressure gauge); u^n = E X^n (BCs and halos, :368-372).
  objective   J = 1/2 sum over owned u faces (y >= y0) of rho_u u^2 vol + the same for v faces,
              at step N.

Adjoint: for n = N..1
  A_n^T lam_n = dJ/dX^n + sum_{m = n+1, n+2} (db_m/du^n - d(A_m X^m)/du^n)^T lam_m,
and dJ/dtheta = explicit - sum_n lam_n^T (dA_n/dtheta X^n - db_n/dtheta), chained through the
property map.  The transposed solves use the transposed assembled operator (the dense/sparse
FP64 transpose-solve reference).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pyoracle as O


def owned_index(g):
    probe = O.mask_owned(g, [np.ones(g.shape(k)) for k in "uvp"])
    return [np.flatnonzero(a.ravel()) for a in probe]


class Layout:
    """Flat vectors of owned degrees of freedom <-> padded triplets."""

    def __init__(self, g):
        self.g = g
        self.idx = owned_index(g)
        self.sizes = [len(i) for i in self.idx]
        self.n = sum(self.sizes)
        self.off = np.cumsum([0] + self.sizes)

    def flat(self, X):
        return np.concatenate([X[k].ravel()[self.idx[k]] for k in range(3)])

    def trip(self, x):
        X = [self.g.zeros(t) for t in "uvp"]
        for k in range(3):
            X[k].ravel()[self.idx[k]] = x[self.off[k]:self.off[k + 1]]
        return X


def assemble(g, lay, props, u_adv, v_adv, dt):
    cols = []
    for col in range(lay.n):
        e = np.zeros(lay.n)
        e[col] = 1.0
        cols.append(lay.flat(O.matvec_eff(g, lay.trip(e), props, u_adv, v_adv, dt, True)))
    return sp.csc_matrix(np.array(cols).T)


def bordered(A, lay):
    """[[A, e_p], [e_p^T, 0]]: the zero-mean pressure gauge (A has the constant pressure null
    space on both sides)."""
    e = np.zeros((lay.n, 1))
    e[lay.off[2]:lay.off[3], 0] = 1.0
    return sp.bmat([[A, sp.csc_matrix(e)], [sp.csc_matrix(e.T), None]]).tocsc()


class Case:
    def __init__(self, g, mu_cell, rho_cell, sigma, gx, gy, dt, nsteps, time_order=2, y0_frac=0.9,
                 u0=None, v0=None, c0=None):
        self.g, self.mu, self.rho, self.sigma = g, mu_cell, rho_cell, sigma
        self.gx, self.gy, self.dt, self.nsteps, self.time_order = gx, gy, dt, nsteps, time_order
        self.y0 = y0_frac * g.ly
        self.u0 = g.zeros("u") if u0 is None else u0
        self.v0 = g.zeros("v") if v0 is None else v0
        self.c0 = g.zeros("p") if c0 is None else c0

    def with_params(self, mu=None, rho=None, sigma=None):
        return Case(self.g, self.mu if mu is None else mu, self.rho if rho is None else rho,
                    self.sigma if sigma is None else sigma, self.gx, self.gy, self.dt, self.nsteps,
                    self.time_order, self.y0 / self.g.ly, self.u0, self.v0, self.c0)


def objective_masks(g, y0):
    """Owned faces with centre y >= y0 (u faces: (j - 1/2) dy; v faces: j dy)."""
    mu_ = O.mask_owned(g, [np.ones(g.shape("u")), np.ones(g.shape("v")), g.zeros("p")])
    ju = np.arange(1 - g.nh, g.ny + g.nh + 1)[:, None]
    jv = np.arange(-g.nh, g.ny + g.nh + 1)[:, None]
    mu_[0] = mu_[0] * ((ju - 0.5) * g.dy >= y0)
    mu_[1] = mu_[1] * (jv * g.dy >= y0)
    return mu_[0], mu_[1]


def forward(case: Case, frozen=None):
    """J and the trajectory.  frozen: the per-step unit-sigma CSF forces of another run (the
    recorded interface), used instead of recomputing them from this run's C -- the map the
    adjoint differentiates, so the one finite differences must use."""
    g = case.g
    lay = Layout(g)
    pr = O.cell_properties(g, case.mu, case.rho)
    u, v = case.u0.copy(), case.v0.copy()
    O.velocity_bcs(g, u, v, 0.0)
    O.halo(g, u, "u"); O.halo(g, v, "v")
    u_prev, v_prev = u.copy(), v.copy()
    c = case.c0.copy()
    traj = []
    for n in range(1, case.nsteps + 1):
        ua = u.copy() if n <= 1 else 2.0 * u - u_prev
        va = v.copy() if n <= 1 else 2.0 * v - v_prev
        unm1, vnm1 = u_prev, v_prev
        u_prev, v_prev = u.copy(), v.copy()
        c = O.vof_advection(g, c, u, v, case.dt)
        if frozen is not None:
            fhu, fhv = frozen[n - 1]["fhu"], frozen[n - 1]["fhv"]
        elif case.sigma > 1e-14:
            cs, _ = O.svof(g, c, 1.5 * min(g.dx, g.dy), 150, 1e-7)
            kc = O.curvature(g, cs)
            fhu, fhv = O.csf(g, c, kc, 1.0)
        else:
            fhu, fhv = g.zeros("u"), g.zeros("v")
        bdf2 = case.time_order == 2 and n >= 2
        dte = case.dt / 1.5 if bdf2 else case.dt
        hu = (2.0 * u - 0.5 * unm1) if bdf2 else u
        hv = (2.0 * v - 0.5 * vnm1) if bdf2 else v
        B = [pr["rho_u"] / case.dt * hu + pr["rho_u"] * case.gx + case.sigma * fhu,
             pr["rho_v"] / case.dt * hv + pr["rho_v"] * case.gy + case.sigma * fhv, g.zeros("p")]
        O.mask_owned(g, B)
        A = assemble(g, lay, pr, ua, va, dte)
        Ab = bordered(A, lay)
        sol = spla.spsolve(Ab, np.concatenate([lay.flat(B), [0.0]]))
        X = lay.trip(sol[:lay.n])
        Xe = O.extend(g, X)
        u, v = Xe[0], Xe[1]
        traj.append(dict(X=X, Xe=Xe, ua=ua, va=va, fhu=fhu, fhv=fhv, hu=hu, hv=hv, bdf2=bdf2, dte=dte, A=A))
    mu_, mv_ = objective_masks(g, case.y0)
    vol = g.dx * g.dy
    J = 0.5 * vol * (np.sum(mu_ * pr["rho_u"] * u * u) + np.sum(mv_ * pr["rho_v"] * v * v))
    return J, traj, pr, lay


def gradient(case: Case):
    """(J, dJ/dmu_cell, dJ/drho_cell, dJ/dsigma, lambdas)."""
    g = case.g
    J, traj, pr, lay = forward(case)
    N = case.nsteps
    vol = g.dx * g.dy
    mu_, mv_ = objective_masks(g, case.y0)
    gc = {k: g.zeros(kind) for k, kind in (("rho_c", "p"), ("rho_u", "u"), ("rho_v", "v"), ("kappa", "p"),
                                           ("zeta", "w"), ("eta", "w"))}
    uN, vN = traj[-1]["Xe"][0], traj[-1]["Xe"][1]
    gc["rho_u"] += 0.5 * vol * mu_ * uN * uN
    gc["rho_v"] += 0.5 * vol * mv_ * vN * vN
    g_sigma = 0.0
    lam = [None] * (N + 1)
    for n in range(N, 0, -1):
        t = traj[n - 1]
        rhs = [g.zeros(k) for k in "uvp"]
        if n == N:
            rhs[0] += vol * mu_ * pr["rho_u"] * uN
            rhs[1] += vol * mv_ * pr["rho_v"] * vN
        for m, c_adv in ((n + 1, 2.0), (n + 2, -1.0)):
            if m > N:
                continue
            tm = traj[m - 1]
            L = lam[m]
            # db_m/du^n: u^n enters B_m as u^{m-1} (coefficient 2 or 1) or u^{m-2} (-1/2)
            if m == n + 1:
                cb = 2.0 if tm["bdf2"] else 1.0
            else:
                cb = -0.5 if tm["bdf2"] else 0.0
            rhs[0] += cb * pr["rho_u"] / case.dt * L[0]
            rhs[1] += cb * pr["rho_v"] / case.dt * L[1]
            # -d(A_m X^m)/du^n through u_adv^m = 2u^{m-1} - u^{m-2} (m >= 2)
            if m >= 2:
                gr = O.operator_coef_grad(g, L, tm["Xe"], pr, tm["ua"], tm["va"], tm["dte"], True, 1.0)
                Ea = O.extend_T(g, [gr["u_adv"], gr["v_adv"], g.zeros("p")])
                rhs[0] -= c_adv * Ea[0]
                rhs[1] -= c_adv * Ea[1]
        O.mask_owned(g, rhs)
        At = bordered(t["A"], lay).T.tocsc()
        sol = spla.spsolve(At, np.concatenate([lay.flat(rhs), [0.0]]))
        lam[n] = lay.trip(sol[:lay.n])
        L = lam[n]
        # - lam^T dA/dcoef X^n
        gr = O.operator_coef_grad(g, L, t["Xe"], pr, t["ua"], t["va"], t["dte"], True, -1.0)
        for k in gc:
            gc[k] += gr[k]
        # + lam^T dB/dcoef: rho_u (hist/dt + g_x), sigma f_hat
        Lm = O.mask_owned(g, [L[0].copy(), L[1].copy(), g.zeros("p")])
        gc["rho_u"] += Lm[0] * (t["hu"] / case.dt + case.gx)
        gc["rho_v"] += Lm[1] * (t["hv"] / case.dt + case.gy)
        g_sigma += np.sum(Lm[0] * t["fhu"]) + np.sum(Lm[1] * t["fhv"])
    g_mu, g_rho = O.cell_properties_T(g, case.mu, gc)
    return J, g_mu, g_rho, g_sigma, lam
