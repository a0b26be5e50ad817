"""The CPU adjoint reference (oracle/adjoint_ref.py) against central finite differences of its
own forward map (exact sparse solves), per parameter family: per-cell mu, per-cell rho and sigma
over a multi-step BDF2 horizon with advection, surface tension and gravity, at 1e-6 relative.
The adjoint solves use the transposed assembled operator (the FP64 transpose-solve reference).
The finite differences hold the interface (C, hence the unit-sigma CSF force) at the base run's
trajectory, the frozen-interface map the adjoint differentiates; without that the difference is
the interface-motion sensitivity the frozen map leaves out (checked to be small, not zero)."""
import numpy as np
import pytest

import pyoracle as O
import adjoint_ref as AR
from cases import droplet_c


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


def fd(case, which, d, eps, frozen):
    if which == "mu":
        jp = AR.forward(case.with_params(mu=case.mu + eps * d), frozen)[0]
        jm = AR.forward(case.with_params(mu=case.mu - eps * d), frozen)[0]
    elif which == "rho":
        jp = AR.forward(case.with_params(rho=case.rho + eps * d), frozen)[0]
        jm = AR.forward(case.with_params(rho=case.rho - eps * d), frozen)[0]
    else:
        jp = AR.forward(case.with_params(sigma=case.sigma + eps * d), frozen)[0]
        jm = AR.forward(case.with_params(sigma=case.sigma - eps * d), frozen)[0]
    return (jp - jm) / (2 * eps)


@pytest.mark.parametrize("per_x,time_order", [(False, 2), (True, 2), (False, 1)],
                         ids=["walls_bdf2", "periodic_x_bdf2", "walls_bdf1"])
def test_adjoint_gradient_matches_central_differences(per_x, time_order):
    case = make_case(per_x, time_order)
    J, g_mu, g_rho, g_sigma, _ = AR.gradient(case)
    frozen = AR.forward(case)[1]
    assert J > 0
    rng = np.random.default_rng(31)
    dmu = case.mu * rng.standard_normal(case.mu.shape)
    drho = case.rho * rng.standard_normal(case.rho.shape)
    for which, d, grad in (("mu", dmu, np.sum(g_mu * dmu)), ("rho", drho, np.sum(g_rho * drho)),
                           ("sigma", 1.0, g_sigma)):
        f = fd(case, which, d, 1e-4, frozen)
        assert abs(f - grad) <= 1e-6 * abs(f), (which, f, grad)
    # the moving interface changes the answer a little: the frozen map is an approximation
    f = fd(case, "rho", drho, 1e-4, None)
    assert 1e-9 < abs(f - np.sum(g_rho * drho)) <= 1e-2 * abs(f)


def test_adjoint_gradient_localised_cells():
    """Single-cell perturbations (interior, near a wall, a ghost cell) of mu and rho."""
    case = make_case()
    J, g_mu, g_rho, g_sigma, _ = AR.gradient(case)
    frozen = AR.forward(case)[1]
    g = case.g
    for (i, j) in ((6, 8), (1, 1), (0, 5)):
        r, c = g.ip(i, j)
        d = np.zeros(case.mu.shape)
        d[r, c] = case.mu[r, c]
        f = fd(case, "mu", d, 1e-4, frozen)
        assert abs(f - g_mu[r, c] * d[r, c]) <= 1e-6 * abs(f) + 1e-14 * abs(J), ("mu", i, j, f, g_mu[r, c] * d[r, c])
        d = np.zeros(case.rho.shape)
        d[r, c] = case.rho[r, c]
        f = fd(case, "rho", d, 1e-4, frozen)
        assert abs(f - g_rho[r, c] * d[r, c]) <= 1e-6 * abs(f) + 1e-14 * abs(J), ("rho", i, j, f, g_rho[r, c] * d[r, c])
