"""GPU tests of the multigrid pressure block (FUGU_PRECOND_MG, no reference counterpart):
V-cycles converge to the dense solution of the reference's pressure operator, and one
V-cycle is a fixed linear map (what the outer Krylov method needs)."""
import numpy as np
import pytest

import pyoracle as O
from cases import interior, rand, two_phase_case
from test_oracle_solvers import neumann_pressure_matrix

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")


@pytest.mark.parametrize("per_x", [False, True])
def test_mg_converges_to_dense_solution(per_x):
    g = O.Grid(32, 24, 1.0, 0.75, per_x=per_x)
    rng = np.random.default_rng(200)
    c, pr, _, _ = two_phase_case(g, rng, rho=(100.0, 1.0))
    if per_x:
        O.halo(g, pr["rho_u"], "u"); O.halo(g, pr["rho_v"], "v")
    rhs = rand(g, "p", rng)
    sl = interior(g, "p")
    rhs[sl] -= rhs[sl].mean()
    A = neumann_pressure_matrix(g, pr["rho_u"], pr["rho_v"])
    x = np.linalg.lstsq(A, rhs[sl].ravel(), rcond=None)[0]
    x -= x.mean()
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly, per_x=per_x)
    errs = []
    for n in (1, 2, 4, 40):
        phi = g.zeros("p")
        ctx.pressure_mg(phi, rhs, pr["rho_u"], pr["rho_v"], n_cycles=n)
        got = phi[sl].ravel() - phi[sl].mean()
        errs.append(np.abs(got - x).max() / np.abs(x).max())
    assert errs[-1] < 1e-9, errs
    assert errs[1] < errs[0] and errs[2] < errs[1], errs
    assert errs[0] < 0.5, errs


def test_mg_vcycle_is_linear():
    g = O.Grid(64, 32, 2.0, 1.0)
    rng = np.random.default_rng(201)
    c, pr, _, _ = two_phase_case(g, rng, rho=(1000.0, 1.0))
    ctx = abi.Context(g.nx, g.ny, g.lx, g.ly)
    b1, b2 = rand(g, "p", rng), rand(g, "p", rng)
    out = []
    for b in (b1, b2, 3.0 * b1 - 2.0 * b2):
        phi = g.zeros("p")
        ctx.pressure_mg(phi, b, pr["rho_u"], pr["rho_v"], n_cycles=1)
        out.append(phi)
    lin = 3.0 * out[0] - 2.0 * out[1]
    assert np.abs(out[2] - lin).max() <= 1e-12 * np.abs(lin).max()
