"""Physics diagnostics of the reference's validation programs, on padded host arrays (numpy, the
reference's layout with i fastest).  Each function restates the reference test's definition.

laplace_jump        mean p inside r < 0.7 R minus mean p outside r > 1.3 R, against sigma / R
                    (tests/test_laplace_coupled.f90:130-158)
curvature_mean      mean kappa over the cells with |r - R| <= 1.5 dx, against 1 / R
                    (tests/test_static_droplet.f90:85-105)
spurious_velocity   max |u|, |v| over the interior points (tests/test_laplace_coupled.f90:160-166)
mass_error          |sum C - sum C_0| / sum C_0 over the interior cells (test_laplace_coupled.f90:119-127)
div_max             max |div u| over the cells (test_laplace_coupled.f90:168-175, op_divergence)
cavity_extrema      min u on the vertical centre line (u faces i = nx/2), min and max v on the
                    horizontal centre line (v faces j = ny/2) (tests/test_lid_driven_cavity.f90:200-242)
"""
from __future__ import annotations

import numpy as np


def _centres(nx, ny, lx, ly):
    dx, dy = lx / nx, ly / ny
    x = (np.arange(1, nx + 1) - 0.5) * dx
    y = (np.arange(1, ny + 1) - 0.5) * dy
    return np.meshgrid(x, y)   # interior cells, shape (ny, nx)


def interior(a, nx, ny, nh=2, kind="p"):
    if kind == "u":
        return a[nh:nh + ny, nh:nh + nx + 1]       # i = 0..nx
    if kind == "v":
        return a[nh:nh + ny + 1, nh:nh + nx]       # j = 0..ny
    return a[nh:nh + ny, nh:nh + nx]


def laplace_jump(p, nx, ny, lx, ly, R, sigma, xc=0.5, yc=0.5, nh=2):
    X, Y = _centres(nx, ny, lx, ly)
    r = np.hypot(X - xc, Y - yc)
    pi = interior(p, nx, ny, nh)
    p_in, p_out = pi[r < 0.7 * R].mean(), pi[r > 1.3 * R].mean()
    dp = p_in - p_out
    exact = sigma / R
    return {"delta_p": float(dp), "exact": exact, "rel_error": float(abs(dp - exact) / exact)}


def curvature_mean(kappa, nx, ny, lx, ly, R, xc=0.5, yc=0.5, nh=2):
    X, Y = _centres(nx, ny, lx, ly)
    r = np.hypot(X - xc, Y - yc)
    sel = np.abs(r - R) <= 1.5 * (lx / nx)
    k = interior(kappa, nx, ny, nh)[sel].mean() if sel.any() else 0.0
    return {"kappa_mean": float(k), "exact": 1.0 / R, "rel_error": float(abs(k - 1.0 / R) * R)}


def spurious_velocity(u, v, nx, ny, nh=2):
    """max |u|, |v| over i = 1..nx, j = 1..ny of both padded arrays (as the reference loop reads them)."""
    return float(max(np.abs(u[nh:nh + ny, nh + 1:nh + nx + 1]).max(), np.abs(v[nh + 1:nh + ny + 1, nh:nh + nx]).max()))


def mass_error(c, c0, nx, ny, nh=2):
    m0 = interior(c0, nx, ny, nh).sum()
    return float(abs(interior(c, nx, ny, nh).sum() - m0) / max(m0, 1e-14))


def div_max(u, v, nx, ny, lx, ly, nh=2):
    dx, dy = lx / nx, ly / ny
    U, V = interior(u, nx, ny, nh, "u"), interior(v, nx, ny, nh, "v")
    d = (U[:, 1:] - U[:, :-1]) / dx + (V[1:, :] - V[:-1, :]) / dy
    return float(np.abs(d).max())


# Botella and Peyret (1998) at Re = 100, quoted by the reference's cavity test (:240-242)
BOTELLA_RE100 = {"u_min": -0.21404, "v_min": -0.25380, "v_max": 0.17957}


def cavity_extrema(u, v, nx, ny, lx=1.0, ly=1.0, nh=2):
    i, j = nx // 2, ny // 2
    ucol = u[nh:nh + ny, i + nh]             # u(i, 1..ny)
    vrow = v[j + nh, nh:nh + nx]             # v(1..nx, j)
    dy, dx = ly / ny, lx / nx
    ju, iv_min, iv_max = int(np.argmin(ucol)), int(np.argmin(vrow)), int(np.argmax(vrow))
    return {"u_min": float(ucol[ju]), "y_at_u_min": (ju + 0.5) * dy,
            "v_min": float(vrow[iv_min]), "x_at_v_min": (iv_min + 0.5) * dx,
            "v_max": float(vrow[iv_max]), "x_at_v_max": (iv_max + 0.5) * dx}
