"""Seeded synthetic workloads (numpy, host side) in the reference's padded layout.

rising_bubble_c: C = 0 inside r(theta) = r0 (1 + 0.01 sum_m a_m cos(m theta + phi_m)), 1 outside,
    sampled at cell centres over the full padded array (the initialiser of src/main.f90:137-151
    moved to (0.5, 0.6) with a seeded perturbation; a_m ~ U[-1, 1] / 8, phi_m ~ U[0, 2 pi)).

inclusions: non-overlapping circular inclusions in a carrier on the unit square -- radii
    U[rmin, rmax], centres by rejection sampling with a gap >= 2h between circles and a wall
    clearance >= r + 2h; carrier C = 1, rho_carrier, mu_carrier; inclusion k: C = 0, rho_incl,
    mu_carrier * 10^U[0, log10(mu_ratio)] (per inclusion).  Returns C and the per-cell mu, rho
    fields over the padded cells.
"""
from __future__ import annotations

import numpy as np


def cell_centres(nx, ny, lx, ly, nh=2):
    dx, dy = lx / nx, ly / ny
    i = np.arange(1 - nh, nx + nh + 1, dtype=np.float64)
    j = np.arange(1 - nh, ny + nh + 1, dtype=np.float64)
    return np.meshgrid((i - 0.5) * dx, (j - 0.5) * dy)


def rising_bubble_c(nx, ny, lx=1.0, ly=2.0, nh=2, r0=0.2, xc=0.5, yc=0.6, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, 8) / 8.0
    ph = rng.uniform(0.0, 2.0 * np.pi, 8)
    X, Y = cell_centres(nx, ny, lx, ly, nh)
    th = np.arctan2(Y - yc, X - xc)
    r = np.hypot(X - xc, Y - yc)
    R = r0 * (1.0 + 0.01 * sum(a[m - 1] * np.cos(m * th + ph[m - 1]) for m in range(1, 9)))
    return np.where(r <= R, 0.0, 1.0)


def inclusions(n, count, rmin, rmax, seed, rho_carrier, mu_carrier, rho_incl, mu_ratio, nh=2,
               max_tries=200000):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    circles = []
    tries = 0
    while len(circles) < count:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {count} inclusions")
        r = rng.uniform(rmin, rmax)
        xc, yc = rng.uniform(r + 2 * h, 1.0 - r - 2 * h, 2)
        if all(np.hypot(xc - x, yc - y) >= r + q + 2 * h for x, y, q in circles):
            circles.append((xc, yc, r))
    mus = mu_carrier * 10.0 ** rng.uniform(0.0, np.log10(mu_ratio), count)
    shape = (n + 2 * nh, n + 2 * nh)
    c = np.ones(shape)
    mu = np.full(shape, mu_carrier)
    rho = np.full(shape, rho_carrier)
    for (xc, yc, r), m in zip(circles, mus):
        # padded index k <-> cell i = k + 1 - nh, centre (i - 1/2) h; test only the bounding box
        k0 = max(int(np.floor((xc - r) / h + 0.5)) - 1 + nh - 1, 0)
        k1 = min(int(np.ceil((xc + r) / h + 0.5)) + 1 + nh, shape[1])
        l0 = max(int(np.floor((yc - r) / h + 0.5)) - 1 + nh - 1, 0)
        l1 = min(int(np.ceil((yc + r) / h + 0.5)) + 1 + nh, shape[0])
        xs = (np.arange(k0, k1) + 1 - nh - 0.5) * h
        ys = (np.arange(l0, l1) + 1 - nh - 0.5) * h
        inside = np.hypot(xs[None, :] - xc, ys[:, None] - yc) <= r
        c[l0:l1, k0:k1][inside] = 0.0
        mu[l0:l1, k0:k1][inside] = m
        rho[l0:l1, k0:k1][inside] = rho_incl
    return c, mu, rho, circles
