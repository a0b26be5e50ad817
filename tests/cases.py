"""Shared case builders for the oracle and GPU parity tests (seeded, small)."""
from __future__ import annotations

import numpy as np

import pyoracle as O


def phys(rho1=1.0, rho2=1.0, mu1=1.0, mu2=1.0, sigma=0.0, gx=0.0, gy=0.0, time_order=1,
         alpha_penalty=1.0e12, cfl_target=0.5):
    return O.OgPhys(rho1, rho2, mu1, mu2, sigma, gx, gy, time_order, alpha_penalty, cfl_target)


def smooth_c(g: O.Grid, rng, n_blobs=3):
    """A volume fraction made of a few circles (sharp 0/1, like the reference initializers),
    sampled at cell centres over the full padded array (src/main.f90:137-151)."""
    X, Y = g.coords("p")
    c = np.zeros(g.shape("p"))
    for _ in range(n_blobs):
        xc, yc = rng.uniform(0.25, 0.75) * g.lx, rng.uniform(0.25, 0.75) * g.ly
        r = rng.uniform(0.1, 0.2) * min(g.lx, g.ly)
        c = np.maximum(c, (np.hypot(X - xc, Y - yc) <= r).astype(np.float64))
    return c


def rand(g: O.Grid, kind, rng, scale=1.0):
    return rng.standard_normal(g.shape(kind)) * scale


def props_from_c(g, c, rho1, rho2, mu1, mu2):
    return O.mixture(g, c, rho1, rho2, mu1, mu2)


def two_phase_case(g, rng, rho=(1000.0, 1.0), mu=(1.0, 0.01)):
    c = smooth_c(g, rng)
    O.halo(g, c, "p")
    pr = props_from_c(g, c, rho[0], rho[1], mu[0], mu[1])
    u_adv, v_adv = rand(g, "u", rng, 0.1), rand(g, "v", rng, 0.1)
    O.velocity_bcs(g, u_adv, v_adv)
    return c, pr, u_adv, v_adv


def droplet_c(g, R=0.25, xc=0.5, yc=0.5, inside=1.0):
    X, Y = g.coords("p")
    r = np.hypot(X - xc, Y - yc)
    return np.where(r <= R, inside, 1.0 - inside).astype(np.float64)


def interior(g, kind):
    """Slices of the owned interior of each field kind (no ghosts)."""
    nh, nx, ny = g.nh, g.nx, g.ny
    if kind == "u":
        return (slice(nh, nh + ny), slice(nh, nh + nx + 1))
    if kind == "v":
        return (slice(nh, nh + ny + 1), slice(nh, nh + nx))
    if kind == "p":
        return (slice(nh, nh + ny), slice(nh, nh + nx))
    return (slice(nh, nh + ny + 1), slice(nh, nh + nx + 1))
