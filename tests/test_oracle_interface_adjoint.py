"""The oracle's transposes of the interface chain (og_vof_advection_T, og_svof_T, og_curvature_T,
og_csf_T: the reverse mode of src/mod_vof.f90:38-194 and src/mod_surface_tension.f90:25-109 at a
given state) pinned by central differences of the oracle's own forward routines:

    < ybar, (F(x + h d) - F(x - h d)) / 2h >  =  < F'(x)^T ybar, d >

for random weights ybar on every output and random directions d on every input, over whole padded
arrays.  CSF is bilinear and the SVOF sweeps linear at a fixed count (exact up to rounding); the
curvature is smooth; VOF is piecewise smooth, so its states keep every face away from the kinks
(monotone profiles strictly inside (0, 1): no clamp, |den| > 1e-12, r > 0, Superbee's r away from
0.5, 1 and 2), with both flow directions on every face kind, on wall and periodic grids and with
both limiters.  A dropped term, a wrong sign or a wrong stencil cell fails these checks."""
import numpy as np
import pytest

import pyoracle as O


def dot(a, b):
    return float(sum(np.vdot(x, y) for x, y in zip(a, b)))


def grids():
    return [O.Grid(14, 11, 1.0, 0.8), O.Grid(12, 10, 1.2, 1.0, per_x=True, per_y=True),
            O.Grid(10, 12, 1.0, 1.1, per_x=True)]


def smooth_c(g, kind="bump"):
    X, Y = g.coords("p")
    if kind == "mono":   # strictly increasing in x and y, curved (r != 1), inside (0.1, 0.9)
        xs, ys = X / g.lx, Y / g.ly
        return 0.1 + 0.35 * (xs + 0.3) ** 2 / 1.69 + 0.35 * (ys + 0.2) ** 2 / 1.44 + 0.05 * xs * ys
    r = np.hypot(X - 0.5 * g.lx, Y - 0.5 * g.ly)
    return 0.5 * (1.0 - np.tanh((r - 0.3) / 0.08))


def test_csf_transpose_central_differences():
    rng = np.random.default_rng(11)
    for g in grids():
        c, kc = smooth_c(g) + 0.1 * rng.standard_normal(g.shape("p")), rng.standard_normal(g.shape("p"))
        dc, dk = rng.standard_normal(c.shape), rng.standard_normal(kc.shape)
        fub, fvb = rng.standard_normal(g.shape("u")), rng.standard_normal(g.shape("v"))
        h, sigma = 1e-4, 0.7
        p = O.csf(g, c + h * dc, kc + h * dk, sigma)
        m = O.csf(g, c - h * dc, kc - h * dk, sigma)
        fd = dot((fub, fvb), ((p[0] - m[0]) / (2 * h), (p[1] - m[1]) / (2 * h)))
        cb, kb = O.csf_T(g, c, kc, sigma, fub, fvb)
        ad = dot((cb, kb), (dc, dk))
        assert abs(fd - ad) <= 1e-10 * max(abs(ad), 1.0), (fd, ad)


def test_curvature_transpose_central_differences():
    rng = np.random.default_rng(12)
    for g in grids():
        cs = smooth_c(g) + 0.02 * rng.standard_normal(g.shape("p"))
        d, kb = rng.standard_normal(cs.shape), rng.standard_normal(cs.shape)
        h = 1e-6
        fd = np.vdot(kb, (O.curvature(g, cs + h * d) - O.curvature(g, cs - h * d)) / (2 * h))
        ad = np.vdot(O.curvature_T(g, cs, kb), d)
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1.0), (fd, ad)


@pytest.mark.parametrize("stop", ["fixed", "converged"])
def test_svof_transpose_central_differences(stop):
    rng = np.random.default_rng(13)
    for g in grids():
        c = smooth_c(g) + 0.1 * rng.standard_normal(g.shape("p"))
        delta = 1.5 * min(g.dx, g.dy)
        iters = 7 if stop == "fixed" else O.svof(g, c, delta, 150, 1e-7)[1]
        assert iters >= 2
        # tol 0: exactly `iters` sweeps, the linear map the transpose is of
        F = lambda x: O.svof(g, x, delta, iters, 0.0)[0]
        d, sb = rng.standard_normal(c.shape), rng.standard_normal(c.shape)
        h = 1e-3
        fd = np.vdot(sb, (F(c + h * d) - F(c - h * d)) / (2 * h))
        ad = np.vdot(O.svof_T(g, sb, delta, iters), d)
        assert abs(fd - ad) <= 1e-10 * max(abs(ad), 1.0), (g.per_x, g.per_y, fd, ad)


@pytest.mark.parametrize("limiter", [0, 1])
def test_vof_transpose_central_differences(limiter):
    rng = np.random.default_rng(14 + limiter)
    O.set_vof_limiter(limiter)
    try:
        for g in grids():
            c = smooth_c(g, "mono")
            # both flow directions on every face kind; |u| dt / h well below 1
            u = 0.6 * np.sign(rng.standard_normal(g.shape("u"))) * (0.5 + rng.random(g.shape("u")))
            v = 0.6 * np.sign(rng.standard_normal(g.shape("v"))) * (0.5 + rng.random(g.shape("v")))
            dt = 0.2 * min(g.dx, g.dy)
            dc, du, dv = (rng.standard_normal(a.shape) for a in (c, u, v))
            cb = rng.standard_normal(c.shape)
            h = 1e-7
            p = O.vof_advection(g, c + h * dc, u + h * du, v + h * dv, dt)
            m = O.vof_advection(g, c - h * dc, u - h * du, v - h * dv, dt)
            fd = np.vdot(cb, (p - m) / (2 * h))
            gc, gu, gv = O.vof_advection_T(g, c, u, v, dt, cb)
            ad = dot((gc, gu, gv), (dc, du, dv))
            assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1.0), (g.per_x, g.per_y, fd, ad)
    finally:
        O.set_vof_limiter(0)


def test_vof_transpose_flat_regions_pass_nothing():
    """The one-sided conventions on a sharp droplet (C = 0 or 1 away from the interface, uniform
    flow): a cell whose 5 x 5 neighbourhood is flat has an unclamped update of exactly its own value
    (0 or 1, where the clamp passes nothing) and faces whose stencils see |den| = 0, so it receives
    no adjoint at all, and no face inside a flat region passes any to the velocity; the interface
    band receives some."""
    g = O.Grid(32, 32)
    X, Y = g.coords("p")
    c = (np.hypot(X - 0.5, Y - 0.5) <= 0.25).astype(np.float64)
    u, v = 0.3 * np.ones(g.shape("u")), -0.2 * np.ones(g.shape("v"))
    nh = g.nh
    cb = np.zeros(g.shape("p"))
    cb[nh:nh + g.ny, nh:nh + g.nx] = 1.0
    gc, gu, gv = O.vof_advection_T(g, c, u, v, 0.01, cb)
    flat = np.zeros(c.shape, bool)
    for j in range(nh + 2, nh + g.ny - 2):
        for i in range(nh + 2, nh + g.nx - 2):
            w = c[j - 2:j + 3, i - 2:i + 3]
            flat[j, i] = w.min() == w.max()
    assert flat.sum() > 50 and not flat.all()
    assert np.all(gc[flat] == 0.0)
    band = ~flat[nh + 2:nh + g.ny - 2, nh + 2:nh + g.nx - 2]
    assert np.abs(gc[nh + 2:nh + g.ny - 2, nh + 2:nh + g.nx - 2][band]).max() > 0.0
    # a u face (column q of the u array) lies between the cells of p columns q - 1 and q; when both
    # are flat its whole stencil is
    for j in range(c.shape[0]):
        for q in range(1, c.shape[1]):
            if flat[j, q - 1] and flat[j, q]:
                assert gu[j, q] == 0.0, (j, q)
