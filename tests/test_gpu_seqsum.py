"""Reference-order sums on adversarial term sequences (coupled_dot, src/mod_bicgstab2.f90:32-72).

With dx = dy = 1 (vol = 1) and b = 1 every term a*b*vol is exactly a, so coupled_dot adds the
chosen values one by one in the reference's loop order (u faces, v faces, cells; j outer, i
inner).  The GPU's parallel reproduction of that sequential sum (chunked binade certification,
exact grid sums of each run, a serial pass over the runs) must return the same double as a plain
Python loop over the same terms, on sequences built to defeat it: random walks around zero, ties
on the rounding grid, heavy cancellation, thirty decades of magnitude, exact powers of two,
zeros, infinities and NaNs.  Sizes span one to ~70 chunks of 4096 terms (up to ~530 chunks of
512); every chunk length the library uses (512, 1024, 4096, and its choice by the sum's length)
is run."""
import numpy as np
import pytest

from cases import interior
import pyoracle as O

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")


def owned_terms(g, a):
    """The terms of coupled_dot(a, 1) with vol = 1, in the reference's order."""
    out = []
    for k, kind in enumerate("uvp"):
        A = a[k][interior(g, kind)]
        if kind == "u":
            A = A[:, 1:] if g.per_x else A[:, 1:-1]
        if kind == "v":
            A = A[1:, :] if g.per_y else A[1:-1, :]
        out.append(A.ravel())
    return np.concatenate(out)


def seq(terms):
    s = 0.0
    for x in terms.tolist():
        s = s + x
    return s


def fill(g, rng, kind):
    def gen(n):
        if kind == "positive":
            return rng.uniform(0, 1, n)
        if kind == "walk":
            return rng.standard_normal(n)
        if kind == "decades":
            return rng.standard_normal(n) * 10.0 ** rng.uniform(-15, 15, n)
        if kind == "ties":
            return np.floor(rng.standard_normal(n) * 8) * 0.125
        if kind == "cancel":
            return np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * (1 + rng.uniform(0, 1e-12, n)) * 1e10
        if kind == "zeros":
            return np.where(rng.uniform(0, 1, n) < 0.3, 0.0, rng.standard_normal(n) * 1e-3 + 1e-3)
        if kind == "pow2":
            return np.ldexp(1.0, rng.integers(-30, 30, n)) * rng.choice([-1.0, 1.0], n)
        if kind == "products":
            return rng.standard_normal(n) * rng.standard_normal(n) * 0.0625
        if kind == "halves":
            x = np.full(n, 0.1)
            x[n // 2:] = -0.1
            return x + rng.standard_normal(n) * 1e-17
        raise ValueError(kind)
    return [gen(int(np.prod(g.shape(k)))).reshape(g.shape(k)) for k in "uvp"]


KINDS = ["positive", "walk", "decades", "ties", "cancel", "zeros", "pow2", "products", "halves"]


@pytest.mark.parametrize("shape", [(30, 20, False, False), (120, 100, True, False), (160, 150, False, True)],
                         ids=["1chunk", "9chunks_px", "18chunks_py"])
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("chunk", ["0", "512", "1024", "4096"])
def test_reference_order_sum_bitwise(monkeypatch, shape, kind, chunk):
    monkeypatch.setenv("FUGU_SEQ_CHUNK", chunk)
    nx, ny, px, py = shape
    g = O.Grid(nx, ny, float(nx), float(ny), per_x=px, per_y=py)   # dx = dy = 1
    ctx = abi.Context(nx, ny, float(nx), float(ny), px, py, sum_order=abi.SUM_REFERENCE)
    rng = np.random.default_rng([nx, ny, KINDS.index(kind)])
    a = fill(g, rng, kind)
    ones = [np.ones(g.shape(k)) for k in "uvp"]
    want = seq(owned_terms(g, a))
    got = ctx.coupled_dot(a, ones)
    assert got == want and np.signbit(got) == np.signbit(want), (got, want)
    ctx.close()


@pytest.mark.parametrize("chunk", ["0", "512", "1024", "4096"])
def test_reference_order_sum_large_and_nonfinite(monkeypatch, chunk):
    """~70 chunks of 4096 terms (a 270k-term sum), then the same with an infinity and a NaN planted."""
    monkeypatch.setenv("FUGU_SEQ_CHUNK", chunk)
    nx = ny = 300
    g = O.Grid(nx, ny, float(nx), float(ny))
    ctx = abi.Context(nx, ny, float(nx), float(ny), sum_order=abi.SUM_REFERENCE)
    rng = np.random.default_rng(7)
    ones = [np.ones(g.shape(k)) for k in "uvp"]
    for kind in ("walk", "products", "positive"):
        a = fill(g, rng, kind)
        want = seq(owned_terms(g, a))
        assert ctx.coupled_dot(a, ones) == want, kind
    a = fill(g, rng, "walk")
    a[2][g.nh + 100, g.nh + 100] = np.inf
    assert ctx.coupled_dot(a, ones) == seq(owned_terms(g, a)) == np.inf
    a[0][g.nh + 50, g.nh + 50] = -np.inf
    assert np.isnan(ctx.coupled_dot(a, ones))
    a = fill(g, rng, "walk")
    a[1][g.nh + 7, g.nh + 9] = np.nan
    assert np.isnan(ctx.coupled_dot(a, ones))
    ctx.close()
