"""fugu_coupled_dot_noalloc (device pointers, no context, scratch in a fixed device array)
against fugu_coupled_dot on the same inputs, in both summation orders: bit-identical results on
two calls in a row, the second on a smaller grid (the shared scratch keeps the first call's
contents), and the size limit of the reference order rejected with an error."""
import numpy as np
import pytest

import pyoracle as O
from cases import rand

pytestmark = pytest.mark.gpu
abi = pytest.importorskip("coupled_codefort_2d_amd.abi")
torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("order", [abi.SUM_REFERENCE, abi.SUM_TREE], ids=["reference", "tree"])
def test_noalloc_dot_matches_context_dot(order):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    for nx, ny, px, py, seed in ((300, 200, False, True, 1), (40, 24, True, False, 2)):
        g = O.Grid(nx, ny, 1.3, 0.9, per_x=px, per_y=py)
        rng = np.random.default_rng(seed)
        a = [rand(g, k, rng) for k in "uvp"]
        b = [rand(g, k, rng) for k in "uvp"]
        ctx = abi.Context(nx, ny, 1.3, 0.9, px, py, sum_order=order)
        want = ctx.coupled_dot(a, b)
        ctx.close()
        da, db = [dev(x) for x in a], [dev(x) for x in b]
        torch.cuda.synchronize()
        grid = abi.FuguGrid(nx, ny, 2, 1.3, 0.9, int(px), int(py))
        abi.coupled_dot_noalloc(grid, [x.data_ptr() for x in da], [x.data_ptr() for x in db], out.data_ptr(), order)
        torch.cuda.synchronize()
        got = float(out.cpu()[0])
        assert got == want, (nx, ny, got, want)
        if order == abi.SUM_REFERENCE:
            O.set_sum_mode(0)
            assert got == O.coupled_dot(g, a, b)


def test_noalloc_dot_rejects_oversized_reference_order():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    g = abi.FuguGrid(600, 600, 2, 1.0, 1.0, 0, 0)   # 1,078,800 owned points > 2^20
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(abi.FuguError, match="scratch limit"):
        abi.coupled_dot_noalloc(g, [x.data_ptr()] * 3, [x.data_ptr()] * 3, x.data_ptr(), abi.SUM_REFERENCE)
