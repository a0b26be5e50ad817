"""Pins of the checkpoint / VTK restatement (oracle/io_ref.py, src/mod_io.f90) against a Fortran
runtime: the fixtures in tests/golden/ were written by a Fortran program compiled with flang
(tests/golden/make_io_fixtures.py) using the reference's edit descriptors and stream layout."""
import os

import numpy as np

import io_ref as IO

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_es_editing_matches_fortran_runtime():
    raw = open(os.path.join(GOLD, "fortran_format_values.bin"), "rb").read()
    n = int(np.frombuffer(raw, "=i4", 1, 0)[0])
    vals = np.frombuffer(raw, "=f8", n, 4).tolist()
    lines = [IO.fortran_es(x) for x in vals]
    lines += [IO.fortran_es(vals[k]) + IO.fortran_es(vals[k + 1]) + IO.fortran_es(vals[k + 2])
              for k in range(0, n - 2, 3)]
    lines.append("DIMENSIONS " + IO.fortran_i(128, 8) + IO.fortran_i(2048, 8) + IO.fortran_i(1, 8))
    lines.append("ORIGIN " + IO.fortran_es(0.5 / 128) + IO.fortran_es(0.5 / 96) + IO.fortran_es(0.0))
    lines.append("SPACING " + IO.fortran_es(1.0 / 128) + IO.fortran_es(1.0 / 96) + IO.fortran_es(1.0 / 128))
    lines.append("POINT_DATA " + IO.fortran_i(4194304, 10))
    lines.append("POINT_DATA " + IO.fortran_i(12, 10))
    want = open(os.path.join(GOLD, "fortran_format_out.txt")).read()
    assert "\n".join(lines) + "\n" == want


def fixture_arrays(nx=4, ny=3, nh=2):
    """The values the fixture program fills its arrays with, in Fortran's storage order."""
    shapes = [(ny + 2 * nh, nx + 1 + 2 * nh), (ny + 1 + 2 * nh, nx + 2 * nh)] * 2 + [(ny + 2 * nh, nx + 2 * nh)] * 2
    out, k = [], 0
    for sh in shapes:
        n = sh[0] * sh[1]
        ks = np.arange(k + 1, k + n + 1)
        out.append((np.sin(0.37 * ks) * 10.0 ** (ks % 7) - 0.5 * ks).reshape(sh))
        k += n
    return out


def test_checkpoint_bytes_match_fortran_stream(tmp_path):
    ref = open(os.path.join(GOLD, "fortran_checkpoint_4x3.bin"), "rb").read()
    assert len(ref) == 24 + 8 * sum(a.size for a in fixture_arrays())
    src = str(tmp_path / "fortran")
    open(IO.checkpoint_name(src), "wb").write(ref)          # as the reference names a rank-0 file
    ok, step, t, arrs = IO.read_checkpoint(src, 4, 3)
    assert ok and step == 10 and t == 0.1
    prefix = str(tmp_path / "chk")
    IO.write_checkpoint(prefix, step, t, *arrs, 4, 3)
    assert open(IO.checkpoint_name(prefix), "rb").read() == ref
    # the arrays hold the fixture program's values to its sin's rounding (one ulp)
    for a, b in zip(arrs, fixture_arrays()):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14)
    # wrong dimensions are refused
    assert IO.read_checkpoint(src, 5, 3)[0] is False
    assert IO.read_checkpoint(str(tmp_path / "missing"), 4, 3)[0] is False
