"""The Fortran drop-in module (coupled_codefort_2d_amd/fortran/fugu2d_hip.f90) driven by a
reference-style Fortran program: two advance_coupled_step calls of a 32x32 static droplet
(reference BiCGStab + PCD, BDF2) and a coupled_dot, through iso_c_binding into the HIP
library, checked against the oracle (u, v, p within 1e-10 relative, C within 1e-12,
identical iteration counts)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "coupled_codefort_2d_amd", "fortran", "_build", "dropin_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran check program not built (no flang)")
def test_fortran_dropin_matches_oracle(tmp_path):
    out = tmp_path / "dropin.bin"
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = O.Grid(32, 32)
    ph = O.OgPhys(1000.0, 1.0, 1.0, 0.01, 1.0, 0.0, 0.0, 2, 1.0e12, 0.5)
    st = O.State(g)
    X, Y = g.coords("p")
    st.c[...] = (np.hypot(X - 0.5, Y - 0.5) <= 0.25).astype(np.float64)
    st.init_properties(ph)
    raw = out.read_bytes()
    off = 0
    shapes = [st.u.shape, st.v.shape, st.p.shape, st.c.shape]
    for step in (1, 2):
        it_f = int(np.frombuffer(raw, np.int32, 1, off)[0]); off += 4
        res_f, dot_f = np.frombuffer(raw, np.float64, 2, off); off += 16
        fields = []
        for shp in shapes:
            n = shp[0] * shp[1]
            fields.append(np.frombuffer(raw, np.float64, n, off).reshape(shp)); off += 8 * n
        it_o, res_o = O.advance_step(st, ph, 1e-4, step, use_bicgstab2=False, tol=1e-8, max_iter=40)
        assert it_f == it_o, (step, it_f, it_o)
        assert abs(res_f - res_o) <= 1e-10 * max(abs(res_o), 1e-300), (step, res_f, res_o)
        for name, a, tol in zip("uvpc", fields, (1e-10, 1e-10, 1e-10, 1e-12)):
            b = getattr(st, name)
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
            assert err <= tol, (step, name, err)
        dot_o = O.coupled_dot(g, [st.u, st.v, st.p], [st.u, st.v, st.p])
        assert abs(dot_f - dot_o) <= 1e-12 * abs(dot_o), (step, dot_f, dot_o)
    assert off == len(raw)
