"""Restatement of the reference's checkpoint and VTK formats (src/mod_io.f90) at one rank.

TEST INFRASTRUCTURE ONLY (the checker of the GPU library's writers and reader).

write_checkpoint / read_checkpoint   src/mod_io.f90:169-225: "<prefix>_r0000.bin", stream access,
    int32 step, real64 t, int32 nx, ny, nh (24 bytes, no record markers), then u, v, u_prev, v_prev,
    p, C as whole padded arrays (column-major, i fastest), native byte order.  The reader refuses a
    missing file or other dimensions (ok = False).
write_vtk   src/mod_io.f90:29-164: legacy ASCII STRUCTURED_POINTS with VOF_C, Pressure, Density,
    Curvature (one '(ES15.7)' line per value) and the cell-centred velocity (one '(3ES15.7)' line per
    point), i fastest.
fortran_es  Fortran ES editing (Fortran 2008, 10.7.2.3.2) as the reference's edit descriptors use it:
    the correctly rounded significand d.ddddddd, an exponent E+dd for |e| <= 99 and +ddd beyond, right
    justified in the field width.  Pinned against a Fortran runtime by tests/golden/.
"""
from __future__ import annotations

import math
import os

import numpy as np


def fortran_es(x: float, w: int = 15, d: int = 7) -> str:
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Inf" if x > 0 else "-Inf"
    else:
        mant, ex = f"{x:.{d}E}".split("E")
        e = int(ex)
        sign = "+" if e >= 0 else "-"
        s = f"{mant}E{sign}{abs(e):02d}" if abs(e) <= 99 else f"{mant}{sign}{abs(e):03d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_i(n: int, w: int) -> str:
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else "*" * w


def checkpoint_name(prefix: str, rank: int = 0) -> str:
    return f"{prefix}_r{rank:04d}.bin"


def write_checkpoint(prefix, step, t, u, v, u_prev, v_prev, p, c, nx, ny, nh=2):
    with open(checkpoint_name(prefix), "wb") as f:
        f.write(np.array([step], "=i4").tobytes() + np.array([t], "=f8").tobytes() +
                np.array([nx, ny, nh], "=i4").tobytes())
        for a in (u, v, u_prev, v_prev, p, c):
            f.write(np.ascontiguousarray(a, dtype="=f8").tobytes())


def read_checkpoint(prefix, nx, ny, nh=2):
    """(ok, step, t, [u, v, u_prev, v_prev, p, c]) with the arrays in the padded (rows, ld) shape."""
    name = checkpoint_name(prefix)
    if not os.path.exists(name):
        return False, None, None, None
    raw = open(name, "rb").read()
    step = int(np.frombuffer(raw, "=i4", 1, 0)[0])
    t = float(np.frombuffer(raw, "=f8", 1, 4)[0])
    nxf, nyf, nhf = (int(x) for x in np.frombuffer(raw, "=i4", 3, 12))
    if (nxf, nyf, nhf) != (nx, ny, nh):
        return False, step, t, None
    shapes = [(ny + 2 * nh, nx + 1 + 2 * nh), (ny + 1 + 2 * nh, nx + 2 * nh)] * 2 + [(ny + 2 * nh, nx + 2 * nh)] * 2
    off, out = 24, []
    for sh in shapes:
        n = sh[0] * sh[1]
        out.append(np.frombuffer(raw, "=f8", n, off).reshape(sh).copy())
        off += 8 * n
    return True, step, t, out


def write_vtk(filename, p, c, rho_c, kappa_curv, u, v, nx, ny, dx, dy, nh=2):
    ins = (slice(nh, nh + ny), slice(nh, nh + nx))
    uc = 0.5 * (u[nh:nh + ny, nh:nh + nx] + u[nh:nh + ny, nh + 1:nh + nx + 1])     # half * (u(i-1) + u(i))
    vc = 0.5 * (v[nh:nh + ny, nh:nh + nx] + v[nh + 1:nh + ny + 1, nh:nh + nx])
    lines = ["# vtk DataFile Version 3.0", "FUGU-2D Coupled Two-Phase Solver Output", "ASCII",
             "DATASET STRUCTURED_POINTS",
             "DIMENSIONS " + fortran_i(nx, 8) + fortran_i(ny, 8) + fortran_i(1, 8),
             "ORIGIN " + fortran_es(0.5 * dx) + fortran_es(0.5 * dy) + fortran_es(0.0),
             "SPACING " + fortran_es(dx) + fortran_es(dy) + fortran_es(dx),
             "POINT_DATA " + fortran_i(nx * ny, 10)]
    for name, a in (("VOF_C", c), ("Pressure", p), ("Density", rho_c), ("Curvature", kappa_curv)):
        lines += [f"SCALARS {name} double 1", "LOOKUP_TABLE default"]
        lines += [fortran_es(x) for x in a[ins].ravel().tolist()]
    lines.append("VECTORS Velocity double")
    lines += [fortran_es(a) + fortran_es(b) + fortran_es(0.0) for a, b in zip(uc.ravel().tolist(), vc.ravel().tolist())]
    with open(filename, "w") as f:
        f.write("\n".join(lines) + "\n")
