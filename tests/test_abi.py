"""CPU checks of the drop-in boundary: the HIP library loads and exports every symbol that
include/fugu2d.h declares (no GPU compute is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "fugu2d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fugu_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_reference_procedures():
    syms = declared_symbols()
    for name in ("fugu_matvec_coupled", "fugu_apply_velocity_bcs", "fugu_coupled_dot",
                 "fugu_bicgstab_solve", "fugu_bicgstab2_solve", "fugu_apply_pcd_precond",
                 "fugu_solve_poisson_pressure", "fugu_solve_helmholtz_u", "fugu_solve_helmholtz_v",
                 "fugu_vof_advection_tvd", "fugu_svof_helmholtz_filter",
                 "fugu_compute_interface_curvature", "fugu_compute_csf_force",
                 "fugu_compute_mixture_densities", "fugu_compute_mixture_viscosities",
                 "fugu_compute_adaptive_dt", "fugu_advance_coupled_step"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    from coupled_codefort_2d_amd import abi
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(abi.EXPORTS) <= set(declared_symbols())
    abi.load()   # argtypes declared for all bound entry points


def test_no_cpu_fallback_without_library(tmp_path):
    from coupled_codefort_2d_amd import abi
    with pytest.raises(FileNotFoundError):
        abi.load.__wrapped__(str(tmp_path / "missing.so")) if hasattr(abi.load, "__wrapped__") else \
            _load_fresh(abi, str(tmp_path / "missing.so"))


def _load_fresh(abi, path):
    saved = abi._lib
    abi._lib = None
    try:
        abi.load(path)
    finally:
        abi._lib = saved
