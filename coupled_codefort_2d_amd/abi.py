"""ctypes binding of libfugu2d_hip.so (include/fugu2d.h) for Python callers (tests, bench).

Raises at import-time use when the HIP library is missing: there is no CPU fallback.
Field arrays are numpy float64 arrays in the reference's padded Fortran layout, C order
with i fastest (shape (rows, ld)), exactly as the C ABI expects.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfugu2d_hip.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class FuguGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nh", C.c_int), ("lx", C.c_double),
                ("ly", C.c_double), ("periodic_x", C.c_int), ("periodic_y", C.c_int)]


class FuguParams(C.Structure):
    _fields_ = [("rho1", C.c_double), ("rho2", C.c_double), ("mu1", C.c_double), ("mu2", C.c_double),
                ("sigma", C.c_double), ("gx", C.c_double), ("gy", C.c_double), ("time_order", C.c_int),
                ("alpha_penalty", C.c_double), ("cfl_target", C.c_double)]


class FuguStepOpts(C.Structure):
    _fields_ = [("use_bicgstab2", C.c_int), ("has_lid", C.c_int), ("u_lid", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("use_amp", C.c_int),
                ("u_solid", C.c_double), ("v_solid", C.c_double), ("solver_flags", C.c_uint),
                ("inner_tol", C.c_double), ("inner_max_iter", C.c_int), ("mg_cycles", C.c_int),
                ("restart", C.c_int)]


class FuguStepStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("outer", C.c_int), ("res", C.c_double), ("true_res", C.c_double),
                ("solve_seconds", C.c_double), ("step_seconds", C.c_double), ("pcg_calls", C.c_int),
                ("vol_before", C.c_double), ("vol_change", C.c_double), ("vol_clamp", C.c_double)]


class FuguAdjointOpts(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("dt", C.c_double), ("y0", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int), ("restart", C.c_int), ("mg_cycles", C.c_int), ("flags", C.c_uint),
                ("checkpoint_every", C.c_int)]


class FuguAdjointStats(C.Structure):
    _fields_ = [("J", C.c_double), ("forward_seconds", C.c_double), ("reverse_seconds", C.c_double),
                ("forward_iters", C.c_int), ("adjoint_iters", C.c_int), ("max_forward_res", C.c_double),
                ("max_adjoint_res", C.c_double)]


ADJ_FORWARD_ONLY = 1
ADJ_REUSE_INTERFACE = 2
ADJ_PRECOND_VMG = 4
ADJ_MIXED_FORWARD = 8
ADJ_MOVING_INTERFACE = 16
ADJ_MIXED_ADJOINT = 32

B2_CONSISTENT_BCS = 1
SOLVE_REFINED = 2
SOLVE_MIXED = 4
PRECOND_MG = 8
PRECOND_VISC = 16
PRECOND_VMG = 32
SOLVE_EXTRAPOLATE = 64

SUM_REFERENCE = 0   # the reference's sequential summation order (default)
SUM_TREE = 1        # double-double tree (fast)
LIMITER_VAN_LEER = 0   # the reference's VOF limiter (default)
LIMITER_SUPERBEE = 1   # opt-in

EXPORTS = [
    "fugu_ctx_create", "fugu_ctx_destroy", "fugu_set_params", "fugu_last_error", "fugu_field_size",
    "fugu_set_sum_order", "fugu_get_sum_order", "fugu_set_vof_limiter",
    "fugu_exchange_halos", "fugu_op_divergence", "fugu_apply_velocity_bcs", "fugu_matvec_coupled",
    "fugu_coupled_dot", "fugu_solve_poisson_pressure", "fugu_solve_helmholtz_u", "fugu_solve_helmholtz_v",
    "fugu_apply_pcd_precond", "fugu_bicgstab_solve", "fugu_bicgstab2_solve", "fugu_vof_advection_tvd",
    "fugu_svof_helmholtz_filter", "fugu_compute_interface_curvature", "fugu_compute_csf_force",
    "fugu_compute_mixture_densities", "fugu_compute_mixture_viscosities", "fugu_compute_adaptive_dt",
    "fugu_advance_coupled_step", "fugu_state_upload", "fugu_state_download", "fugu_field_download",
    "fugu_state_init_properties", "fugu_step", "fugu_kernel_timing", "fugu_kernel_time",
    "fugu_solve_pressure_mg", "fugu_matvec_transpose", "fugu_apply_pcd_transpose", "fugu_adjoint_solve",
    "fugu_cell_properties", "fugu_adjoint_gradient", "fugu_coupled_dot_noalloc",
    "fugu_checkpoint_write", "fugu_checkpoint_read", "fugu_write_vtk",
    "fugu_vof_advection_transpose", "fugu_svof_transpose", "fugu_curvature_transpose", "fugu_csf_transpose",
]

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library; raises FileNotFoundError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m coupled_codefort_2d_amd.build`")
    L = C.CDLL(path)
    _declare(L)
    _lib = L
    return L


def _declare(L):
    def s(name, args, res=C.c_int):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    P = C.c_void_p
    s("fugu_ctx_create", [C.POINTER(FuguGrid), C.POINTER(FuguParams), C.POINTER(P)])
    s("fugu_ctx_destroy", [P])
    s("fugu_set_params", [P, C.POINTER(FuguParams)])
    s("fugu_set_sum_order", [P, C.c_int])
    s("fugu_get_sum_order", [P])
    s("fugu_set_vof_limiter", [P, C.c_int])
    s("fugu_last_error", [], C.c_char_p)
    s("fugu_field_size", [P, C.c_int], C.c_longlong)
    s("fugu_exchange_halos", [P, _D, C.c_int])
    s("fugu_op_divergence", [P, _D, _D, _D])
    s("fugu_apply_velocity_bcs", [P, _D, _D, _D])
    s("fugu_matvec_coupled", [P] + [_D] * 14 + [C.c_double, C.c_int] + [_D] * 4)
    s("fugu_coupled_dot", [P] + [_D] * 7)
    # device pointers (integers) for the allocation-free variant
    s("fugu_coupled_dot_noalloc", [C.POINTER(FuguGrid)] + [C.c_void_p] * 7 + [C.c_int, C.c_void_p])
    s("fugu_solve_poisson_pressure", [P, _D, _D, _D, _D, C.c_double, C.c_int, _I, _D])
    s("fugu_solve_helmholtz_u", [P] + [_D] * 6 + [C.c_double, C.c_int, _D])
    s("fugu_solve_helmholtz_v", [P] + [_D] * 6 + [C.c_double, C.c_int, _D])
    s("fugu_apply_pcd_precond", [P] + [_D] * 13 + [C.c_double, _D, _D])
    kry = [P] + [_D] * 14 + [C.c_double, C.c_double, C.c_int, _I, _D, _D, _D, _D, _D, _D]
    s("fugu_bicgstab_solve", kry)
    s("fugu_bicgstab2_solve", kry + [C.c_uint])
    s("fugu_vof_advection_tvd", [P, _D, _D, _D, _D, C.c_double, _D, _D])
    s("fugu_svof_helmholtz_filter", [P, _D, _D, C.c_double, C.c_int, C.c_double, _I])
    s("fugu_compute_interface_curvature", [P, _D, _D])
    s("fugu_compute_csf_force", [P, _D, _D, _D, _D])
    s("fugu_compute_mixture_densities", [P, _D, _D, _D, _D])
    s("fugu_compute_mixture_viscosities", [P, _D, _D, _D, _D])
    s("fugu_compute_adaptive_dt", [P, _D, _D, C.c_double, _D])
    s("fugu_advance_coupled_step", [P] + [_D] * 15 + [C.c_double, C.c_int, C.POINTER(FuguStepOpts), _D, _D, _I, _D])
    s("fugu_state_upload", [P] + [_D] * 8)
    s("fugu_state_download", [P] + [_D] * 6)
    s("fugu_field_download", [P, C.c_char_p, _D])
    s("fugu_state_init_properties", [P])
    s("fugu_checkpoint_write", [P, C.c_char_p, C.c_int, C.c_double])
    s("fugu_checkpoint_read", [P, C.c_char_p, _I, _D, _I])
    s("fugu_write_vtk", [P, C.c_char_p])
    s("fugu_step", [P, C.c_double, C.c_int, C.POINTER(FuguStepOpts), C.POINTER(FuguStepStats)])
    s("fugu_kernel_timing", [P, C.c_int])
    s("fugu_solve_pressure_mg", [P, _D, _D, _D, _D, C.c_int])
    s("fugu_kernel_time", [P, C.c_char_p, _D, C.POINTER(C.c_longlong), _D])
    s("fugu_matvec_transpose", [P] + [_D] * 14 + [C.c_double, C.c_int])
    s("fugu_vof_advection_transpose", [P, _D, _D, _D, C.c_double, _D, _D, _D, _D])
    s("fugu_svof_transpose", [P, _D, C.c_double, C.c_int, C.c_double, _D, _D, _I])
    s("fugu_curvature_transpose", [P, _D, _D, _D])
    s("fugu_csf_transpose", [P, _D, _D, C.c_double, _D, _D, _D, _D])
    s("fugu_apply_pcd_transpose", [P] + [_D] * 13 + [C.c_double, C.c_double, C.c_int, C.c_double])
    s("fugu_adjoint_solve", [P] + [_D] * 14 + [C.c_double, C.c_double, C.c_int, _I, _D])
    s("fugu_cell_properties", [P] + [_D] * 8)
    s("fugu_adjoint_gradient", [P, C.POINTER(FuguAdjointOpts), _D, _D, C.c_double, _D, _D, _D, _D, _D, _D,
                                C.POINTER(FuguAdjointStats)])


def coupled_dot_noalloc(grid: FuguGrid, a_dev, b_dev, out_dev, sum_order=SUM_REFERENCE, stream=None):
    """fugu_coupled_dot_noalloc on device addresses (ints): a_dev, b_dev are (u, v, p) triples,
    out_dev one double; asynchronous on `stream` (an int hipStream_t or None)."""
    rc = load().fugu_coupled_dot_noalloc(C.byref(grid), *[C.c_void_p(x) for x in a_dev],
                                         *[C.c_void_p(x) for x in b_dev], C.c_void_p(out_dev), int(sum_order),
                                         C.c_void_p(stream) if stream else None)
    return check(rc, "fugu_coupled_dot_noalloc")


def ptr(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class FuguError(RuntimeError):
    pass


def check(rc, what):
    if rc < 0:
        msg = load().fugu_last_error().decode(errors="replace")
        raise FuguError(f"{what} failed ({rc}): {msg}")
    return rc


class Context:
    """One device-resident FUGU-2D domain (fugu_ctx)."""

    def __init__(self, nx, ny, lx=1.0, ly=1.0, per_x=False, per_y=False, params=None, nh=2, sum_order=None):
        L = load()
        self.nx, self.ny, self.nh, self.lx, self.ly = nx, ny, nh, lx, ly
        self.per_x, self.per_y = per_x, per_y
        self.grid = FuguGrid(nx, ny, nh, lx, ly, int(per_x), int(per_y))
        self.params = params if params is not None else FuguParams(1, 1, 1, 1, 0, 0, 0, 1, 1e12, 0.5)
        h = C.c_void_p()
        check(L.fugu_ctx_create(C.byref(self.grid), C.byref(self.params), C.byref(h)), "fugu_ctx_create")
        self.h = h
        if sum_order is not None:
            self.set_sum_order(sum_order)

    def close(self):
        if getattr(self, "h", None):
            load().fugu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sum_order(self, order):
        """SUM_REFERENCE (the reference's sequential sums, bit-identical) or SUM_TREE (fast)."""
        check(load().fugu_set_sum_order(self.h, int(order)), "fugu_set_sum_order")

    def set_vof_limiter(self, limiter):
        check(load().fugu_set_vof_limiter(self.h, int(limiter)), "fugu_set_vof_limiter")

    def sum_order(self):
        return check(load().fugu_get_sum_order(self.h), "fugu_get_sum_order")

    def set_params(self, params):
        self.params = params
        check(load().fugu_set_params(self.h, C.byref(params)), "fugu_set_params")

    # -- fine API ------------------------------------------------------------------
    def halo(self, a, kind):
        check(load().fugu_exchange_halos(self.h, ptr(a), {"u": 0, "v": 1, "p": 2}[kind]), "halo")

    def divergence(self, u, v, out):
        check(load().fugu_op_divergence(self.h, ptr(u), ptr(v), ptr(out)), "divergence")

    def velocity_bcs(self, u, v, u_lid=None):
        lid = C.byref(C.c_double(u_lid)) if u_lid is not None else None
        check(load().fugu_apply_velocity_bcs(self.h, ptr(u), ptr(v), lid), "velocity_bcs")

    def matvec(self, X, Y, props, u_adv, v_adv, dt, include_advection=True, pen=None, mdot=None):
        pu, pv = pen if pen is not None else (None, None)
        mu, mv = mdot if mdot is not None else (None, None)
        check(load().fugu_matvec_coupled(self.h, ptr(X[0]), ptr(X[1]), ptr(X[2]), ptr(Y[0]), ptr(Y[1]), ptr(Y[2]),
                                         ptr(u_adv), ptr(v_adv), ptr(props["rho_c"]), ptr(props["rho_u"]),
                                         ptr(props["rho_v"]), ptr(props["kappa"]), ptr(props["zeta"]),
                                         ptr(props["eta"]), dt, int(include_advection), ptr(pu), ptr(pv),
                                         ptr(mu), ptr(mv)), "matvec")

    # -- adjoint (no counterpart in the reference) ---------------------------------------
    def matvec_transpose(self, L, O, props, u_adv, v_adv, dt, include_advection=True):
        check(load().fugu_matvec_transpose(self.h, ptr(L[0]), ptr(L[1]), ptr(L[2]), ptr(O[0]), ptr(O[1]), ptr(O[2]),
                                           ptr(u_adv), ptr(v_adv), ptr(props["rho_c"]), ptr(props["rho_u"]),
                                           ptr(props["rho_v"]), ptr(props["kappa"]), ptr(props["zeta"]),
                                           ptr(props["eta"]), dt, int(include_advection)), "matvec_transpose")

    def pcd_transpose(self, Y, Z, props, u_adv, v_adv, dt, ptol=1e-14, pmax=5000, visc=0.0):
        check(load().fugu_apply_pcd_transpose(self.h, ptr(Y[0]), ptr(Y[1]), ptr(Y[2]), ptr(Z[0]), ptr(Z[1]),
                                              ptr(Z[2]), ptr(u_adv), ptr(v_adv), ptr(props["rho_u"]),
                                              ptr(props["rho_v"]), ptr(props["kappa"]), ptr(props["zeta"]),
                                              ptr(props["eta"]), dt, ptol, pmax, visc), "pcd_transpose")

    def adjoint_solve(self, L, G, props, u_adv, v_adv, dt, tol=1e-10, max_iter=400):
        it, res = C.c_int(0), C.c_double(0)
        rc = check(load().fugu_adjoint_solve(self.h, ptr(L[0]), ptr(L[1]), ptr(L[2]), ptr(G[0]), ptr(G[1]),
                                             ptr(G[2]), ptr(u_adv), ptr(v_adv), ptr(props["rho_c"]),
                                             ptr(props["rho_u"]), ptr(props["rho_v"]), ptr(props["kappa"]),
                                             ptr(props["zeta"]), ptr(props["eta"]), dt, tol, max_iter,
                                             C.byref(it), C.byref(res)), "adjoint_solve")
        return it.value, res.value, rc

    def cell_properties(self, mu_cell, rho_cell, out):
        check(load().fugu_cell_properties(self.h, ptr(mu_cell), ptr(rho_cell), ptr(out["rho_c"]), ptr(out["rho_u"]),
                                          ptr(out["rho_v"]), ptr(out["kappa"]), ptr(out["zeta"]), ptr(out["eta"])),
              "cell_properties")

    def adjoint_gradient(self, opts: FuguAdjointOpts, mu_cell, rho_cell, sigma, c0, u0, v0, g_mu=None, g_rho=None):
        """Returns (J, g_sigma, stats, rc); g_mu, g_rho (padded p-kind arrays) are filled in place."""
        gs = C.c_double(0)
        st = FuguAdjointStats()
        rc = check(load().fugu_adjoint_gradient(self.h, C.byref(opts), ptr(mu_cell), ptr(rho_cell), sigma, ptr(c0),
                                                ptr(u0), ptr(v0), ptr(g_mu), ptr(g_rho), C.byref(gs), C.byref(st)),
                   "adjoint_gradient")
        return st.J, gs.value, st, rc

    def coupled_dot(self, a, b):
        out = C.c_double(0)
        check(load().fugu_coupled_dot(self.h, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(b[0]), ptr(b[1]), ptr(b[2]),
                                      C.byref(out)), "coupled_dot")
        return out.value

    def poisson_pcg(self, phi, rhs, rho_u, rho_v, tol=1e-5, max_iter=150):
        it, res = C.c_int(0), C.c_double(0)
        check(load().fugu_solve_poisson_pressure(self.h, ptr(phi), ptr(rhs), ptr(rho_u), ptr(rho_v), tol,
                                                 max_iter, C.byref(it), C.byref(res)), "poisson")
        return it.value, res.value

    def pressure_mg(self, phi, rhs, rho_u, rho_v, n_cycles=1):
        check(load().fugu_solve_pressure_mg(self.h, ptr(phi), ptr(rhs), ptr(rho_u), ptr(rho_v), n_cycles),
              "pressure_mg")

    def helmholtz_u(self, u_sol, rhs, props, dt, n_sweeps=5, pen_u=None):
        check(load().fugu_solve_helmholtz_u(self.h, ptr(u_sol), ptr(rhs), ptr(props["rho_u"]), ptr(props["kappa"]),
                                            ptr(props["zeta"]), ptr(props["eta"]), dt, n_sweeps, ptr(pen_u)), "hu")

    def helmholtz_v(self, v_sol, rhs, props, dt, n_sweeps=5, pen_v=None):
        check(load().fugu_solve_helmholtz_v(self.h, ptr(v_sol), ptr(rhs), ptr(props["rho_v"]), ptr(props["kappa"]),
                                            ptr(props["zeta"]), ptr(props["eta"]), dt, n_sweeps, ptr(pen_v)), "hv")

    def pcd(self, Z, R, props, u_adv, v_adv, dt, pen=None):
        pu, pv = pen if pen is not None else (None, None)
        check(load().fugu_apply_pcd_precond(self.h, ptr(Z[0]), ptr(Z[1]), ptr(Z[2]), ptr(R[0]), ptr(R[1]), ptr(R[2]),
                                            ptr(u_adv), ptr(v_adv), ptr(props["rho_u"]), ptr(props["rho_v"]),
                                            ptr(props["kappa"]), ptr(props["zeta"]), ptr(props["eta"]), dt,
                                            ptr(pu), ptr(pv)), "pcd")

    def krylov(self, X, B, props, u_adv, v_adv, dt, tol=1e-5, max_iter=100, order=2, u_lid=None, pen=None,
               mdot=None, flags=0):
        it, res = C.c_int(0), C.c_double(0)
        pu, pv = pen if pen is not None else (None, None)
        mu, mv = mdot if mdot is not None else (None, None)
        lid = C.byref(C.c_double(u_lid)) if u_lid is not None else None
        args = [self.h, ptr(X[0]), ptr(X[1]), ptr(X[2]), ptr(B[0]), ptr(B[1]), ptr(B[2]), ptr(u_adv), ptr(v_adv),
                ptr(props["rho_c"]), ptr(props["rho_u"]), ptr(props["rho_v"]), ptr(props["kappa"]),
                ptr(props["zeta"]), ptr(props["eta"]), dt, tol, max_iter, C.byref(it), C.byref(res), lid,
                ptr(pu), ptr(pv), ptr(mu), ptr(mv)]
        if order == 2:
            check(load().fugu_bicgstab2_solve(*args, flags), "bicgstab2")
        else:
            check(load().fugu_bicgstab_solve(*args), "bicgstab")
        return it.value, res.value

    def vof_advection(self, c, out, u, v, dt, fx=None, fy=None):
        check(load().fugu_vof_advection_tvd(self.h, ptr(c), ptr(out), ptr(u), ptr(v), dt, ptr(fx), ptr(fy)), "vof")

    def svof(self, c_raw, cs, delta, max_iter=100, tol=1e-6):
        it = C.c_int(0)
        check(load().fugu_svof_helmholtz_filter(self.h, ptr(c_raw), ptr(cs), delta, max_iter, tol, C.byref(it)),
              "svof")
        return it.value

    def curvature(self, cs, kc):
        check(load().fugu_compute_interface_curvature(self.h, ptr(cs), ptr(kc)), "curvature")

    def csf(self, c_raw, kc, fu, fv):
        check(load().fugu_compute_csf_force(self.h, ptr(c_raw), ptr(kc), ptr(fu), ptr(fv)), "csf")

    # transposes of the interface chain (outputs overwritten)
    def vof_advection_T(self, c_in, u, v, dt, cbar_out, cbar_in, ubar, vbar):
        check(load().fugu_vof_advection_transpose(self.h, ptr(c_in), ptr(u), ptr(v), dt, ptr(cbar_out), ptr(cbar_in),
                                                  ptr(ubar), ptr(vbar)), "vof_T")

    def svof_T(self, c_raw, delta, csbar, crawbar, max_iter=100, tol=1e-6):
        it = C.c_int(0)
        check(load().fugu_svof_transpose(self.h, ptr(c_raw), delta, max_iter, tol, ptr(csbar), ptr(crawbar),
                                         C.byref(it)), "svof_T")
        return it.value

    def curvature_T(self, cs, kbar, csbar):
        check(load().fugu_curvature_transpose(self.h, ptr(cs), ptr(kbar), ptr(csbar)), "curvature_T")

    def csf_T(self, c_raw, kc, sigma, fubar, fvbar, cbar, kbar):
        check(load().fugu_csf_transpose(self.h, ptr(c_raw), ptr(kc), sigma, ptr(fubar), ptr(fvbar), ptr(cbar),
                                        ptr(kbar)), "csf_T")

    def mixture(self, c, rho_c, rho_u, rho_v, kappa, zeta, eta):
        check(load().fugu_compute_mixture_densities(self.h, ptr(c), ptr(rho_c), ptr(rho_u), ptr(rho_v)), "dens")
        check(load().fugu_compute_mixture_viscosities(self.h, ptr(c), ptr(kappa), ptr(zeta), ptr(eta)), "visc")

    def adaptive_dt(self, u, v, dt_max):
        out = C.c_double(0)
        check(load().fugu_compute_adaptive_dt(self.h, ptr(u), ptr(v), dt_max, C.byref(out)), "adaptive_dt")
        return out.value

    def advance_step(self, st, dt, step_idx, opts: FuguStepOpts, phi_s=None):
        it, res = C.c_int(0), C.c_double(0)
        psu, psv = phi_s if phi_s is not None else (None, None)
        rc = check(load().fugu_advance_coupled_step(
            self.h, ptr(st.u), ptr(st.v), ptr(st.p), ptr(st.c), ptr(st.u_prev), ptr(st.v_prev), ptr(st.rho_c),
            ptr(st.rho_u), ptr(st.rho_v), ptr(st.kappa), ptr(st.zeta), ptr(st.eta), ptr(st.kappa_curv),
            ptr(st.f_u), ptr(st.f_v), dt, step_idx, C.byref(opts), ptr(psu), ptr(psv), C.byref(it),
            C.byref(res)), "advance_step")
        return it.value, res.value, rc

    # -- device-resident API ----------------------------------------------------------
    def upload(self, u, v, u_prev, v_prev, p, c, phi_s=None):
        psu, psv = phi_s if phi_s is not None else (None, None)
        check(load().fugu_state_upload(self.h, ptr(u), ptr(v), ptr(u_prev), ptr(v_prev), ptr(p), ptr(c),
                                       ptr(psu), ptr(psv)), "upload")

    def download(self, u, v, u_prev, v_prev, p, c):
        check(load().fugu_state_download(self.h, ptr(u), ptr(v), ptr(u_prev), ptr(v_prev), ptr(p), ptr(c)),
              "download")

    def field(self, name, out):
        check(load().fugu_field_download(self.h, name.encode(), ptr(out)), "field_download")
        return out

    def checkpoint_write(self, prefix, step, t):
        check(load().fugu_checkpoint_write(self.h, prefix.encode(), int(step), float(t)), "checkpoint_write")

    def checkpoint_read(self, prefix):
        """(ok, step, t)."""
        st, t, ok = C.c_int(0), C.c_double(0), C.c_int(0)
        check(load().fugu_checkpoint_read(self.h, prefix.encode(), C.byref(st), C.byref(t), C.byref(ok)),
              "checkpoint_read")
        return bool(ok.value), st.value, t.value

    def write_vtk(self, filename):
        check(load().fugu_write_vtk(self.h, filename.encode()), "write_vtk")

    def init_properties(self):
        check(load().fugu_state_init_properties(self.h), "init_properties")

    def kernel_timing(self, enable=True):
        check(load().fugu_kernel_timing(self.h, int(enable)), "kernel_timing")

    def kernel_time(self, name):
        """(total_ms, launches, algorithmic bytes) of an instrumented kernel since timing was enabled."""
        t, n, b = C.c_double(0), C.c_longlong(0), C.c_double(0)
        check(load().fugu_kernel_time(self.h, name.encode(), C.byref(t), C.byref(n), C.byref(b)), "kernel_time")
        return t.value, n.value, b.value

    def step(self, dt, step_idx, opts: FuguStepOpts):
        st = FuguStepStats()
        rc = check(load().fugu_step(self.h, dt, step_idx, C.byref(opts), C.byref(st)), "step")
        return st, rc


def step_opts(use_bicgstab2=True, u_lid=None, tol=1e-6, max_iter=500, use_amp=False, u_solid=0.0, v_solid=0.0,
              solver_flags=0, inner_tol=0.0, inner_max_iter=0, mg_cycles=0, restart=0):
    return FuguStepOpts(int(use_bicgstab2), int(u_lid is not None), float(u_lid or 0.0), tol, max_iter,
                        int(use_amp), u_solid, v_solid, solver_flags, inner_tol, inner_max_iter, mg_cycles,
                        restart)
