"""ctypes front end of the CPU oracle (oracle/libfugu2d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker.  Nothing in the product package
imports this module.

Fields are numpy float64 arrays holding the reference's Fortran padded arrays
in C order with the i index fastest, i.e. shape (rows, ld):
    u  (0-nh:nx+nh, 1-nh:ny+nh)  -> (ny+2nh,   nx+1+2nh)
    v  (1-nh:nx+nh, 0-nh:ny+nh)  -> (ny+1+2nh, nx+2nh)
    p  (1-nh:nx+nh, 1-nh:ny+nh)  -> (ny+2nh,   nx+2nh)
    w  (0-nh:nx+nh, 0-nh:ny+nh)  -> (ny+1+2nh, nx+1+2nh)   (vertices)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfugu2d_oracle.so")


class OgGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nh", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double),
                ("per_x", C.c_int), ("per_y", C.c_int)]


class OgPhys(C.Structure):
    _fields_ = [("rho1", C.c_double), ("rho2", C.c_double), ("mu1", C.c_double),
                ("mu2", C.c_double), ("sigma", C.c_double), ("gx", C.c_double),
                ("gy", C.c_double), ("time_order", C.c_int),
                ("alpha_penalty", C.c_double), ("cfl_target", C.c_double)]


class OgStepOpts(C.Structure):
    _fields_ = [("use_bicgstab2", C.c_int), ("has_lid", C.c_int), ("u_lid", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("use_amp", C.c_int),
                ("u_solid", C.c_double), ("v_solid", C.c_double), ("consistent", C.c_int)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc), return the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_G = C.POINTER(OgGrid)


def _declare(L):
    def s(name, args, res=None):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    s("og_set_sum_mode", [C.c_int]); s("og_get_sum_mode", [], C.c_int)
    s("og_set_pcd_pcg", [C.c_double, C.c_int], None)
    s("og_set_vof_limiter", [C.c_int], None)
    s("og_superbee", [C.c_double], C.c_double)
    s("og_halo_p", [_G, _D]); s("og_halo_u", [_G, _D]); s("og_halo_v", [_G, _D])
    s("og_divergence", [_G, _D, _D, _D])
    s("og_grad_pressure", [_G, _D, _D, _D])
    s("og_viscous_stress", [_G] + [_D] * 7)
    s("og_advection", [_G] + [_D] * 11)
    s("og_velocity_bcs", [_G, _D, _D, C.c_double])
    s("og_matvec_coupled", [_G] + [_D] * 14 + [C.c_double, C.c_int] + [_D] * 4)
    s("og_coupled_dot", [_G] + [_D] * 6, C.c_double)
    s("og_poisson_pcg", [_G, _D, _D, _D, _D, C.c_double, C.c_int, _I, _D])
    s("og_helmholtz_u", [_G] + [_D] * 6 + [C.c_double, C.c_int, _D])
    s("og_helmholtz_v", [_G] + [_D] * 6 + [C.c_double, C.c_int, _D])
    s("og_pcd", [_G] + [_D] * 13 + [C.c_double, _D, _D])
    krylov = [_G] + [_D] * 14 + [C.c_double, C.c_double, C.c_int, _I, _D, C.c_int, C.c_double] + [_D] * 4
    s("og_bicgstab", krylov)
    s("og_bicgstab2", krylov)
    s("og_bicgstab2_ex", krylov + [C.c_int])
    s("og_van_leer", [C.c_double], C.c_double)
    s("og_vof_advection", [_G, _D, _D, _D, _D, C.c_double, _D, _D])
    s("og_svof", [_G, _D, _D, C.c_double, C.c_int, C.c_double, _I])
    s("og_curvature", [_G, _D, _D])
    s("og_csf", [_G, _D, _D, C.c_double, _D, _D])
    s("og_harmonic_viscosity", [C.c_double] * 3, C.c_double)
    s("og_mixture_densities", [_G, _D, C.c_double, C.c_double, _D, _D, _D])
    s("og_mixture_viscosities", [_G, _D, C.c_double, C.c_double, _D, _D, _D])
    s("og_adaptive_dt", [_G, C.POINTER(OgPhys), _D, _D, C.c_double], C.c_double)
    s("og_advance_step", [_G, C.POINTER(OgPhys), C.POINTER(OgStepOpts)] + [_D] * 15 +
      [C.c_double, C.c_int, _D, _D, _I, _D])
    # adjoint building blocks
    s("og_mask_owned", [_G, _D, _D, _D])
    for k in ("p", "u", "v"):
        s(f"og_halo_{k}_T", [_G, _D])
    s("og_velocity_bcs_T", [_G, _D, _D])
    s("og_extend", [_G, _D, _D, _D]); s("og_extend_T", [_G, _D, _D, _D])
    s("og_matvec_eff", [_G] + [_D] * 14 + [C.c_double, C.c_int])
    s("og_matvec_T", [_G] + [_D] * 14 + [C.c_double, C.c_int])
    s("og_pcd_lin", [_G] + [_D] * 13 + [C.c_double, C.c_double, C.c_int, C.c_double])
    s("og_pcd_lin_T", [_G] + [_D] * 13 + [C.c_double, C.c_double, C.c_int, C.c_double])
    s("og_operator_coef_grad", [_G] + [_D] * 12 + [C.c_double, C.c_int, C.c_double] + [_D] * 8)
    s("og_cell_properties", [_G] + [_D] * 8)
    s("og_cell_properties_T", [_G] + [_D] * 9)
    s("og_get_vof_limiter", [], C.c_int)
    s("og_vof_advection_T", [_G, _D, _D, _D, C.c_double, _D, _D, _D, _D])
    s("og_svof_T", [_G, _D, C.c_double, C.c_int, _D])
    s("og_curvature_T", [_G, _D, _D, _D])
    s("og_csf_T", [_G, _D, _D, C.c_double, _D, _D, _D, _D])


def ptr(a):
    """Pointer to a C-contiguous float64 array, or NULL for None."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class Grid:
    """Grid geometry and padded-array shapes (single device domain)."""

    def __init__(self, nx, ny, lx=1.0, ly=1.0, nh=2, per_x=False, per_y=False):
        self.nx, self.ny, self.nh = int(nx), int(ny), int(nh)
        self.lx, self.ly = float(lx), float(ly)
        self.dx = self.lx / self.nx          # src/mod_mpi.f90:104-105
        self.dy = self.ly / self.ny
        self.per_x, self.per_y = bool(per_x), bool(per_y)

    @property
    def c(self):
        return OgGrid(self.nx, self.ny, self.nh, self.dx, self.dy, int(self.per_x), int(self.per_y))

    def shape(self, kind):
        nx, ny, nh = self.nx, self.ny, self.nh
        return {"u": (ny + 2 * nh, nx + 1 + 2 * nh), "v": (ny + 1 + 2 * nh, nx + 2 * nh),
                "p": (ny + 2 * nh, nx + 2 * nh), "w": (ny + 1 + 2 * nh, nx + 1 + 2 * nh)}[kind]

    def zeros(self, kind):
        return np.zeros(self.shape(kind), dtype=np.float64)

    # index of Fortran (i, j) in the numpy array of each kind
    def iu(self, i, j): return (j - 1 + self.nh, i + self.nh)
    def iv(self, i, j): return (j + self.nh, i - 1 + self.nh)
    def ip(self, i, j): return (j - 1 + self.nh, i - 1 + self.nh)
    def iw(self, i, j): return (j + self.nh, i + self.nh)

    # Fortran index ranges of each kind (full padded ranges)
    def irange(self, kind):
        nx, nh = self.nx, self.nh
        return range(-nh, nx + nh + 1) if kind in ("u", "w") else range(1 - nh, nx + nh + 1)

    def jrange(self, kind):
        ny, nh = self.ny, self.nh
        return range(-nh, ny + nh + 1) if kind in ("v", "w") else range(1 - nh, ny + nh + 1)

    def coords(self, kind):
        """Physical coordinates (X, Y) of every padded point of a field kind."""
        i = np.array(list(self.irange(kind)), dtype=np.float64)
        j = np.array(list(self.jrange(kind)), dtype=np.float64)
        if kind == "u":
            x, y = i * self.dx, (j - 0.5) * self.dy
        elif kind == "v":
            x, y = (i - 0.5) * self.dx, j * self.dy
        elif kind == "p":
            x, y = (i - 0.5) * self.dx, (j - 0.5) * self.dy
        else:
            x, y = i * self.dx, j * self.dy
        return np.meshgrid(x, y)


# --------------------------------------------------------------------------
# Thin Python wrappers (allocate outputs, return them)

def set_sum_mode(mode: int):
    """0 (default): the reference's plain sequential sums; 1: double-double accumulation (the
    restatement of the GPU library's fast tree order, FUGU_SUM_TREE)."""
    lib().og_set_sum_mode(int(mode))


def set_pcd_pcg(tol=1e-5, max_iter=150):
    """The PCD's inner PCG stopping rule (default: the reference's 1e-5 / 150)."""
    lib().og_set_pcd_pcg(float(tol), int(max_iter))


def set_vof_limiter(lim: int):
    """0: van Leer (the reference's, default); 1: Superbee."""
    lib().og_set_vof_limiter(int(lim))


def get_sum_mode() -> int:
    return int(lib().og_get_sum_mode())


def halo(g: Grid, a, kind):
    getattr(lib(), "og_halo_" + kind)(C.byref(g.c), ptr(a))


def velocity_bcs(g, u, v, u_lid=0.0):
    lib().og_velocity_bcs(C.byref(g.c), ptr(u), ptr(v), u_lid)


def mixture(g, c, rho1, rho2, mu1, mu2):
    rc, ru, rv = g.zeros("p"), g.zeros("u"), g.zeros("v")
    ka, ze, et = g.zeros("p"), g.zeros("w"), g.zeros("w")
    lib().og_mixture_densities(C.byref(g.c), ptr(c), rho1, rho2, ptr(rc), ptr(ru), ptr(rv))
    lib().og_mixture_viscosities(C.byref(g.c), ptr(c), mu1, mu2, ptr(ka), ptr(ze), ptr(et))
    return dict(rho_c=rc, rho_u=ru, rho_v=rv, kappa=ka, zeta=ze, eta=et)


def matvec(g, X, props, u_adv, v_adv, dt, include_advection=True, pen=None, mdot=None):
    """Returns (Y, X_after) -- X is modified in place by the periodic halo exchange."""
    Xu, Xv, Xp = X
    Y = (g.zeros("u"), g.zeros("v"), g.zeros("p"))
    pu, pv = pen if pen is not None else (None, None)
    mu, mv = mdot if mdot is not None else (None, None)
    lib().og_matvec_coupled(C.byref(g.c), ptr(Xu), ptr(Xv), ptr(Xp), ptr(Y[0]), ptr(Y[1]), ptr(Y[2]),
                            ptr(u_adv), ptr(v_adv), ptr(props["rho_c"]), ptr(props["rho_u"]),
                            ptr(props["rho_v"]), ptr(props["kappa"]), ptr(props["zeta"]),
                            ptr(props["eta"]), dt, int(include_advection), ptr(pu), ptr(pv),
                            ptr(mu), ptr(mv))
    return Y


def coupled_dot(g, a, b):
    return lib().og_coupled_dot(C.byref(g.c), ptr(a[0]), ptr(a[1]), ptr(a[2]),
                                ptr(b[0]), ptr(b[1]), ptr(b[2]))


def poisson_pcg(g, phi, rhs, rho_u, rho_v, tol=1e-5, max_iter=150):
    it = C.c_int(0)
    res = C.c_double(0.0)
    lib().og_poisson_pcg(C.byref(g.c), ptr(phi), ptr(rhs), ptr(rho_u), ptr(rho_v), tol, max_iter,
                         C.byref(it), C.byref(res))
    return it.value, res.value


def helmholtz_u(g, u_sol, rhs, props, dt, n_sweeps=5, pen_u=None):
    lib().og_helmholtz_u(C.byref(g.c), ptr(u_sol), ptr(rhs), ptr(props["rho_u"]), ptr(props["kappa"]),
                         ptr(props["zeta"]), ptr(props["eta"]), dt, n_sweeps, ptr(pen_u))


def helmholtz_v(g, v_sol, rhs, props, dt, n_sweeps=5, pen_v=None):
    lib().og_helmholtz_v(C.byref(g.c), ptr(v_sol), ptr(rhs), ptr(props["rho_v"]), ptr(props["kappa"]),
                         ptr(props["zeta"]), ptr(props["eta"]), dt, n_sweeps, ptr(pen_v))


def pcd(g, Z, props, u_adv, v_adv, dt, pen=None):
    R = (g.zeros("u"), g.zeros("v"), g.zeros("p"))
    pu, pv = pen if pen is not None else (None, None)
    lib().og_pcd(C.byref(g.c), ptr(Z[0]), ptr(Z[1]), ptr(Z[2]), ptr(R[0]), ptr(R[1]), ptr(R[2]),
                 ptr(u_adv), ptr(v_adv), ptr(props["rho_u"]), ptr(props["rho_v"]),
                 ptr(props["kappa"]), ptr(props["zeta"]), ptr(props["eta"]), dt, ptr(pu), ptr(pv))
    return R


def krylov(g, X, B, props, u_adv, v_adv, dt, tol=1e-5, max_iter=100, order=2,
           u_lid=None, pen=None, mdot=None, consistent=False):
    """In-place solve of X; returns (iters, rel_res).  order=1: BiCGStab, 2: BiCGStab(2);
    consistent=True adds the velocity BCs on r and s that the reference's BiCGStab(2) omits."""
    it = C.c_int(0)
    res = C.c_double(0.0)
    pu, pv = pen if pen is not None else (None, None)
    mu, mv = mdot if mdot is not None else (None, None)
    args = [C.byref(g.c), ptr(X[0]), ptr(X[1]), ptr(X[2]), ptr(B[0]), ptr(B[1]), ptr(B[2]),
            ptr(u_adv), ptr(v_adv), ptr(props["rho_c"]), ptr(props["rho_u"]), ptr(props["rho_v"]),
            ptr(props["kappa"]), ptr(props["zeta"]), ptr(props["eta"]), dt, tol, max_iter,
            C.byref(it), C.byref(res), int(u_lid is not None), float(u_lid or 0.0),
            ptr(pu), ptr(pv), ptr(mu), ptr(mv)]
    if order == 2:
        lib().og_bicgstab2_ex(*args, int(consistent))
    else:
        lib().og_bicgstab(*args)
    return it.value, res.value


def vof_advection(g, c, u, v, dt, want_flux=False):
    out = g.zeros("p")
    fx = g.zeros("u") if want_flux else None
    fy = g.zeros("v") if want_flux else None
    lib().og_vof_advection(C.byref(g.c), ptr(c), ptr(out), ptr(u), ptr(v), dt, ptr(fx), ptr(fy))
    return (out, fx, fy) if want_flux else out


def svof(g, c_raw, delta, max_iter=100, tol=1e-6):
    cs = g.zeros("p")
    it = C.c_int(0)
    lib().og_svof(C.byref(g.c), ptr(c_raw), ptr(cs), delta, max_iter, tol, C.byref(it))
    return cs, it.value


def curvature(g, cs):
    kc = g.zeros("p")
    lib().og_curvature(C.byref(g.c), ptr(cs), ptr(kc))
    return kc


def csf(g, c_raw, kc, sigma):
    fu, fv = g.zeros("u"), g.zeros("v")
    lib().og_csf(C.byref(g.c), ptr(c_raw), ptr(kc), sigma, ptr(fu), ptr(fv))
    return fu, fv


# Transposes of the interface chain (reverse mode at the given state; fresh outputs).
def vof_advection_T(g, c_in, u, v, dt, cbar_out):
    cb, ub, vb = g.zeros("p"), g.zeros("u"), g.zeros("v")
    lib().og_vof_advection_T(C.byref(g.c), ptr(c_in), ptr(u), ptr(v), dt, ptr(cbar_out), ptr(cb), ptr(ub), ptr(vb))
    return cb, ub, vb


def svof_T(g, csbar, delta, iters):
    cb = g.zeros("p")
    lib().og_svof_T(C.byref(g.c), ptr(csbar), delta, int(iters), ptr(cb))
    return cb


def curvature_T(g, cs, kbar):
    sb = g.zeros("p")
    lib().og_curvature_T(C.byref(g.c), ptr(cs), ptr(kbar), ptr(sb))
    return sb


def csf_T(g, c_raw, kc, sigma, fubar, fvbar):
    cb, kb = g.zeros("p"), g.zeros("p")
    lib().og_csf_T(C.byref(g.c), ptr(c_raw), ptr(kc), sigma, ptr(fubar), ptr(fvbar), ptr(cb), ptr(kb))
    return cb, kb


class State:
    """The padded state arrays of advance_coupled_step (src/mod_solver_coupled.f90:137-160)."""

    NAMES_U = ("u", "u_prev", "rho_u", "f_u")
    NAMES_V = ("v", "v_prev", "rho_v", "f_v")
    NAMES_P = ("p", "c", "rho_c", "kappa", "kappa_curv")
    NAMES_W = ("zeta", "eta")

    def __init__(self, g: Grid):
        self.g = g
        for n in self.NAMES_U:
            setattr(self, n, g.zeros("u"))
        for n in self.NAMES_V:
            setattr(self, n, g.zeros("v"))
        for n in self.NAMES_P:
            setattr(self, n, g.zeros("p"))
        for n in self.NAMES_W:
            setattr(self, n, g.zeros("w"))

    def copy(self):
        s = State(self.g)
        for n in self.NAMES_U + self.NAMES_V + self.NAMES_P + self.NAMES_W:
            setattr(s, n, getattr(self, n).copy())
        return s

    def init_properties(self, phys: OgPhys):
        pr = mixture(self.g, self.c, phys.rho1, phys.rho2, phys.mu1, phys.mu2)
        for k, a in pr.items():
            getattr(self, k)[...] = a

    def props(self):
        return dict(rho_c=self.rho_c, rho_u=self.rho_u, rho_v=self.rho_v,
                    kappa=self.kappa, zeta=self.zeta, eta=self.eta)


def advance_step(st: State, phys: OgPhys, dt, step_idx, use_bicgstab2=True, u_lid=None,
                 tol=1e-6, max_iter=500, use_amp=False, phi_s=None, u_solid=0.0, v_solid=0.0, consistent=False):
    o = OgStepOpts(int(use_bicgstab2), int(u_lid is not None), float(u_lid or 0.0), tol, max_iter,
                   int(use_amp), u_solid, v_solid, int(consistent))
    it = C.c_int(0)
    res = C.c_double(0.0)
    psu, psv = phi_s if phi_s is not None else (None, None)
    lib().og_advance_step(C.byref(st.g.c), C.byref(phys), C.byref(o),
                          ptr(st.u), ptr(st.v), ptr(st.p), ptr(st.c), ptr(st.u_prev), ptr(st.v_prev),
                          ptr(st.rho_c), ptr(st.rho_u), ptr(st.rho_v), ptr(st.kappa), ptr(st.zeta),
                          ptr(st.eta), ptr(st.kappa_curv), ptr(st.f_u), ptr(st.f_v), dt, step_idx,
                          ptr(psu), ptr(psv), C.byref(it), C.byref(res))
    return it.value, res.value


def adaptive_dt(g, phys, u, v, dt_max):
    return lib().og_adaptive_dt(C.byref(g.c), C.byref(phys), ptr(u), ptr(v), dt_max)


# ---- adjoint building blocks (fugu2d_adjoint.c) -----------------------------------------
PROP_NAMES = ("rho_c", "rho_u", "rho_v", "kappa", "zeta", "eta")


def mask_owned(g, X):
    """Zero the entries outside the owned degrees of freedom, in place (triplet list)."""
    lib().og_mask_owned(C.byref(g.c), ptr(X[0]), ptr(X[1]), ptr(X[2]))
    return X


def halo_T(g, a, kind):
    getattr(lib(), f"og_halo_{kind}_T")(C.byref(g.c), ptr(a))


def velocity_bcs_T(g, u, v):
    lib().og_velocity_bcs_T(C.byref(g.c), ptr(u), ptr(v))


def extend(g, X):
    """E: owned triplet -> the padded operand the matvec sees (copy)."""
    Y = [a.copy() for a in X]
    lib().og_extend(C.byref(g.c), ptr(Y[0]), ptr(Y[1]), ptr(Y[2]))
    return Y


def extend_T(g, X):
    Y = [a.copy() for a in X]
    lib().og_extend_T(C.byref(g.c), ptr(Y[0]), ptr(Y[1]), ptr(Y[2]))
    return Y


def _props_args(pr):
    return [ptr(pr[k]) for k in PROP_NAMES]


def matvec_eff(g, X, props, u_adv, v_adv, dt, include_advection=True):
    Y = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    lib().og_matvec_eff(C.byref(g.c), *[ptr(a) for a in X], *[ptr(a) for a in Y], ptr(u_adv), ptr(v_adv),
                        *_props_args(props), dt, int(include_advection))
    return Y


def matvec_T(g, L, props, u_adv, v_adv, dt, include_advection=True):
    Y = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    lib().og_matvec_T(C.byref(g.c), *[ptr(a) for a in L], *[ptr(a) for a in Y], ptr(u_adv), ptr(v_adv),
                      *_props_args(props), dt, int(include_advection))
    return Y


def _pcd_props(pr):
    return [ptr(pr[k]) for k in ("rho_u", "rho_v", "kappa", "zeta", "eta")]


def pcd_lin(g, Z, props, u_adv, v_adv, dt, ptol=1e-14, pmax=5000, visc=0.0):
    R = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    lib().og_pcd_lin(C.byref(g.c), *[ptr(a) for a in Z], *[ptr(a) for a in R], ptr(u_adv), ptr(v_adv),
                     *_pcd_props(props), dt, ptol, pmax, visc)
    return R


def pcd_lin_T(g, Y, props, u_adv, v_adv, dt, ptol=1e-14, pmax=5000, visc=0.0):
    Z = [g.zeros("u"), g.zeros("v"), g.zeros("p")]
    lib().og_pcd_lin_T(C.byref(g.c), *[ptr(a) for a in Y], *[ptr(a) for a in Z], ptr(u_adv), ptr(v_adv),
                       *_pcd_props(props), dt, ptol, pmax, visc)
    return Z


def operator_coef_grad(g, L, Xext, props, u_adv, v_adv, dt, include_advection=True, s=1.0):
    """Gradients of s * lam^T A(coef; X) with respect to rho_c, rho_u, rho_v, kappa, zeta, eta
    and the advecting velocity (u_adv, v_adv); Xext is the padded operand (extend(X))."""
    out = dict(rho_c=g.zeros("p"), rho_u=g.zeros("u"), rho_v=g.zeros("v"), kappa=g.zeros("p"),
               zeta=g.zeros("w"), eta=g.zeros("w"), u_adv=g.zeros("u"), v_adv=g.zeros("v"))
    lib().og_operator_coef_grad(C.byref(g.c), ptr(L[0]), ptr(L[1]), ptr(Xext[0]), ptr(Xext[1]), ptr(u_adv),
                                ptr(v_adv), *_props_args(props), dt, int(include_advection), s,
                                *[ptr(out[k]) for k in ("rho_c", "rho_u", "rho_v", "kappa", "zeta", "eta",
                                                         "u_adv", "v_adv")])
    return out


def cell_properties(g, mu_cell, rho_cell):
    pr = dict(rho_c=g.zeros("p"), rho_u=g.zeros("u"), rho_v=g.zeros("v"), kappa=g.zeros("p"),
              zeta=g.zeros("w"), eta=g.zeros("w"))
    lib().og_cell_properties(C.byref(g.c), ptr(mu_cell), ptr(rho_cell), *_props_args(pr))
    return pr


def cell_properties_T(g, mu_cell, gcoef):
    """Transpose of the Jacobian of cell_properties at mu_cell: coefficient gradients ->
    (g_mu, g_rho) per padded cell."""
    gc = {k: gcoef[k].copy() for k in PROP_NAMES}
    g_mu, g_rho = g.zeros("p"), g.zeros("p")
    lib().og_cell_properties_T(C.byref(g.c), ptr(mu_cell), *_props_args(gc), ptr(g_mu), ptr(g_rho))
    return g_mu, g_rho
