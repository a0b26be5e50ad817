/*
 * fugu2d_oracle.c -- CPU restatement of the FUGU-2D per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * cites the reference file:lines it restates.  Loops, operation order and
 * branch structure follow the reference so that, compiled with
 * -ffp-contract=off, element-wise results are bit-identical to the reference
 * arithmetic.  Reductions are sequential (j outer, i inner) as in the
 * reference at one MPI rank by default (og_set_sum_mode(0)); og_set_sum_mode(1)
 * switches the solver sums to a double-double accumulation, the restatement of the
 * GPU library's fast tree order.
 *
 * Pinning (how this file is checked against something other than itself) is
 * listed per routine in tests/test_oracle_*.py.
 */
#include "fugu2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DECL                                                          \
  const int nx = g->nx, ny = g->ny, nh = g->nh;                       \
  const size_t ldu = (size_t)(nx + 1 + 2 * nh), ldv = (size_t)(nx + 2 * nh); \
  const size_t ldp = (size_t)(nx + 2 * nh), ldw = (size_t)(nx + 1 + 2 * nh); \
  (void)ldu; (void)ldv; (void)ldp; (void)ldw; (void)nx; (void)ny; (void)nh

#define U(a, i, j) (a)[(size_t)((i) + nh) + (size_t)((j) - 1 + nh) * ldu]
#define V(a, i, j) (a)[(size_t)((i) - 1 + nh) + (size_t)((j) + nh) * ldv]
#define P(a, i, j) (a)[(size_t)((i) - 1 + nh) + (size_t)((j) - 1 + nh) * ldp]
#define W(a, i, j) (a)[(size_t)((i) + nh) + (size_t)((j) + nh) * ldw]

size_t og_size_u(const og_grid* g) { return (size_t)(g->nx + 1 + 2 * g->nh) * (size_t)(g->ny + 2 * g->nh); }
size_t og_size_v(const og_grid* g) { return (size_t)(g->nx + 2 * g->nh) * (size_t)(g->ny + 1 + 2 * g->nh); }
size_t og_size_p(const og_grid* g) { return (size_t)(g->nx + 2 * g->nh) * (size_t)(g->ny + 2 * g->nh); }
size_t og_size_w(const og_grid* g) { return (size_t)(g->nx + 1 + 2 * g->nh) * (size_t)(g->ny + 1 + 2 * g->nh); }

/* Summation of the solver's global reductions (coupled_dot, the PCG's means and dots).
 * Mode 0 (default): plain sequential sums `s = s + term` from zero in the reference's loop
 * order (j outer, i inner; u, v, p), i.e. the reference's own roundings.  Mode 1: a
 * double-double accumulation (TwoSum, then renormalisation), accurate to about
 * n u^2 sum|x|, which nearly always rounds to the same double as the GPU's double-double
 * tree of the same terms (not always: it is not a correctly rounded sum). */
static int g_sum_mode = 0;
void og_set_sum_mode(int mode) { g_sum_mode = mode; }
int og_get_sum_mode(void) { return g_sum_mode; }
typedef struct { double hi, lo; } acc_t;
static void acc_add(acc_t* a, double x) {
  if (!g_sum_mode) { a->hi = a->hi + x; return; }
  const double s = a->hi + x;
  const double v = s - a->hi;
  double e = (a->hi - (s - v)) + (x - v);
  e = e + (a->lo + 0.0);
  a->hi = s + e;
  a->lo = e - (a->hi - s);
}
static double acc_val(acc_t a) { return g_sum_mode ? a.hi + a.lo : a.hi; }

static double* zalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }
static double fmin_f(double a, double b) { return (b < a) ? b : a; } /* Fortran min(a,b) */
static double fmax_f(double a, double b) { return (b > a) ? b : a; } /* Fortran max(a,b) */
static double clamp01(double x) { return fmax_f(0.0, fmin_f(1.0, x)); }

/* ------------------------------------------------------------------------ */
/* Halo exchange at one rank: src/mod_mpi.f90:113-153 (scalar), 159-199 (u),
 * 205-245 (v).  Neighbour = self when periodic, MPI_PROC_NULL otherwise.
 * x pass over the full j range, then y pass over the full i range.          */
void og_halo_p(const og_grid* g, double* a) {
  DECL;
  if (g->per_x) {
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) P(a, 1 - nh + k, j) = P(a, nx - nh + 1 + k, j);
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) P(a, nx + 1 + k, j) = P(a, 1 + k, j);
  }
  if (g->per_y) {
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) P(a, i, 1 - nh + k) = P(a, i, ny - nh + 1 + k);
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) P(a, i, ny + 1 + k) = P(a, i, 1 + k);
  }
}

void og_halo_u(const og_grid* g, double* a) {
  DECL;
  if (g->per_x) {
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) U(a, 1 - nh + k, j) = U(a, nx - nh + 1 + k, j);
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) U(a, nx + 1 + k, j) = U(a, 1 + k, j);
  }
  if (g->per_y) {
    for (int k = 0; k < nh; ++k)
      for (int i = -nh; i <= nx + nh; ++i) U(a, i, 1 - nh + k) = U(a, i, ny - nh + 1 + k);
    for (int k = 0; k < nh; ++k)
      for (int i = -nh; i <= nx + nh; ++i) U(a, i, ny + 1 + k) = U(a, i, 1 + k);
  }
}

void og_halo_v(const og_grid* g, double* a) {
  DECL;
  if (g->per_x) {
    for (int j = -nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) V(a, 1 - nh + k, j) = V(a, nx - nh + 1 + k, j);
    for (int j = -nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) V(a, nx + 1 + k, j) = V(a, 1 + k, j);
  }
  if (g->per_y) {
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) V(a, i, 1 - nh + k) = V(a, i, ny - nh + 1 + k);
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) V(a, i, ny + 1 + k) = V(a, i, 1 + k);
  }
}

/* ------------------------------------------------------------------------ */
/* src/mod_operators.f90:34-53 */
void og_divergence(const og_grid* g, const double* u, const double* v, double* div) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  memset(div, 0, og_size_p(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      P(div, i, j) = (U(u, i, j) - U(u, i - 1, j)) * inv_dx + (V(v, i, j) - V(v, i, j - 1)) * inv_dy;
}

/* src/mod_operators.f90:60-88 */
void og_grad_pressure(const og_grid* g, const double* p, double* gu, double* gv) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  memset(gu, 0, og_size_u(g) * sizeof(double));
  memset(gv, 0, og_size_v(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) U(gu, i, j) = (P(p, i + 1, j) - P(p, i, j)) * inv_dx;
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) V(gv, i, j) = (P(p, i, j + 1) - P(p, i, j)) * inv_dy;
}

/* src/mod_operators.f90:99-162 */
void og_viscous_stress(const og_grid* g, const double* u, const double* v,
                       const double* kappa, const double* zeta, const double* eta,
                       double* visc_u, double* visc_v) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  /* T_xx(0:nx+1,1:ny), T_yy(1:nx,0:ny+1), T_xy/T_yx(0:nx,0:ny) */
  const int lxx = nx + 2, lyy = nx, lxy = nx + 1;
  double* Txx = zalloc((size_t)lxx * ny);
  double* Tyy = zalloc((size_t)lyy * (ny + 2));
  double* Txy = zalloc((size_t)lxy * (ny + 1));
  double* Tyx = zalloc((size_t)lxy * (ny + 1));
#define TXX(i, j) Txx[(size_t)(i) + (size_t)((j) - 1) * lxx]
#define TYY(i, j) Tyy[(size_t)((i) - 1) + (size_t)(j) * lyy]
#define TXY(i, j) Txy[(size_t)(i) + (size_t)(j) * lxy]
#define TYX(i, j) Tyx[(size_t)(i) + (size_t)(j) * lxy]
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx + 1; ++i) TXX(i, j) = P(kappa, i, j) * (U(u, i, j) - U(u, i - 1, j)) * inv_dx;
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 1; i <= nx; ++i) TYY(i, j) = P(kappa, i, j) * (V(v, i, j) - V(v, i, j - 1)) * inv_dy;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double du_dy = (U(u, i, j + 1) - U(u, i, j)) * inv_dy;
      const double dv_dx = (V(v, i + 1, j) - V(v, i, j)) * inv_dx;
      TXY(i, j) = (W(zeta, i, j) - W(eta, i, j)) * du_dy + W(eta, i, j) * dv_dx;
      TYX(i, j) = W(eta, i, j) * du_dy + (W(zeta, i, j) - W(eta, i, j)) * dv_dx;
    }
  memset(visc_u, 0, og_size_u(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      U(visc_u, i, j) = -((TXX(i + 1, j) - TXX(i, j)) * inv_dx + (TXY(i, j) - TXY(i, j - 1)) * inv_dy);
  memset(visc_v, 0, og_size_v(g) * sizeof(double));
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      V(visc_v, i, j) = -((TYX(i, j) - TYX(i - 1, j)) * inv_dx + (TYY(i, j + 1) - TYY(i, j)) * inv_dy);
#undef TXX
#undef TYY
#undef TXY
#undef TYX
  free(Txx); free(Tyy); free(Txy); free(Tyx);
}

/* src/mod_operators.f90:168-249 */
void og_advection(const og_grid* g, const double* u_adv, const double* v_adv,
                  const double* rho_c, const double* rho_u, const double* rho_v,
                  const double* u, const double* v, double* adv_u, double* adv_v,
                  const double* mdot_u, const double* mdot_v) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const int use_cons = (mdot_u != NULL) && (mdot_v != NULL);
  const int i_u_max = g->per_x ? nx : nx - 1;
  const int j_v_max = g->per_y ? ny : ny - 1;
  memset(adv_u, 0, og_size_u(g) * sizeof(double));
  memset(adv_v, 0, og_size_v(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= i_u_max; ++i) {
      double mE, mW, mN, mS;
      if (use_cons) {
        mE = half * (U(mdot_u, i, j) + U(mdot_u, i + 1, j));
        mW = half * (U(mdot_u, i - 1, j) + U(mdot_u, i, j));
        mN = half * (V(mdot_v, i, j) + V(mdot_v, i + 1, j));
        mS = half * (V(mdot_v, i, j - 1) + V(mdot_v, i + 1, j - 1));
      } else {
        mE = P(rho_c, i + 1, j) * half * (U(u_adv, i, j) + U(u_adv, i + 1, j));
        mW = P(rho_c, i, j) * half * (U(u_adv, i - 1, j) + U(u_adv, i, j));
        mN = half * (V(rho_v, i, j) + V(rho_v, i + 1, j)) * half * (V(v_adv, i, j) + V(v_adv, i + 1, j));
        mS = half * (V(rho_v, i, j - 1) + V(rho_v, i + 1, j - 1)) * half *
             (V(v_adv, i, j - 1) + V(v_adv, i + 1, j - 1));
      }
      const double uE = half * (U(u, i, j) + U(u, i + 1, j));
      const double uW = half * (U(u, i - 1, j) + U(u, i, j));
      const double uN = half * (U(u, i, j) + U(u, i, j + 1));
      const double uS = half * (U(u, i, j - 1) + U(u, i, j));
      U(adv_u, i, j) = (mE * uE - mW * uW) * inv_dx + (mN * uN - mS * uS) * inv_dy;
    }
  for (int j = 1; j <= j_v_max; ++j)
    for (int i = 1; i <= nx; ++i) {
      double mE, mW, mN, mS;
      if (use_cons) {
        mE = half * (U(mdot_u, i, j) + U(mdot_u, i, j + 1));
        mW = half * (U(mdot_u, i - 1, j) + U(mdot_u, i - 1, j + 1));
        mN = half * (V(mdot_v, i, j) + V(mdot_v, i, j + 1));
        mS = half * (V(mdot_v, i, j - 1) + V(mdot_v, i, j));
      } else {
        mE = half * (U(rho_u, i, j) + U(rho_u, i, j + 1)) * half * (U(u_adv, i, j) + U(u_adv, i, j + 1));
        mW = half * (U(rho_u, i - 1, j) + U(rho_u, i - 1, j + 1)) * half *
             (U(u_adv, i - 1, j) + U(u_adv, i - 1, j + 1));
        mN = P(rho_c, i, j + 1) * half * (V(v_adv, i, j) + V(v_adv, i, j + 1));
        mS = P(rho_c, i, j) * half * (V(v_adv, i, j - 1) + V(v_adv, i, j));
      }
      const double vE = half * (V(v, i, j) + V(v, i + 1, j));
      const double vW = half * (V(v, i - 1, j) + V(v, i, j));
      const double vN = half * (V(v, i, j) + V(v, i, j + 1));
      const double vS = half * (V(v, i, j - 1) + V(v, i, j));
      V(adv_v, i, j) = (mE * vE - mW * vW) * inv_dx + (mN * vN - mS * vS) * inv_dy;
    }
}

/* src/mod_operators.f90:254-333 (x walls first, then y walls) */
void og_velocity_bcs(const og_grid* g, double* u, double* v, double u_lid) {
  DECL;
  if (!g->per_x) {
    for (int j = 1 - nh; j <= ny + nh; ++j) {
      U(u, 0, j) = 0.0;
      for (int i = 1; i <= nh; ++i) U(u, -i, j) = -U(u, i, j);
    }
    for (int j = -nh; j <= ny + nh; ++j)
      for (int i = 1; i <= nh; ++i) V(v, 1 - i, j) = -V(v, i, j);
    for (int j = 1 - nh; j <= ny + nh; ++j) {
      U(u, nx, j) = 0.0;
      for (int i = 1; i <= nh; ++i) U(u, nx + i, j) = -U(u, nx - i, j);
    }
    for (int j = -nh; j <= ny + nh; ++j)
      for (int i = 1; i <= nh; ++i) V(v, nx + i, j) = -V(v, nx - i + 1, j);
  }
  if (!g->per_y) {
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      V(v, i, 0) = 0.0;
      for (int j = 1; j <= nh; ++j) V(v, i, -j) = -V(v, i, j);
    }
    for (int i = -nh; i <= nx + nh; ++i)
      for (int j = 1; j <= nh; ++j) U(u, i, 1 - j) = -U(u, i, j);
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      V(v, i, ny) = 0.0;
      for (int j = 1; j <= nh; ++j) V(v, i, ny + j) = -V(v, i, ny - j);
    }
    for (int i = -nh; i <= nx + nh; ++i)
      for (int j = 1; j <= nh; ++j) U(u, i, ny + j) = 2.0 * u_lid - U(u, i, ny - j + 1);
  }
}

/* src/mod_operators.f90:341-443 */
void og_matvec_coupled(const og_grid* g, double* Xu, double* Xv, double* Xp,
                       double* Yu, double* Yv, double* Yp,
                       const double* u_adv, const double* v_adv,
                       const double* rho_c, const double* rho_u, const double* rho_v,
                       const double* kappa, const double* zeta, const double* eta,
                       double dt, int include_advection,
                       const double* pen_u, const double* pen_v,
                       const double* mdot_u, const double* mdot_v) {
  DECL;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const int j_v_max = g->per_y ? ny : ny - 1;
  const size_t su = og_size_u(g), sv = og_size_v(g);
  double* visc_u = zalloc(su); double* adv_u = zalloc(su); double* gp_u = zalloc(su);
  double* visc_v = zalloc(sv); double* adv_v = zalloc(sv); double* gp_v = zalloc(sv);
  og_halo_u(g, Xu);
  og_halo_v(g, Xv);
  og_halo_p(g, Xp);
  og_viscous_stress(g, Xu, Xv, kappa, zeta, eta, visc_u, visc_v);
  og_grad_pressure(g, Xp, gp_u, gp_v);
  if (include_advection) {
    const int cons = (mdot_u != NULL) && (mdot_v != NULL);
    og_advection(g, u_adv, v_adv, rho_c, rho_u, rho_v, Xu, Xv, adv_u, adv_v,
                 cons ? mdot_u : NULL, cons ? mdot_v : NULL);
  }
  memset(Yu, 0, su * sizeof(double));
  for (int j = 1; j <= ny; ++j) {
    for (int i = 1; i <= i_u_max; ++i) {
      U(Yu, i, j) = (U(rho_u, i, j) / dt) * U(Xu, i, j) + U(adv_u, i, j) + U(visc_u, i, j) + U(gp_u, i, j);
      if (pen_u) U(Yu, i, j) = U(Yu, i, j) + U(pen_u, i, j) * U(Xu, i, j);
    }
    if (!g->per_x) {
      U(Yu, 0, j) = U(Xu, 0, j);
      U(Yu, nx, j) = U(Xu, nx, j);
    }
  }
  memset(Yv, 0, sv * sizeof(double));
  for (int j = 1; j <= j_v_max; ++j)
    for (int i = 1; i <= nx; ++i) {
      V(Yv, i, j) = (V(rho_v, i, j) / dt) * V(Xv, i, j) + V(adv_v, i, j) + V(visc_v, i, j) + V(gp_v, i, j);
      if (pen_v) V(Yv, i, j) = V(Yv, i, j) + V(pen_v, i, j) * V(Xv, i, j);
    }
  if (!g->per_y)
    for (int i = 1; i <= nx; ++i) {
      V(Yv, i, 0) = V(Xv, i, 0);
      V(Yv, i, ny) = V(Xv, i, ny);
    }
  og_divergence(g, Xu, Xv, Yp);
  free(visc_u); free(adv_u); free(gp_u); free(visc_v); free(adv_v); free(gp_v);
}

/* src/mod_bicgstab2.f90:32-72 */
double og_coupled_dot(const og_grid* g, const double* u1, const double* v1, const double* p1,
                      const double* u2, const double* v2, const double* p2) {
  DECL;
  const double vol = g->dx * g->dy;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const int j_v_max = g->per_y ? ny : ny - 1;
  acc_t s = {0.0, 0.0};
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= i_u_max; ++i) acc_add(&s, U(u1, i, j) * U(u2, i, j) * vol);
  for (int j = 1; j <= j_v_max; ++j)
    for (int i = 1; i <= nx; ++i) acc_add(&s, V(v1, i, j) * V(v2, i, j) * vol);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) acc_add(&s, P(p1, i, j) * P(p2, i, j) * vol);
  return acc_val(s);
}

/* ------------------------------------------------------------------------ */
/* src/mod_subsolvers.f90:378-416 */
static void neumann_mirror_p(const og_grid* g, double* phi) {
  DECL;
  if (!g->per_x) {
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int i = 1; i <= nh; ++i) P(phi, 1 - i, j) = P(phi, i, j);
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int i = 1; i <= nh; ++i) P(phi, nx + i, j) = P(phi, nx - i + 1, j);
  }
  if (!g->per_y) {
    for (int i = 1 - nh; i <= nx + nh; ++i)
      for (int j = 1; j <= nh; ++j) P(phi, i, 1 - j) = P(phi, i, j);
    for (int i = 1 - nh; i <= nx + nh; ++i)
      for (int j = 1; j <= nh; ++j) P(phi, i, ny + j) = P(phi, i, ny - j + 1);
  }
}

/* Coefficients of the pressure operator at cell (i,j): mod_subsolvers.f90:65-78 */
static void pcg_coeffs(const og_grid* g, const double* rho_u, const double* rho_v, int i, int j,
                       double inv_dx2, double inv_dy2, double* cE, double* cW, double* cN, double* cS) {
  DECL;
  *cE = inv_dx2 / U(rho_u, i, j);
  *cW = inv_dx2 / U(rho_u, i - 1, j);
  *cN = inv_dy2 / V(rho_v, i, j);
  *cS = inv_dy2 / V(rho_v, i, j - 1);
  if (!g->per_x) {
    if (i == 1) *cW = 0.0;
    if (i == nx) *cE = 0.0;
  }
  if (!g->per_y) {
    if (j == 1) *cS = 0.0;
    if (j == ny) *cN = 0.0;
  }
}

/* src/mod_subsolvers.f90:32-246 */
void og_poisson_pcg(const og_grid* g, double* phi, const double* rhs,
                    const double* rho_u, const double* rho_v,
                    double tol, int max_iter, int* iter_out, double* res_out) {
  DECL;
  const size_t sp = og_size_p(g);
  double* r = zalloc(sp); double* z = zalloc(sp); double* p = zalloc(sp);
  double* Ap = zalloc(sp); double* dinv = zalloc(sp); double* rw = zalloc(sp);
  double res_norm = 0.0, rhs_norm, rz_old, rz_new, pAp, alpha, beta, loc;
  const double inv_dx2 = 1.0 / (g->dx * g->dx), inv_dy2 = 1.0 / (g->dy * g->dy);
  const int walls_both = (!g->per_x) && (!g->per_y);
  int iter;

  for (size_t k = 0; k < sp; ++k) dinv[k] = 1.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      double cE, cW, cN, cS;
      pcg_coeffs(g, rho_u, rho_v, i, j, inv_dx2, inv_dy2, &cE, &cW, &cN, &cS);
      const double d = cE + cW + cN + cS;
      P(dinv, i, j) = (d > 1.0e-14) ? 1.0 / d : 1.0;
    }

  memcpy(rw, rhs, sp * sizeof(double));
  if (walls_both) {
    acc_t s = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) acc_add(&s, P(rhs, i, j));
    const double mean = acc_val(s) / (double)(nx * ny);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) P(rw, i, j) = P(rhs, i, j) - mean;
  }
  {
    acc_t a = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) acc_add(&a, P(rw, i, j) * P(rw, i, j));
    loc = acc_val(a);
  }
  rhs_norm = sqrt(loc);
  if (rhs_norm < 1.0e-14) {
    memset(phi, 0, sp * sizeof(double));
    if (iter_out) *iter_out = 0;
    if (res_out) *res_out = 0.0;
    goto done;
  }

  og_halo_p(g, phi);
  neumann_mirror_p(g, phi);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      double cE, cW, cN, cS;
      pcg_coeffs(g, rho_u, rho_v, i, j, inv_dx2, inv_dy2, &cE, &cW, &cN, &cS);
      P(Ap, i, j) = (cE + cW + cN + cS) * P(phi, i, j) -
                    (cE * P(phi, i + 1, j) + cW * P(phi, i - 1, j) + cN * P(phi, i, j + 1) + cS * P(phi, i, j - 1));
      P(r, i, j) = P(rw, i, j) - P(Ap, i, j);
      P(z, i, j) = P(dinv, i, j) * P(r, i, j);
      P(p, i, j) = P(z, i, j);
    }
  {
    acc_t a = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) acc_add(&a, P(r, i, j) * P(z, i, j));
    rz_old = acc_val(a);
  }

  for (iter = 1; iter <= max_iter; ++iter) {
    og_halo_p(g, p);
    neumann_mirror_p(g, p);
    acc_t a1 = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double cE, cW, cN, cS;
        pcg_coeffs(g, rho_u, rho_v, i, j, inv_dx2, inv_dy2, &cE, &cW, &cN, &cS);
        P(Ap, i, j) = (cE + cW + cN + cS) * P(p, i, j) -
                      (cE * P(p, i + 1, j) + cW * P(p, i - 1, j) + cN * P(p, i, j + 1) + cS * P(p, i, j - 1));
        acc_add(&a1, P(p, i, j) * P(Ap, i, j));
      }
    pAp = acc_val(a1);
    if (fabs(pAp) < 1.0e-30) break;
    alpha = rz_old / pAp;
    acc_t a2 = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        P(phi, i, j) = P(phi, i, j) + alpha * P(p, i, j);
        P(r, i, j) = P(r, i, j) - alpha * P(Ap, i, j);
        P(z, i, j) = P(dinv, i, j) * P(r, i, j);
        acc_add(&a2, P(r, i, j) * P(r, i, j));
      }
    res_norm = sqrt(acc_val(a2));
    if (res_norm / rhs_norm < tol) break;
    acc_t a3 = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) acc_add(&a3, P(r, i, j) * P(z, i, j));
    rz_new = acc_val(a3);
    beta = rz_new / rz_old;
    rz_old = rz_new;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) P(p, i, j) = P(z, i, j) + beta * P(p, i, j);
  }

  if (walls_both) {
    acc_t s = {0.0, 0.0};
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) acc_add(&s, P(phi, i, j));
    const double mean = acc_val(s) / (double)(nx * ny);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) P(phi, i, j) = P(phi, i, j) - mean;
  }
  /* Fortran DO: on normal completion iter = max_iter+1 */
  if (iter_out) *iter_out = (iter < max_iter) ? iter : max_iter;
  if (res_out) *res_out = res_norm / rhs_norm;
done:
  free(r); free(z); free(p); free(Ap); free(dinv); free(rw);
}

/* src/mod_subsolvers.f90:252-309 */
void og_helmholtz_u(const og_grid* g, double* u_sol, const double* rhs_u,
                    const double* rho_u, const double* kappa, const double* zeta,
                    const double* eta, double dt, int n_sweeps, const double* pen_u) {
  DECL;
  const double omega = 0.8;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const double inv_dx2 = 1.0 / (g->dx * g->dx), inv_dy2 = 1.0 / (g->dy * g->dy);
  const size_t su = og_size_u(g);
  double* u_old = zalloc(su);
  for (int sweep = 1; sweep <= n_sweeps; ++sweep) {
    og_halo_u(g, u_sol);
    memcpy(u_old, u_sol, su * sizeof(double));
    for (int j = 1; j <= ny; ++j) {
      for (int i = 1; i <= i_u_max; ++i) {
        const double cE = P(kappa, i + 1, j) * inv_dx2;
        const double cW = P(kappa, i, j) * inv_dx2;
        const double cN = (W(zeta, i, j) - W(eta, i, j)) * inv_dy2;
        const double cS = (W(zeta, i, j - 1) - W(eta, i, j - 1)) * inv_dy2;
        double diag = (U(rho_u, i, j) / dt) + cE + cW + cN + cS;
        if (pen_u) diag = diag + U(pen_u, i, j);
        const double off = cE * U(u_old, i + 1, j) + cW * U(u_old, i - 1, j) +
                           cN * U(u_old, i, j + 1) + cS * U(u_old, i, j - 1);
        if (diag > 1.0e-15)
          U(u_sol, i, j) = (1.0 - omega) * U(u_old, i, j) + omega * ((U(rhs_u, i, j) + off) / diag);
      }
      if (!g->per_x) {
        U(u_sol, 0, j) = U(rhs_u, 0, j);
        U(u_sol, nx, j) = U(rhs_u, nx, j);
      }
    }
  }
  free(u_old);
}

/* src/mod_subsolvers.f90:314-372 */
void og_helmholtz_v(const og_grid* g, double* v_sol, const double* rhs_v,
                    const double* rho_v, const double* kappa, const double* zeta,
                    const double* eta, double dt, int n_sweeps, const double* pen_v) {
  DECL;
  const double omega = 0.8;
  const int j_v_max = g->per_y ? ny : ny - 1;
  const double inv_dx2 = 1.0 / (g->dx * g->dx), inv_dy2 = 1.0 / (g->dy * g->dy);
  const size_t sv = og_size_v(g);
  double* v_old = zalloc(sv);
  for (int sweep = 1; sweep <= n_sweeps; ++sweep) {
    og_halo_v(g, v_sol);
    memcpy(v_old, v_sol, sv * sizeof(double));
    for (int j = 1; j <= j_v_max; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double cE = (W(zeta, i, j) - W(eta, i, j)) * inv_dx2;
        const double cW = (W(zeta, i - 1, j) - W(eta, i - 1, j)) * inv_dx2;
        const double cN = P(kappa, i, j + 1) * inv_dy2;
        const double cS = P(kappa, i, j) * inv_dy2;
        double diag = (V(rho_v, i, j) / dt) + cE + cW + cN + cS;
        if (pen_v) diag = diag + V(pen_v, i, j);
        const double off = cE * V(v_old, i + 1, j) + cW * V(v_old, i - 1, j) +
                           cN * V(v_old, i, j + 1) + cS * V(v_old, i, j - 1);
        if (diag > 1.0e-15)
          V(v_sol, i, j) = (1.0 - omega) * V(v_old, i, j) + omega * ((V(rhs_v, i, j) + off) / diag);
      }
    if (!g->per_y)
      for (int i = 1; i <= nx; ++i) {
        V(v_sol, i, 0) = V(rhs_v, i, 0);
        V(v_sol, i, ny) = V(rhs_v, i, ny);
      }
  }
  free(v_old);
}

/* src/mod_pcd_precond.f90:14-104 */
/* The PCD's inner PCG stopping rule: the reference's 1e-5 / 150 (mod_pcd_precond.f90:50).  A
 * pin may tighten it so that the preconditioner becomes a fixed linear map (to rounding) and the
 * Krylov recurrences around it can be checked against closed forms and dense solves. */
static double g_pcd_pcg_tol = 1.0e-5;
static int g_pcd_pcg_max = 150;
void og_set_pcd_pcg(double tol, int max_iter) { g_pcd_pcg_tol = tol; g_pcd_pcg_max = max_iter; }

void og_pcd(const og_grid* g, const double* zu, const double* zv, const double* zp,
            double* ru, double* rv, double* rp,
            const double* u_adv, const double* v_adv,
            const double* rho_u, const double* rho_v,
            const double* kappa, const double* zeta, const double* eta, double dt,
            const double* pen_u, const double* pen_v) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const size_t su = og_size_u(g), sv = og_size_v(g), sp = og_size_p(g);
  double* phi = zalloc(sp);
  double* zuw = zalloc(su);
  double* zvw = zalloc(sv);
  memset(ru, 0, su * sizeof(double));
  memset(rv, 0, sv * sizeof(double));
  memset(rp, 0, sp * sizeof(double));

  og_poisson_pcg(g, phi, zp, rho_u, rho_v, g_pcd_pcg_tol, g_pcd_pcg_max, NULL, NULL);
  og_halo_p(g, phi);

  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double u_avg = half * (U(u_adv, i, j) + U(u_adv, i - 1, j));
      const double v_avg = half * (V(v_adv, i, j) + V(v_adv, i, j - 1));
      const double conv = u_avg * (P(phi, i + 1, j) - P(phi, i - 1, j)) * (half * inv_dx) +
                          v_avg * (P(phi, i, j + 1) - P(phi, i, j - 1)) * (half * inv_dy);
      P(rp, i, j) = -((1.0 / dt) * P(phi, i, j) + conv);
    }
  og_halo_p(g, rp);

  for (int j = 1; j <= ny; ++j) {
    for (int i = 1; i <= nx - 1; ++i) U(zuw, i, j) = U(zu, i, j) - (P(rp, i + 1, j) - P(rp, i, j)) * inv_dx;
    U(zuw, 0, j) = U(zu, 0, j);
    U(zuw, nx, j) = U(zu, nx, j);
  }
  for (int j = 1; j <= ny - 1; ++j)
    for (int i = 1; i <= nx; ++i) V(zvw, i, j) = V(zv, i, j) - (P(rp, i, j + 1) - P(rp, i, j)) * inv_dy;
  for (int i = 1; i <= nx; ++i) {
    V(zvw, i, 0) = V(zv, i, 0);
    V(zvw, i, ny) = V(zv, i, ny);
  }

  og_helmholtz_v(g, rv, zvw, rho_v, kappa, zeta, eta, dt, 3, pen_v);
  og_halo_v(g, rv);

  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx - 1; ++i) {
      const double cross = -inv_dy * (W(eta, i, j) * (V(rv, i + 1, j) - V(rv, i, j)) * inv_dx -
                                      W(eta, i, j - 1) * (V(rv, i + 1, j - 1) - V(rv, i, j - 1)) * inv_dx);
      U(zuw, i, j) = U(zuw, i, j) - cross;
    }

  og_helmholtz_u(g, ru, zuw, rho_u, kappa, zeta, eta, dt, 3, pen_u);
  og_halo_u(g, ru);
  free(phi); free(zuw); free(zvw);
}

/* ------------------------------------------------------------------------ */
/* Triplet helpers for whole-array Fortran updates (ghosts included). */
typedef struct { double *u, *v, *p; } trip;

static trip trip_alloc(const og_grid* g) {
  trip t = { zalloc(og_size_u(g)), zalloc(og_size_v(g)), zalloc(og_size_p(g)) };
  return t;
}
static void trip_free(trip t) { free(t.u); free(t.v); free(t.p); }

static void mv(const og_grid* g, trip x, trip y, const double* u_adv, const double* v_adv,
               const double* rho_c, const double* rho_u, const double* rho_v,
               const double* kappa, const double* zeta, const double* eta, double dt,
               const double* pen_u, const double* pen_v, const double* mdot_u, const double* mdot_v) {
  og_matvec_coupled(g, x.u, x.v, x.p, y.u, y.v, y.p, u_adv, v_adv, rho_c, rho_u, rho_v, kappa, zeta, eta,
                    dt, 1, pen_u, pen_v, mdot_u, mdot_v);
}
static void pc(const og_grid* g, trip z, trip r, const double* u_adv, const double* v_adv,
               const double* rho_u, const double* rho_v, const double* kappa, const double* zeta,
               const double* eta, double dt, const double* pen_u, const double* pen_v) {
  og_pcd(g, z.u, z.v, z.p, r.u, r.v, r.p, u_adv, v_adv, rho_u, rho_v, kappa, zeta, eta, dt, pen_u, pen_v);
}
static double cdot(const og_grid* g, trip a, trip b) { return og_coupled_dot(g, a.u, a.v, a.p, b.u, b.v, b.p); }

#define FOR_TRIP(g, body)                                                   \
  do {                                                                      \
    size_t n_; double *d_[3]; size_t s_[3] = {og_size_u(g), og_size_v(g), og_size_p(g)}; \
    (void)d_;                                                               \
    for (int c_ = 0; c_ < 3; ++c_) {                                        \
      n_ = s_[c_];                                                          \
      for (size_t k = 0; k < n_; ++k) { body }                              \
    }                                                                       \
  } while (0)
#define C3(t) (c_ == 0 ? (t).u : c_ == 1 ? (t).v : (t).p)

/* src/mod_bicgstab2.f90:77-264 */
void og_bicgstab(const og_grid* g, double* Xu, double* Xv, double* Xp,
                 const double* Bu, const double* Bv, const double* Bp,
                 const double* u_adv, const double* v_adv,
                 const double* rho_c, const double* rho_u, const double* rho_v,
                 const double* kappa, const double* zeta, const double* eta, double dt,
                 double tol, int max_iter, int* iter_out, double* res_out,
                 int has_lid, double u_lid,
                 const double* pen_u, const double* pen_v,
                 const double* mdot_u, const double* mdot_v) {
  trip X = {Xu, Xv, Xp};
  trip B = {(double*)Bu, (double*)Bv, (double*)Bp};
  trip bh = trip_alloc(g), r = trip_alloc(g), r0 = trip_alloc(g), p = trip_alloc(g);
  trip v = trip_alloc(g), s = trip_alloc(g), t = trip_alloc(g), z = trip_alloc(g);
  double norm_b, norm_r, rel_res, rho_old, rho_new, alpha, beta, omega, rv, tt, ts;
  int k;
#define MV(x, y) mv(g, x, y, u_adv, v_adv, rho_c, rho_u, rho_v, kappa, zeta, eta, dt, pen_u, pen_v, mdot_u, mdot_v)
#define PC(zz, rr) pc(g, zz, rr, u_adv, v_adv, rho_u, rho_v, kappa, zeta, eta, dt, pen_u, pen_v)

  og_velocity_bcs(g, X.u, X.v, 0.0);
  PC(B, bh);
  norm_b = sqrt(cdot(g, bh, bh));
  if (norm_b < 1.0e-14) norm_b = 1.0;
  MV(X, z);
  PC(z, v);
  FOR_TRIP(g, { C3(r)[k] = C3(bh)[k] - C3(v)[k]; });
  FOR_TRIP(g, { C3(r0)[k] = C3(r)[k]; C3(p)[k] = C3(r)[k]; });
  rho_old = cdot(g, r0, r);
  norm_r = sqrt(cdot(g, r, r));
  rel_res = norm_r / norm_b;
  alpha = 0.0;
  omega = 1.0;
  (void)alpha;
  for (k = 1; k <= max_iter; ++k) {
    if (rel_res < tol) break;
    og_velocity_bcs(g, p.u, p.v, 0.0);
    MV(p, z);
    PC(z, v);
    og_velocity_bcs(g, v.u, v.v, 0.0);
    rv = cdot(g, r0, v);
    if (fabs(rv) < 1.0e-30) break;
    alpha = rho_old / rv;
    FOR_TRIP(g, { C3(s)[k] = C3(r)[k] - alpha * C3(v)[k]; });
    og_velocity_bcs(g, s.u, s.v, 0.0);
    MV(s, z);
    PC(z, t);
    og_velocity_bcs(g, t.u, t.v, 0.0);
    tt = cdot(g, t, t);
    ts = cdot(g, t, s);
    omega = (tt > 1.0e-30) ? ts / tt : 0.0;
    FOR_TRIP(g, { C3(X)[k] = C3(X)[k] + alpha * C3(p)[k] + omega * C3(s)[k]; });
    FOR_TRIP(g, { C3(r)[k] = C3(s)[k] - omega * C3(t)[k]; });
    norm_r = sqrt(cdot(g, r, r));
    rel_res = norm_r / norm_b;
    if (rel_res < tol) break;
    if (fabs(omega) < 1.0e-30) break;
    rho_new = cdot(g, r0, r);
    if (fabs(rho_old) < 1.0e-30) break;
    beta = (rho_new / rho_old) * (alpha / omega);
    rho_old = rho_new;
    FOR_TRIP(g, { C3(p)[k] = C3(r)[k] + beta * (C3(p)[k] - omega * C3(v)[k]); });
    og_velocity_bcs(g, X.u, X.v, 0.0);
  }
  if (has_lid) og_velocity_bcs(g, X.u, X.v, u_lid);
  if (iter_out) *iter_out = (k < max_iter) ? k : max_iter;
  if (res_out) *res_out = rel_res;
  (void)B;
  trip_free(bh); trip_free(r); trip_free(r0); trip_free(p);
  trip_free(v); trip_free(s); trip_free(t); trip_free(z);
}

/* src/mod_bicgstab2.f90:270-570.  consistent != 0 adds velocity BCs on r and s before
 * their matvecs (absent from the reference at :430 and :489); 0 is the reference. */
void og_bicgstab2_ex(const og_grid* g, double* Xu, double* Xv, double* Xp,
                     const double* Bu, const double* Bv, const double* Bp,
                     const double* u_adv, const double* v_adv,
                     const double* rho_c, const double* rho_u, const double* rho_v,
                     const double* kappa, const double* zeta, const double* eta, double dt,
                     double tol, int max_iter, int* iter_out, double* res_out,
                     int has_lid, double u_lid,
                     const double* pen_u, const double* pen_v,
                     const double* mdot_u, const double* mdot_v, int consistent) {
  trip X = {Xu, Xv, Xp};
  trip B = {(double*)Bu, (double*)Bv, (double*)Bp};
  trip bh = trip_alloc(g), r = trip_alloc(g), rt = trip_alloc(g), uu = trip_alloc(g);
  trip vv = trip_alloc(g), ww = trip_alloc(g), ss = trip_alloc(g), tt = trip_alloc(g);
  trip z = trip_alloc(g), q = trip_alloc(g);
  double norm_b, norm_r, rel_res, xi0, xi1, alpha, beta, gam;
  double om1, om2, lam, nu, tau, det, c1, c2;
  int k;

  og_velocity_bcs(g, X.u, X.v, 0.0);
  PC(B, bh);
  norm_b = sqrt(cdot(g, bh, bh));
  if (norm_b < 1.0e-14) norm_b = 1.0;
  MV(X, z);
  PC(z, q);
  FOR_TRIP(g, { C3(r)[k] = C3(bh)[k] - C3(q)[k]; });
  FOR_TRIP(g, { C3(rt)[k] = C3(r)[k]; });
  /* u_vec = v_vec = 0 (already zero from calloc) */
  xi0 = 1.0; om1 = 0.0; om2 = 1.0; alpha = 0.0;
  norm_r = sqrt(cdot(g, r, r));
  rel_res = norm_r / norm_b;

  for (k = 1; k <= max_iter; ++k) {
    if (rel_res < tol) break;
    if (fabs(om2) > 1.0e-14) xi0 = -om2 * xi0;
    else if (fabs(om1) > 1.0e-14) xi0 = -om1 * xi0;
    else xi0 = -xi0;

    /* even step */
    xi1 = cdot(g, rt, r);
    if (fabs(xi0) < 1.0e-30) break;
    beta = alpha * (xi1 / xi0);
    xi0 = xi1;
    FOR_TRIP(g, { C3(uu)[k] = C3(r)[k] - beta * C3(uu)[k]; });
    og_velocity_bcs(g, uu.u, uu.v, 0.0);
    MV(uu, z);
    PC(z, vv);
    og_velocity_bcs(g, vv.u, vv.v, 0.0);
    gam = cdot(g, rt, vv);
    if (fabs(gam) < 1.0e-30) break;
    alpha = xi0 / gam;
    FOR_TRIP(g, { C3(r)[k] = C3(r)[k] - alpha * C3(vv)[k]; });
    if (consistent) og_velocity_bcs(g, r.u, r.v, 0.0);
    MV(r, z);
    PC(z, ss);
    og_velocity_bcs(g, ss.u, ss.v, 0.0);
    FOR_TRIP(g, { C3(X)[k] = C3(X)[k] + alpha * C3(uu)[k]; });

    /* odd step */
    xi1 = cdot(g, rt, ss);
    if (fabs(xi0) < 1.0e-30) break;
    beta = alpha * (xi1 / xi0);
    xi0 = xi1;
    FOR_TRIP(g, { C3(vv)[k] = C3(ss)[k] - beta * C3(vv)[k]; });
    og_velocity_bcs(g, vv.u, vv.v, 0.0);
    MV(vv, z);
    PC(z, ww);
    og_velocity_bcs(g, ww.u, ww.v, 0.0);
    gam = cdot(g, rt, ww);
    if (fabs(gam) < 1.0e-30) break;
    alpha = xi0 / gam;
    FOR_TRIP(g, { C3(uu)[k] = C3(r)[k] - beta * C3(uu)[k]; });
    FOR_TRIP(g, { C3(r)[k] = C3(r)[k] - alpha * C3(vv)[k]; });
    FOR_TRIP(g, { C3(ss)[k] = C3(ss)[k] - alpha * C3(ww)[k]; });
    if (consistent) og_velocity_bcs(g, ss.u, ss.v, 0.0);
    MV(ss, z);
    PC(z, tt);
    og_velocity_bcs(g, tt.u, tt.v, 0.0);

    /* GCR(2) */
    om1 = cdot(g, r, ss);
    lam = cdot(g, ss, ss);
    nu = cdot(g, ss, tt);
    tau = cdot(g, tt, tt);
    om2 = cdot(g, r, tt);
    if (lam * tau > 1.0e-30) {
      det = lam * tau - nu * nu;
      if (det > 1.0e-14 * (lam * tau)) {
        c1 = om1; c2 = om2;
        om1 = (c1 * tau - c2 * nu) / det;
        om2 = (c2 * lam - c1 * nu) / det;
      } else if (lam > 1.0e-30) {
        om1 = om1 / lam; om2 = 0.0;
      } else {
        om1 = 0.0; om2 = 0.0;
      }
    } else if (lam > 1.0e-30) {
      om1 = om1 / lam; om2 = 0.0;
    } else {
      om1 = 0.0; om2 = 0.0;
    }
    FOR_TRIP(g, { C3(X)[k] = C3(X)[k] + om1 * C3(r)[k] + om2 * C3(ss)[k] + alpha * C3(uu)[k]; });
    FOR_TRIP(g, { C3(r)[k] = C3(r)[k] - om1 * C3(ss)[k] - om2 * C3(tt)[k]; });
    FOR_TRIP(g, { C3(uu)[k] = C3(uu)[k] - om1 * C3(vv)[k] - om2 * C3(ww)[k]; });
    norm_r = sqrt(cdot(g, r, r));
    rel_res = norm_r / norm_b;
    og_velocity_bcs(g, X.u, X.v, 0.0);
  }
  if (has_lid) og_velocity_bcs(g, X.u, X.v, u_lid);
  if (iter_out) *iter_out = (k < max_iter) ? k : max_iter;
  if (res_out) *res_out = rel_res;
  (void)B;
  trip_free(bh); trip_free(r); trip_free(rt); trip_free(uu); trip_free(vv);
  trip_free(ww); trip_free(ss); trip_free(tt); trip_free(z); trip_free(q);
#undef MV
#undef PC
}

void og_bicgstab2(const og_grid* g, double* Xu, double* Xv, double* Xp,
                  const double* Bu, const double* Bv, const double* Bp,
                  const double* u_adv, const double* v_adv,
                  const double* rho_c, const double* rho_u, const double* rho_v,
                  const double* kappa, const double* zeta, const double* eta, double dt,
                  double tol, int max_iter, int* iter_out, double* res_out,
                  int has_lid, double u_lid,
                  const double* pen_u, const double* pen_v,
                  const double* mdot_u, const double* mdot_v) {
  og_bicgstab2_ex(g, Xu, Xv, Xp, Bu, Bv, Bp, u_adv, v_adv, rho_c, rho_u, rho_v, kappa, zeta, eta, dt,
                  tol, max_iter, iter_out, res_out, has_lid, u_lid, pen_u, pen_v, mdot_u, mdot_v, 0);
}

/* ------------------------------------------------------------------------ */
/* src/mod_vof.f90:25-33 */
double og_van_leer(double r) { return (r <= 0.0) ? 0.0 : (2.0 * r) / (1.0 + r); }

/* Superbee (P. L. Roe, Annu. Rev. Fluid Mech. 18 (1986) 337-365; P. K. Sweby, SIAM J. Numer. Anal.
 * 21 (1984) 995-1011): psi(r) = max(0, min(2r, 1), min(r, 2)), the upper edge of Sweby's TVD region.
 * Not in the reference (whose code is van Leer, src/mod_vof.f90:25-33); an opt-in limiter. */
double og_superbee(double r) {
  const double a = fmin_f(2.0 * r, 1.0), b = fmin_f(r, 2.0);
  return fmax_f(0.0, fmax_f(a, b));
}
/* The face limiter of og_vof_advection: 0 van Leer (the reference), 1 Superbee. */
static int g_vof_limiter = 0;
void og_set_vof_limiter(int lim) { g_vof_limiter = lim; }
int og_get_vof_limiter(void) { return g_vof_limiter; }
static double vof_psi(double r) { return g_vof_limiter == 1 ? og_superbee(r) : og_van_leer(r); }

/* src/mod_vof.f90:38-140 */
void og_vof_advection(const og_grid* g, const double* c_in, double* c_out,
                      const double* u, const double* v, double dt,
                      double* flux_x, double* flux_y) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const int lfx = nx + 1;
  double* Fx = zalloc((size_t)(nx + 1) * ny);       /* (0:nx, 1:ny) */
  double* Fy = zalloc((size_t)nx * (ny + 1));       /* (1:nx, 0:ny) */
#define FX(i, j) Fx[(size_t)(i) + (size_t)((j) - 1) * lfx]
#define FY(i, j) Fy[(size_t)((i) - 1) + (size_t)(j) * nx]
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double uf = U(u, i, j);
      double num, den, rr, cfl, cf;
      if (uf >= 0.0) {
        num = P(c_in, i, j) - P(c_in, i - 1, j);
        den = P(c_in, i + 1, j) - P(c_in, i, j);
        rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
        cfl = fabs(uf) * dt * inv_dx;
        cf = P(c_in, i, j) + half * (1.0 - cfl) * vof_psi(rr) * den;
      } else {
        num = P(c_in, i + 2, j) - P(c_in, i + 1, j);
        den = P(c_in, i + 1, j) - P(c_in, i, j);
        rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
        cfl = fabs(uf) * dt * inv_dx;
        cf = P(c_in, i + 1, j) - half * (1.0 - cfl) * vof_psi(rr) * den;
      }
      FX(i, j) = uf * clamp01(cf);
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double vf = V(v, i, j);
      double num, den, rr, cfl, cf;
      if (vf >= 0.0) {
        num = P(c_in, i, j) - P(c_in, i, j - 1);
        den = P(c_in, i, j + 1) - P(c_in, i, j);
        rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
        cfl = fabs(vf) * dt * inv_dy;
        cf = P(c_in, i, j) + half * (1.0 - cfl) * vof_psi(rr) * den;
      } else {
        num = P(c_in, i, j + 2) - P(c_in, i, j + 1);
        den = P(c_in, i, j + 1) - P(c_in, i, j);
        rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
        cfl = fabs(vf) * dt * inv_dy;
        cf = P(c_in, i, j + 1) - half * (1.0 - cfl) * vof_psi(rr) * den;
      }
      FY(i, j) = vf * clamp01(cf);
    }
  memcpy(c_out, c_in, og_size_p(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      P(c_out, i, j) = P(c_in, i, j) - dt * ((FX(i, j) - FX(i - 1, j)) * inv_dx + (FY(i, j) - FY(i, j - 1)) * inv_dy);
      P(c_out, i, j) = clamp01(P(c_out, i, j));
    }
  if (flux_x) {
    memset(flux_x, 0, og_size_u(g) * sizeof(double));
    for (int j = 1; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) U(flux_x, i, j) = FX(i, j);
    og_halo_u(g, flux_x);
  }
  if (flux_y) {
    memset(flux_y, 0, og_size_v(g) * sizeof(double));
    for (int j = 0; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) V(flux_y, i, j) = FY(i, j);
    og_halo_v(g, flux_y);
  }
  og_halo_p(g, c_out);
#undef FX
#undef FY
  free(Fx); free(Fy);
}

/* src/mod_vof.f90:145-194 (+ apply_neumann_bc :199-237) */
void og_svof(const og_grid* g, const double* c_raw, double* c_smooth, double delta,
             int max_iter, double tol, int* iter_out) {
  DECL;
  const double cx = (delta / g->dx) * (delta / g->dx);
  const double cy = (delta / g->dy) * (delta / g->dy);
  const double diag = 1.0 + 2.0 * (cx + cy);
  const size_t sp = og_size_p(g);
  double* c_old = zalloc(sp);
  int it;
  memcpy(c_smooth, c_raw, sp * sizeof(double));
  for (it = 1; it <= max_iter; ++it) {
    og_halo_p(g, c_smooth);
    neumann_mirror_p(g, c_smooth);
    memcpy(c_old, c_smooth, sp * sizeof(double));
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        P(c_smooth, i, j) = (P(c_raw, i, j) + cx * (P(c_old, i + 1, j) + P(c_old, i - 1, j)) +
                             cy * (P(c_old, i, j + 1) + P(c_old, i, j - 1))) / diag;
    double res = 0.0;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) res = fmax_f(res, fabs(P(c_smooth, i, j) - P(c_old, i, j)));
    if (res < tol) break;
  }
  og_halo_p(g, c_smooth);
  neumann_mirror_p(g, c_smooth);
  if (iter_out) *iter_out = (it < max_iter) ? it : max_iter;
  free(c_old);
}

/* src/mod_surface_tension.f90:25-72 */
void og_curvature(const og_grid* g, const double* cs, double* kc) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, eps = 1.0e-12;
  const size_t su = og_size_u(g), sv = og_size_v(g);
  double* nxu = zalloc(su);
  double* nyv = zalloc(sv);
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double gx = (P(cs, i + 1, j) - P(cs, i, j)) * inv_dx;
      const double gy = 0.25 * (P(cs, i, j + 1) + P(cs, i + 1, j + 1) - P(cs, i, j - 1) - P(cs, i + 1, j - 1)) * inv_dy;
      const double ng = sqrt(gx * gx + gy * gy + eps);
      U(nxu, i, j) = gx / ng;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double gx = 0.25 * (P(cs, i + 1, j) + P(cs, i + 1, j + 1) - P(cs, i - 1, j) - P(cs, i - 1, j + 1)) * inv_dx;
      const double gy = (P(cs, i, j + 1) - P(cs, i, j)) * inv_dy;
      const double ng = sqrt(gx * gx + gy * gy + eps);
      V(nyv, i, j) = gy / ng;
    }
  memset(kc, 0, og_size_p(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      P(kc, i, j) = -((U(nxu, i, j) - U(nxu, i - 1, j)) * inv_dx + (V(nyv, i, j) - V(nyv, i, j - 1)) * inv_dy);
  og_halo_p(g, kc);
  free(nxu); free(nyv);
}

/* src/mod_surface_tension.f90:78-109 */
void og_csf(const og_grid* g, const double* c_raw, const double* kc, double sigma,
            double* fu, double* fv) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  memset(fu, 0, og_size_u(g) * sizeof(double));
  memset(fv, 0, og_size_v(g) * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double kf = half * (P(kc, i, j) + P(kc, i + 1, j));
      U(fu, i, j) = sigma * kf * (P(c_raw, i + 1, j) - P(c_raw, i, j)) * inv_dx;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double kf = half * (P(kc, i, j) + P(kc, i, j + 1));
      V(fv, i, j) = sigma * kf * (P(c_raw, i, j + 1) - P(c_raw, i, j)) * inv_dy;
    }
}

/* ------------------------------------------------------------------------ */
/* src/mod_viscosity.f90:28-46 */
double og_harmonic_viscosity(double c, double mu1, double mu2) {
  const double cc = clamp01(c);
  if (fabs(mu1 - mu2) < 1.0e-14) return mu1;
  const double den = mu2 * cc + mu1 * (1.0 - cc);
  if (den > 1.0e-15) return (mu1 * mu2) / den;
  return fmin_f(mu1, mu2);
}

/* src/mod_viscosity.f90:54-93 */
void og_mixture_densities(const og_grid* g, const double* c, double rho1, double rho2,
                          double* rho_c, double* rho_u, double* rho_v) {
  DECL;
  const double half = 0.5;
  for (int j = 1 - nh; j <= ny + nh; ++j)
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      const double cv = clamp01(P(c, i, j));
      P(rho_c, i, j) = rho1 * cv + rho2 * (1.0 - cv);
    }
  for (int j = 1 - nh; j <= ny + nh; ++j) {
    for (int i = 1 - nh; i <= nx + nh - 1; ++i) U(rho_u, i, j) = half * (P(rho_c, i, j) + P(rho_c, i + 1, j));
    U(rho_u, -nh, j) = U(rho_u, 1 - nh, j);
    U(rho_u, nx + nh, j) = U(rho_u, nx + nh - 1, j);
  }
  for (int i = 1 - nh; i <= nx + nh; ++i) {
    for (int j = 1 - nh; j <= ny + nh - 1; ++j) V(rho_v, i, j) = half * (P(rho_c, i, j) + P(rho_c, i, j + 1));
    V(rho_v, i, -nh) = V(rho_v, i, 1 - nh);
    V(rho_v, i, ny + nh) = V(rho_v, i, ny + nh - 1);
  }
}

/* src/mod_viscosity.f90:105-151 */
void og_mixture_viscosities(const og_grid* g, const double* c, double mu1, double mu2,
                            double* kappa, double* zeta, double* eta) {
  DECL;
  for (int j = 1 - nh; j <= ny + nh; ++j)
    for (int i = 1 - nh; i <= nx + nh; ++i) P(kappa, i, j) = 2.0 * og_harmonic_viscosity(P(c, i, j), mu1, mu2);
  for (int j = 1 - nh; j <= ny + nh - 1; ++j)
    for (int i = 1 - nh; i <= nx + nh - 1; ++i) {
      const double cv = 0.25 * (P(c, i, j) + P(c, i + 1, j) + P(c, i, j + 1) + P(c, i + 1, j + 1));
      const double m = og_harmonic_viscosity(cv, mu1, mu2);
      W(zeta, i, j) = 2.0 * m;
      W(eta, i, j) = m;
    }
  for (int j = 1 - nh; j <= ny + nh - 1; ++j) {
    W(zeta, -nh, j) = W(zeta, 1 - nh, j);
    W(zeta, nx + nh, j) = W(zeta, nx + nh - 1, j);
    W(eta, -nh, j) = W(eta, 1 - nh, j);
    W(eta, nx + nh, j) = W(eta, nx + nh - 1, j);
  }
  for (int i = -nh; i <= nx + nh; ++i) {
    W(zeta, i, -nh) = W(zeta, i, 1 - nh);
    W(zeta, i, ny + nh) = W(zeta, i, ny + nh - 1);
    W(eta, i, -nh) = W(eta, i, 1 - nh);
    W(eta, i, ny + nh) = W(eta, i, ny + nh - 1);
  }
}

/* src/mod_solver_coupled.f90:41-91 */
double og_adaptive_dt(const og_grid* g, const og_phys* ph, const double* u, const double* v,
                      double dt_max) {
  DECL;
  const double pi = 3.141592653589793238462643383279502884197;
  const double h = fmin_f(g->dx, g->dy);
  double um = 0.0, vm = 0.0, dt_cfl, dt_loc;
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) um = fmax_f(um, fabs(U(u, i, j)));
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) vm = fmax_f(vm, fabs(V(v, i, j)));
  if (um / g->dx + vm / g->dy > 1.0e-12) dt_cfl = ph->cfl_target / (um / g->dx + vm / g->dy);
  else dt_cfl = dt_max;
  dt_loc = fmin_f(dt_max, dt_cfl);
  if (ph->sigma > 1.0e-14) {
    const double dt_cap = 0.5 * sqrt((ph->rho1 + ph->rho2) * (h * h * h) / (4.0 * pi * ph->sigma));
    dt_loc = fmin_f(dt_loc, dt_cap);
  }
  const double gm = sqrt(ph->gx * ph->gx + ph->gy * ph->gy);
  if (gm > 1.0e-14) {
    const double dt_g = 0.5 * sqrt(h / gm);
    dt_loc = fmin_f(dt_loc, dt_g);
  }
  return dt_loc;
}

/* src/mod_solver_coupled.f90:137-380 */
void og_advance_step(const og_grid* g, const og_phys* ph, const og_step_opts* o,
                     double* u, double* v, double* p, double* c,
                     double* u_prev, double* v_prev,
                     double* rho_c, double* rho_u, double* rho_v,
                     double* kappa, double* zeta, double* eta,
                     double* kappa_curv, double* f_u, double* f_v,
                     double dt, int step_idx,
                     const double* phi_s_u, const double* phi_s_v,
                     int* iter_out, double* res_out) {
  DECL;
  const size_t su = og_size_u(g), sv = og_size_v(g), sp = og_size_p(g);
  const int has_pen = (phi_s_u != NULL) && (phi_s_v != NULL);
  const int use_bdf2 = (ph->time_order == 2) && (step_idx >= 2);
  const double dt_eff = use_bdf2 ? dt / 1.5 : dt;
  const double u_lid = o->has_lid ? o->u_lid : 0.0;
  const double half = 0.5;
  double *un = zalloc(su), *vn = zalloc(sv), *unm1 = zalloc(su), *vnm1 = zalloc(sv);
  double *uadv = zalloc(su), *vadv = zalloc(sv), *cnew = zalloc(sp), *cs = zalloc(sp);
  double *Bu = zalloc(su), *Bv = zalloc(sv), *Bp = zalloc(sp);
  double *fcx = NULL, *fcy = NULL, *mu_ = NULL, *mv_ = NULL, *penu = NULL, *penv = NULL;
  int it = 0;
  double res = 0.0;

  memcpy(un, u, su * sizeof(double));
  memcpy(vn, v, sv * sizeof(double));
  memcpy(unm1, u_prev, su * sizeof(double));
  memcpy(vnm1, v_prev, sv * sizeof(double));
  if (step_idx <= 1) {
    memcpy(uadv, un, su * sizeof(double));
    memcpy(vadv, vn, sv * sizeof(double));
  } else {
    for (size_t k = 0; k < su; ++k) uadv[k] = 2.0 * un[k] - unm1[k];
    for (size_t k = 0; k < sv; ++k) vadv[k] = 2.0 * vn[k] - vnm1[k];
  }
  memcpy(u_prev, un, su * sizeof(double));
  memcpy(v_prev, vn, sv * sizeof(double));

  if (o->use_amp) {
    fcx = zalloc(su); fcy = zalloc(sv); mu_ = zalloc(su); mv_ = zalloc(sv);
    og_vof_advection(g, c, cnew, un, vn, dt, fcx, fcy);
    for (size_t k = 0; k < su; ++k) mu_[k] = (ph->rho1 - ph->rho2) * fcx[k] + ph->rho2 * uadv[k];
    for (size_t k = 0; k < sv; ++k) mv_[k] = (ph->rho1 - ph->rho2) * fcy[k] + ph->rho2 * vadv[k];
  } else {
    og_vof_advection(g, c, cnew, un, vn, dt, NULL, NULL);
  }
  memcpy(c, cnew, sp * sizeof(double));

  og_mixture_densities(g, c, ph->rho1, ph->rho2, rho_c, rho_u, rho_v);
  og_mixture_viscosities(g, c, ph->mu1, ph->mu2, kappa, zeta, eta);

  if (ph->sigma > 1.0e-14) {
    const double delta = 1.5 * fmin_f(g->dx, g->dy);
    og_svof(g, c, cs, delta, 150, 1.0e-7, NULL);
    og_curvature(g, cs, kappa_curv);
    og_csf(g, c, kappa_curv, ph->sigma, f_u, f_v);
  } else {
    memset(kappa_curv, 0, sp * sizeof(double));
    memset(f_u, 0, su * sizeof(double));
    memset(f_v, 0, sv * sizeof(double));
  }

  if (has_pen) {
    penu = zalloc(su); penv = zalloc(sv);
    for (size_t k = 0; k < su; ++k) penu[k] = ph->alpha_penalty * phi_s_u[k];
    for (size_t k = 0; k < sv; ++k) penv[k] = ph->alpha_penalty * phi_s_v[k];
  }

  {
    const int i_u_max = g->per_x ? nx : nx - 1;
    const int j_v_max = g->per_y ? ny : ny - 1;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= i_u_max; ++i) {
        if (use_bdf2)
          U(Bu, i, j) = (U(rho_u, i, j) / dt) * (2.0 * U(un, i, j) - half * U(unm1, i, j)) +
                        U(rho_u, i, j) * ph->gx + U(f_u, i, j);
        else
          U(Bu, i, j) = (U(rho_u, i, j) / dt) * U(un, i, j) + U(rho_u, i, j) * ph->gx + U(f_u, i, j);
        if (has_pen) U(Bu, i, j) = U(Bu, i, j) + U(penu, i, j) * o->u_solid;
      }
    if (o->has_lid && !g->per_y)
      for (int i = 1; i <= i_u_max; ++i)
        U(Bu, i, ny) = U(Bu, i, ny) + (2.0 * (W(zeta, i, ny) - W(eta, i, ny)) / (g->dy * g->dy)) * u_lid;
    for (int j = 1; j <= j_v_max; ++j)
      for (int i = 1; i <= nx; ++i) {
        if (use_bdf2)
          V(Bv, i, j) = (V(rho_v, i, j) / dt) * (2.0 * V(vn, i, j) - half * V(vnm1, i, j)) +
                        V(rho_v, i, j) * ph->gy + V(f_v, i, j);
        else
          V(Bv, i, j) = (V(rho_v, i, j) / dt) * V(vn, i, j) + V(rho_v, i, j) * ph->gy + V(f_v, i, j);
        if (has_pen) V(Bv, i, j) = V(Bv, i, j) + V(penv, i, j) * o->v_solid;
      }
  }

  if (o->use_bicgstab2)
    og_bicgstab2_ex(g, u, v, p, Bu, Bv, Bp, uadv, vadv, rho_c, rho_u, rho_v, kappa, zeta, eta, dt_eff,
                    o->tol, o->max_iter, &it, &res, 1, u_lid, penu, penv,
                    o->use_amp ? mu_ : NULL, o->use_amp ? mv_ : NULL, o->consistent);
  else
    og_bicgstab(g, u, v, p, Bu, Bv, Bp, uadv, vadv, rho_c, rho_u, rho_v, kappa, zeta, eta, dt_eff,
                o->tol, o->max_iter, &it, &res, 1, u_lid, penu, penv,
                o->use_amp ? mu_ : NULL, o->use_amp ? mv_ : NULL);

  og_halo_u(g, u);
  og_halo_v(g, v);
  og_halo_p(g, p);
  og_velocity_bcs(g, u, v, u_lid);
  if (iter_out) *iter_out = it;
  if (res_out) *res_out = res;

  free(un); free(vn); free(unm1); free(vnm1); free(uadv); free(vadv); free(cnew); free(cs);
  free(Bu); free(Bv); free(Bp);
  free(fcx); free(fcy); free(mu_); free(mv_); free(penu); free(penv);
}
