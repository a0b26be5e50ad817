/*
 * fugu2d_adjoint.c -- CPU restatement of the adjoint building blocks (no counterpart in the
 * reference: it has no adjoint).  TEST INFRASTRUCTURE ONLY, like fugu2d_oracle.c.
 *
 * The linear operator of one coupled step, as the solver sees it, acts on the owned degrees of
 * freedom (u: j=1..ny, i=1..i_u_max; v: i=1..nx, j=1..j_v_max; p: interior):
 *
 *     A_eff X = R  A_pad  H  B  E0  X
 *
 * E0 embeds the owned values (zero elsewhere), B is the linear part of apply_velocity_bcs
 * (src/mod_operators.f90:254-333, lid value 0), H the single-rank halo exchange
 * (src/mod_mpi.f90:113-245), A_pad the padded matvec (src/mod_operators.f90:341-443) and R the
 * restriction to owned rows.  Every transpose here is the reverse-mode of the forward loops
 * (statements in reverse order, each "dst = f(src)" turned into "src_bar += f'(src) dst_bar"),
 * so that the dot-product identity <A x, y> = <x, A^T y> holds to rounding.
 *
 * The linearised PCD (src/mod_pcd_precond.f90:14-104) replaces the truncated PCG by a
 * projected pressure solve to a caller-given tolerance, phi = Pi A_p^+ Pi z_p (Pi: zero-mean
 * projection, applied for every boundary type since A_p is singular for all of them), so it is
 * a fixed linear map whose transpose is well defined.
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

static double* zalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }

/* ------------------------------------------------------------------------ */
void og_mask_owned(const og_grid* g, double* u, double* v, double* p) {
  DECL;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const int j_v_max = g->per_y ? ny : ny - 1;
  if (u)
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int i = -nh; i <= nx + nh; ++i)
        if (!(j >= 1 && j <= ny && i >= 1 && i <= i_u_max)) U(u, i, j) = 0.0;
  if (v)
    for (int j = -nh; j <= ny + nh; ++j)
      for (int i = 1 - nh; i <= nx + nh; ++i)
        if (!(j >= 1 && j <= j_v_max && i >= 1 && i <= nx)) V(v, i, j) = 0.0;
  if (p)
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int i = 1 - nh; i <= nx + nh; ++i)
        if (!(j >= 1 && j <= ny && i >= 1 && i <= nx)) P(p, i, j) = 0.0;
}

/* Transposed halo exchange: the y pass undone first, then the x pass; a copy dst = src
 * becomes src_bar += dst_bar, dst_bar = 0. */
void og_halo_p_T(const og_grid* g, double* a) {
  DECL;
  if (g->per_y) {
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) { P(a, i, 1 + k) += P(a, i, ny + 1 + k); P(a, i, ny + 1 + k) = 0.0; }
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) { P(a, i, ny - nh + 1 + k) += P(a, i, 1 - nh + k); P(a, i, 1 - nh + k) = 0.0; }
  }
  if (g->per_x) {
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) { P(a, 1 + k, j) += P(a, nx + 1 + k, j); P(a, nx + 1 + k, j) = 0.0; }
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) { P(a, nx - nh + 1 + k, j) += P(a, 1 - nh + k, j); P(a, 1 - nh + k, j) = 0.0; }
  }
}

void og_halo_u_T(const og_grid* g, double* a) {
  DECL;
  if (g->per_y) {
    for (int k = 0; k < nh; ++k)
      for (int i = -nh; i <= nx + nh; ++i) { U(a, i, 1 + k) += U(a, i, ny + 1 + k); U(a, i, ny + 1 + k) = 0.0; }
    for (int k = 0; k < nh; ++k)
      for (int i = -nh; i <= nx + nh; ++i) { U(a, i, ny - nh + 1 + k) += U(a, i, 1 - nh + k); U(a, i, 1 - nh + k) = 0.0; }
  }
  if (g->per_x) {
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) { U(a, 1 + k, j) += U(a, nx + 1 + k, j); U(a, nx + 1 + k, j) = 0.0; }
    for (int j = 1 - nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) { U(a, nx - nh + 1 + k, j) += U(a, 1 - nh + k, j); U(a, 1 - nh + k, j) = 0.0; }
  }
}

void og_halo_v_T(const og_grid* g, double* a) {
  DECL;
  if (g->per_y) {
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) { V(a, i, 1 + k) += V(a, i, ny + 1 + k); V(a, i, ny + 1 + k) = 0.0; }
    for (int k = 0; k < nh; ++k)
      for (int i = 1 - nh; i <= nx + nh; ++i) { V(a, i, ny - nh + 1 + k) += V(a, i, 1 - nh + k); V(a, i, 1 - nh + k) = 0.0; }
  }
  if (g->per_x) {
    for (int j = -nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) { V(a, 1 + k, j) += V(a, nx + 1 + k, j); V(a, nx + 1 + k, j) = 0.0; }
    for (int j = -nh; j <= ny + nh; ++j)
      for (int k = 0; k < nh; ++k) { V(a, nx - nh + 1 + k, j) += V(a, 1 - nh + k, j); V(a, 1 - nh + k, j) = 0.0; }
  }
}

/* Transpose of the linear part of og_velocity_bcs: its statements in reverse order. */
void og_velocity_bcs_T(const og_grid* g, double* u, double* v) {
  DECL;
  if (!g->per_y) {
    for (int i = -nh; i <= nx + nh; ++i)
      for (int j = 1; j <= nh; ++j) { U(u, i, ny - j + 1) -= U(u, i, ny + j); U(u, i, ny + j) = 0.0; }
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      for (int j = 1; j <= nh; ++j) { V(v, i, ny - j) -= V(v, i, ny + j); V(v, i, ny + j) = 0.0; }
      V(v, i, ny) = 0.0;
    }
    for (int i = -nh; i <= nx + nh; ++i)
      for (int j = 1; j <= nh; ++j) { U(u, i, j) -= U(u, i, 1 - j); U(u, i, 1 - j) = 0.0; }
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      for (int j = 1; j <= nh; ++j) { V(v, i, j) -= V(v, i, -j); V(v, i, -j) = 0.0; }
      V(v, i, 0) = 0.0;
    }
  }
  if (!g->per_x) {
    for (int j = -nh; j <= ny + nh; ++j)
      for (int i = 1; i <= nh; ++i) { V(v, nx - i + 1, j) -= V(v, nx + i, j); V(v, nx + i, j) = 0.0; }
    for (int j = 1 - nh; j <= ny + nh; ++j) {
      for (int i = 1; i <= nh; ++i) { U(u, nx - i, j) -= U(u, nx + i, j); U(u, nx + i, j) = 0.0; }
      U(u, nx, j) = 0.0;
    }
    for (int j = -nh; j <= ny + nh; ++j)
      for (int i = 1; i <= nh; ++i) { V(v, i, j) -= V(v, 1 - i, j); V(v, 1 - i, j) = 0.0; }
    for (int j = 1 - nh; j <= ny + nh; ++j) {
      for (int i = 1; i <= nh; ++i) { U(u, i, j) -= U(u, -i, j); U(u, -i, j) = 0.0; }
      U(u, 0, j) = 0.0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* A_eff X: owned X -> owned Y (inputs are not modified). */
void og_matvec_eff(const og_grid* g, const double* Xu, const double* Xv, const double* Xp,
                   double* Yu, double* Yv, double* Yp, const double* u_adv, const double* v_adv,
                   const double* rho_c, const double* rho_u, const double* rho_v,
                   const double* kappa, const double* zeta, const double* eta, double dt, int include_adv) {
  const size_t su = og_size_u(g), sv = og_size_v(g), sp = og_size_p(g);
  double* xu = zalloc(su); double* xv = zalloc(sv); double* xp = zalloc(sp);
  memcpy(xu, Xu, su * sizeof(double)); memcpy(xv, Xv, sv * sizeof(double)); memcpy(xp, Xp, sp * sizeof(double));
  og_mask_owned(g, xu, xv, xp);
  og_velocity_bcs(g, xu, xv, 0.0);
  og_matvec_coupled(g, xu, xv, xp, Yu, Yv, Yp, u_adv, v_adv, rho_c, rho_u, rho_v, kappa, zeta, eta, dt,
                    include_adv, NULL, NULL, NULL, NULL);
  og_mask_owned(g, Yu, Yv, Yp);
  free(xu); free(xv); free(xp);
}

/* Reverse-mode of the padded matvec rows restricted to owned rows: L (owned rows) -> Xbar
 * (padded, accumulated).  Mirrors og_matvec_coupled / og_viscous_stress / og_advection /
 * og_grad_pressure / og_divergence term by term. */
static void matvec_pad_T(const og_grid* g, const double* Lu, const double* Lv, const double* Lp,
                         double* Ub, double* Vb, double* Pb, const double* u_adv, const double* v_adv,
                         const double* rho_c, const double* rho_u, const double* rho_v, const double* kappa,
                         const double* zeta, const double* eta, double dt, int include_adv) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const int j_v_max = g->per_y ? ny : ny - 1;
  const int lxx = nx + 2, lyy = nx, lxy = nx + 1;
  double* Txx = zalloc((size_t)lxx * ny);
  double* Tyy = zalloc((size_t)lyy * (ny + 2));
  double* Txy = zalloc((size_t)lxy * (ny + 1));
  double* Tyx = zalloc((size_t)lxy * (ny + 1));
#define TXX(i, j) Txx[(size_t)(i) + (size_t)((j) - 1) * lxx]
#define TYY(i, j) Tyy[(size_t)((i) - 1) + (size_t)(j) * lyy]
#define TXY(i, j) Txy[(size_t)(i) + (size_t)(j) * lxy]
#define TYX(i, j) Tyx[(size_t)(i) + (size_t)(j) * lxy]
  /* p rows: divergence */
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double a = P(Lp, i, j);
      U(Ub, i, j) += a * inv_dx; U(Ub, i - 1, j) -= a * inv_dx;
      V(Vb, i, j) += a * inv_dy; V(Vb, i, j - 1) -= a * inv_dy;
    }
  /* u rows */
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= i_u_max; ++i) {
      const double a = U(Lu, i, j);
      U(Ub, i, j) += (U(rho_u, i, j) / dt) * a;
      P(Pb, i + 1, j) += a * inv_dx; P(Pb, i, j) -= a * inv_dx;
      TXX(i + 1, j) -= a * inv_dx; TXX(i, j) += a * inv_dx;
      TXY(i, j) -= a * inv_dy; TXY(i, j - 1) += a * inv_dy;
      if (include_adv) {
        const double mE = P(rho_c, i + 1, j) * half * (U(u_adv, i, j) + U(u_adv, i + 1, j));
        const double mW = P(rho_c, i, j) * half * (U(u_adv, i - 1, j) + U(u_adv, i, j));
        const double mN = half * (V(rho_v, i, j) + V(rho_v, i + 1, j)) * half * (V(v_adv, i, j) + V(v_adv, i + 1, j));
        const double mS = half * (V(rho_v, i, j - 1) + V(rho_v, i + 1, j - 1)) * half *
                          (V(v_adv, i, j - 1) + V(v_adv, i + 1, j - 1));
        U(Ub, i, j) += a * half * (mE - mW) * inv_dx + a * half * (mN - mS) * inv_dy;
        U(Ub, i + 1, j) += a * half * mE * inv_dx;
        U(Ub, i - 1, j) -= a * half * mW * inv_dx;
        U(Ub, i, j + 1) += a * half * mN * inv_dy;
        U(Ub, i, j - 1) -= a * half * mS * inv_dy;
      }
    }
  /* v rows */
  for (int j = 1; j <= j_v_max; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double a = V(Lv, i, j);
      V(Vb, i, j) += (V(rho_v, i, j) / dt) * a;
      P(Pb, i, j + 1) += a * inv_dy; P(Pb, i, j) -= a * inv_dy;
      TYX(i, j) -= a * inv_dx; TYX(i - 1, j) += a * inv_dx;
      TYY(i, j + 1) -= a * inv_dy; TYY(i, j) += a * inv_dy;
      if (include_adv) {
        const double mE = half * (U(rho_u, i, j) + U(rho_u, i, j + 1)) * half * (U(u_adv, i, j) + U(u_adv, i, j + 1));
        const double mW = half * (U(rho_u, i - 1, j) + U(rho_u, i - 1, j + 1)) * half *
                          (U(u_adv, i - 1, j) + U(u_adv, i - 1, j + 1));
        const double mN = P(rho_c, i, j + 1) * half * (V(v_adv, i, j) + V(v_adv, i, j + 1));
        const double mS = P(rho_c, i, j) * half * (V(v_adv, i, j - 1) + V(v_adv, i, j));
        V(Vb, i, j) += a * half * (mE - mW) * inv_dx + a * half * (mN - mS) * inv_dy;
        V(Vb, i + 1, j) += a * half * mE * inv_dx;
        V(Vb, i - 1, j) -= a * half * mW * inv_dx;
        V(Vb, i, j + 1) += a * half * mN * inv_dy;
        V(Vb, i, j - 1) -= a * half * mS * inv_dy;
      }
    }
  /* stresses */
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx + 1; ++i) {
      const double t = P(kappa, i, j) * TXX(i, j) * inv_dx;
      U(Ub, i, j) += t; U(Ub, i - 1, j) -= t;
    }
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double t = P(kappa, i, j) * TYY(i, j) * inv_dy;
      V(Vb, i, j) += t; V(Vb, i, j - 1) -= t;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double zme = W(zeta, i, j) - W(eta, i, j), et = W(eta, i, j);
      const double dub = (zme * TXY(i, j) + et * TYX(i, j)) * inv_dy;
      const double dvb = (et * TXY(i, j) + zme * TYX(i, j)) * inv_dx;
      U(Ub, i, j + 1) += dub; U(Ub, i, j) -= dub;
      V(Vb, i + 1, j) += dvb; V(Vb, i, j) -= dvb;
    }
#undef TXX
#undef TYY
#undef TXY
#undef TYX
  free(Txx); free(Tyy); free(Txy); free(Tyx);
}

/* A_eff^T L: owned L -> owned Xbar. */
void og_matvec_T(const og_grid* g, const double* Lu, const double* Lv, const double* Lp,
                 double* Xu, double* Xv, double* Xp, const double* u_adv, const double* v_adv,
                 const double* rho_c, const double* rho_u, const double* rho_v,
                 const double* kappa, const double* zeta, const double* eta, double dt, int include_adv) {
  const size_t su = og_size_u(g), sv = og_size_v(g), sp = og_size_p(g);
  double* lu = zalloc(su); double* lv = zalloc(sv); double* lp = zalloc(sp);
  memcpy(lu, Lu, su * sizeof(double)); memcpy(lv, Lv, sv * sizeof(double)); memcpy(lp, Lp, sp * sizeof(double));
  og_mask_owned(g, lu, lv, lp);
  memset(Xu, 0, su * sizeof(double)); memset(Xv, 0, sv * sizeof(double)); memset(Xp, 0, sp * sizeof(double));
  matvec_pad_T(g, lu, lv, lp, Xu, Xv, Xp, u_adv, v_adv, rho_c, rho_u, rho_v, kappa, zeta, eta, dt, include_adv);
  og_halo_p_T(g, Xp);
  og_halo_v_T(g, Xv);
  og_halo_u_T(g, Xu);
  og_velocity_bcs_T(g, Xu, Xv);
  og_mask_owned(g, Xu, Xv, Xp);
  free(lu); free(lv); free(lp);
}

/* ------------------------------------------------------------------------ */
/* Linearised PCD and its transpose. */

/* phi = Pi A_p^+ Pi z (interior written, ghosts 0): projected Jacobi-PCG to tol. */
static void pressure_solve_proj(const og_grid* g, double* phi, const double* z, const double* rho_u,
                                const double* rho_v, double tol, int max_iter) {
  DECL;
  const size_t sp = og_size_p(g);
  double* rhs = zalloc(sp);
  double mean = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) mean += P(z, i, j);
  mean /= (double)(nx * ny);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) P(rhs, i, j) = P(z, i, j) - mean;
  memset(phi, 0, sp * sizeof(double));
  og_poisson_pcg(g, phi, rhs, rho_u, rho_v, tol, max_iter, NULL, NULL);
  mean = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) mean += P(phi, i, j);
  mean /= (double)(nx * ny);
  memset(rhs, 0, sp * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) P(rhs, i, j) = P(phi, i, j) - mean;
  memcpy(phi, rhs, sp * sizeof(double));
  free(rhs);
}

/* Transposed damped-Jacobi sweeps of og_helmholtz_u (n sweeps from zero): xbar (adjoint of the
 * output, consumed) -> rhs_bar (accumulated). */
static void helmholtz_u_T(const og_grid* g, double* xbar, double* rhs_bar, const double* rho_u,
                          const double* kappa, const double* zeta, const double* eta, double dt, int n) {
  DECL;
  const double omega = 0.8;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const double inv_dx2 = 1.0 / (g->dx * g->dx), inv_dy2 = 1.0 / (g->dy * g->dy);
  const size_t su = og_size_u(g);
  double* hb = zalloc(su);
  for (int sweep = n; sweep >= 1; --sweep) {
    memcpy(hb, xbar, su * sizeof(double));     /* entries the sweep leaves alone pass through */
    for (int j = 1; j <= ny; ++j) {
      for (int i = 1; i <= i_u_max; ++i) U(hb, i, j) = 0.0;
      if (!g->per_x) { U(hb, 0, j) = 0.0; U(hb, nx, j) = 0.0; }
    }
    for (int j = 1; j <= ny; ++j) {
      for (int i = 1; i <= i_u_max; ++i) {
        const double cE = P(kappa, i + 1, j) * inv_dx2;
        const double cW = P(kappa, i, j) * inv_dx2;
        const double cN = (W(zeta, i, j) - W(eta, i, j)) * inv_dy2;
        const double cS = (W(zeta, i, j - 1) - W(eta, i, j - 1)) * inv_dy2;
        const double diag = (U(rho_u, i, j) / dt) + cE + cW + cN + cS;
        const double a = U(xbar, i, j);
        U(hb, i, j) += (1.0 - omega) * a;
        const double s = omega * a / diag;
        U(rhs_bar, i, j) += s;
        U(hb, i + 1, j) += cE * s; U(hb, i - 1, j) += cW * s;
        U(hb, i, j + 1) += cN * s; U(hb, i, j - 1) += cS * s;
      }
      if (!g->per_x) {
        U(rhs_bar, 0, j) += U(xbar, 0, j);
        U(rhs_bar, nx, j) += U(xbar, nx, j);
      }
    }
    og_halo_u_T(g, hb);
    memcpy(xbar, hb, su * sizeof(double));
  }
  free(hb);
}

static void helmholtz_v_T(const og_grid* g, double* xbar, double* rhs_bar, const double* rho_v,
                          const double* kappa, const double* zeta, const double* eta, double dt, int n) {
  DECL;
  const double omega = 0.8;
  const int j_v_max = g->per_y ? ny : ny - 1;
  const double inv_dx2 = 1.0 / (g->dx * g->dx), inv_dy2 = 1.0 / (g->dy * g->dy);
  const size_t sv = og_size_v(g);
  double* hb = zalloc(sv);
  for (int sweep = n; sweep >= 1; --sweep) {
    memcpy(hb, xbar, sv * sizeof(double));
    for (int j = 1; j <= j_v_max; ++j)
      for (int i = 1; i <= nx; ++i) V(hb, i, j) = 0.0;
    if (!g->per_y)
      for (int i = 1; i <= nx; ++i) { V(hb, i, 0) = 0.0; V(hb, i, ny) = 0.0; }
    for (int j = 1; j <= j_v_max; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double cE = (W(zeta, i, j) - W(eta, i, j)) * inv_dx2;
        const double cW = (W(zeta, i - 1, j) - W(eta, i - 1, j)) * inv_dx2;
        const double cN = P(kappa, i, j + 1) * inv_dy2;
        const double cS = P(kappa, i, j) * inv_dy2;
        const double diag = (V(rho_v, i, j) / dt) + cE + cW + cN + cS;
        const double a = V(xbar, i, j);
        V(hb, i, j) += (1.0 - omega) * a;
        const double s = omega * a / diag;
        V(rhs_bar, i, j) += s;
        V(hb, i + 1, j) += cE * s; V(hb, i - 1, j) += cW * s;
        V(hb, i, j + 1) += cN * s; V(hb, i, j - 1) += cS * s;
      }
    if (!g->per_y)
      for (int i = 1; i <= nx; ++i) {
        V(rhs_bar, i, 0) += V(xbar, i, 0);
        V(rhs_bar, i, ny) += V(xbar, i, ny);
      }
    og_halo_v_T(g, hb);
    memcpy(xbar, hb, sv * sizeof(double));
  }
  free(hb);
}

/* Forward linearised PCD: owned z -> owned r.  visc: coefficient of the viscous Schur term
 * r_p -= visc * kappa * z_p (0 = the reference's PCD). */
void og_pcd_lin(const og_grid* g, const double* zu_in, const double* zv_in, const double* zp_in,
                double* ru, double* rv, double* rp, const double* u_adv, const double* v_adv,
                const double* rho_u, const double* rho_v, const double* kappa, const double* zeta,
                const double* eta, double dt, double ptol, int pmax, double visc) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const size_t su = og_size_u(g), sv = og_size_v(g), sp = og_size_p(g);
  double* zu = zalloc(su); double* zv = zalloc(sv); double* zp = zalloc(sp);
  memcpy(zu, zu_in, su * sizeof(double)); memcpy(zv, zv_in, sv * sizeof(double)); memcpy(zp, zp_in, sp * sizeof(double));
  og_mask_owned(g, zu, zv, zp);
  double* phi = zalloc(sp); double* zuw = zalloc(su); double* zvw = zalloc(sv);
  memset(ru, 0, su * sizeof(double)); memset(rv, 0, sv * sizeof(double)); memset(rp, 0, sp * sizeof(double));
  pressure_solve_proj(g, phi, zp, rho_u, rho_v, ptol, pmax);
  og_halo_p(g, phi);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double u_avg = half * (U(u_adv, i, j) + U(u_adv, i - 1, j));
      const double v_avg = half * (V(v_adv, i, j) + V(v_adv, i, j - 1));
      const double conv = u_avg * (P(phi, i + 1, j) - P(phi, i - 1, j)) * (half * inv_dx) +
                          v_avg * (P(phi, i, j + 1) - P(phi, i, j - 1)) * (half * inv_dy);
      P(rp, i, j) = -((1.0 / dt) * P(phi, i, j) + conv) - visc * P(kappa, i, j) * P(zp, i, j);
    }
  og_halo_p(g, rp);
  for (int j = 1; j <= ny; ++j) {
    for (int i = 1; i <= nx - 1; ++i) U(zuw, i, j) = U(zu, i, j) - (P(rp, i + 1, j) - P(rp, i, j)) * inv_dx;
    U(zuw, 0, j) = U(zu, 0, j);
    U(zuw, nx, j) = U(zu, nx, j);
  }
  for (int j = 1; j <= ny - 1; ++j)
    for (int i = 1; i <= nx; ++i) V(zvw, i, j) = V(zv, i, j) - (P(rp, i, j + 1) - P(rp, i, j)) * inv_dy;
  for (int i = 1; i <= nx; ++i) { V(zvw, i, 0) = V(zv, i, 0); V(zvw, i, ny) = V(zv, i, ny); }
  og_helmholtz_v(g, rv, zvw, rho_v, kappa, zeta, eta, dt, 3, NULL);
  og_halo_v(g, rv);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx - 1; ++i) {
      const double cross = -inv_dy * (W(eta, i, j) * (V(rv, i + 1, j) - V(rv, i, j)) * inv_dx -
                                      W(eta, i, j - 1) * (V(rv, i + 1, j - 1) - V(rv, i, j - 1)) * inv_dx);
      U(zuw, i, j) = U(zuw, i, j) - cross;
    }
  og_helmholtz_u(g, ru, zuw, rho_u, kappa, zeta, eta, dt, 3, NULL);
  og_halo_u(g, ru);
  og_mask_owned(g, ru, rv, rp);
  free(zu); free(zv); free(zp); free(phi); free(zuw); free(zvw);
}

/* Transposed linearised PCD: owned y -> owned zbar, the reverse of og_pcd_lin. */
void og_pcd_lin_T(const og_grid* g, const double* yu, const double* yv, const double* yp,
                  double* zu, double* zv, double* zp, const double* u_adv, const double* v_adv,
                  const double* rho_u, const double* rho_v, const double* kappa, const double* zeta,
                  const double* eta, double dt, double ptol, int pmax, double visc) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const size_t su = og_size_u(g), sv = og_size_v(g), sp = og_size_p(g);
  double* rub = zalloc(su); double* rvb = zalloc(sv); double* rpb = zalloc(sp);
  double* zuwb = zalloc(su); double* zvwb = zalloc(sv); double* phib = zalloc(sp); double* phs = zalloc(sp);
  memcpy(rub, yu, su * sizeof(double)); memcpy(rvb, yv, sv * sizeof(double)); memcpy(rpb, yp, sp * sizeof(double));
  og_mask_owned(g, rub, rvb, rpb);
  memset(zu, 0, su * sizeof(double)); memset(zv, 0, sv * sizeof(double)); memset(zp, 0, sp * sizeof(double));
  /* ru = H_u J_u(zuw) */
  og_halo_u_T(g, rub);
  helmholtz_u_T(g, rub, zuwb, rho_u, kappa, zeta, eta, dt, 3);
  /* zuw -= cross(rv) */
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx - 1; ++i) {
      const double a = U(zuwb, i, j);
      const double cN = a * inv_dy * W(eta, i, j) * inv_dx;
      const double cS = a * inv_dy * W(eta, i, j - 1) * inv_dx;
      V(rvb, i + 1, j) += cN; V(rvb, i, j) -= cN;
      V(rvb, i + 1, j - 1) -= cS; V(rvb, i, j - 1) += cS;
    }
  /* rv = H_v J_v(zvw) */
  og_halo_v_T(g, rvb);
  helmholtz_v_T(g, rvb, zvwb, rho_v, kappa, zeta, eta, dt, 3);
  /* zvw, zuw from z and rp */
  for (int i = 1; i <= nx; ++i) { V(zv, i, 0) += V(zvwb, i, 0); V(zv, i, ny) += V(zvwb, i, ny); }
  for (int j = 1; j <= ny - 1; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double a = V(zvwb, i, j);
      V(zv, i, j) += a;
      P(rpb, i, j + 1) -= a * inv_dy; P(rpb, i, j) += a * inv_dy;
    }
  for (int j = 1; j <= ny; ++j) {
    U(zu, 0, j) += U(zuwb, 0, j);
    U(zu, nx, j) += U(zuwb, nx, j);
    for (int i = 1; i <= nx - 1; ++i) {
      const double a = U(zuwb, i, j);
      U(zu, i, j) += a;
      P(rpb, i + 1, j) -= a * inv_dx; P(rpb, i, j) += a * inv_dx;
    }
  }
  /* rp from phi (and z_p) */
  og_halo_p_T(g, rpb);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double b = P(rpb, i, j);
      const double u_avg = half * (U(u_adv, i, j) + U(u_adv, i - 1, j));
      const double v_avg = half * (V(v_adv, i, j) + V(v_adv, i, j - 1));
      P(phib, i, j) -= (1.0 / dt) * b;
      P(phib, i + 1, j) -= b * u_avg * (half * inv_dx); P(phib, i - 1, j) += b * u_avg * (half * inv_dx);
      P(phib, i, j + 1) -= b * v_avg * (half * inv_dy); P(phib, i, j - 1) += b * v_avg * (half * inv_dy);
      P(zp, i, j) -= visc * P(kappa, i, j) * b;
    }
  og_halo_p_T(g, phib);
  /* phi = Pi A_p^+ Pi z_p is symmetric: the same solve on the interior of phib */
  og_mask_owned(g, NULL, NULL, phib);
  pressure_solve_proj(g, phs, phib, rho_u, rho_v, ptol, pmax);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) P(zp, i, j) += P(phs, i, j);
  og_mask_owned(g, zu, zv, zp);
  free(rub); free(rvb); free(rpb); free(zuwb); free(zvwb); free(phib); free(phs);
}

/* ------------------------------------------------------------------------ */
/* Gradients of the bilinear form lam^T A(coef; X) with respect to its coefficient fields,
 * accumulated with the given sign (s = -1 for the Lagrangian term -lam^T A X).  lam: owned
 * rows; X: the padded operand as the matvec saw it (BCs and halos applied).  Outputs are
 * padded arrays of the coefficient's kind: g_rho_u (u), g_rho_v (v), g_rho_c (p),
 * g_kappa (p), g_zeta (w), g_eta (w), and, for the advecting velocity, g_uadv (u), g_vadv (v). */
void og_operator_coef_grad(const og_grid* g, const double* Lu, const double* Lv,
                           const double* Xu, const double* Xv,
                           const double* u_adv, const double* v_adv, const double* rho_c,
                           const double* rho_u, const double* rho_v, const double* kappa,
                           const double* zeta, const double* eta, double dt, int include_adv, double s,
                           double* g_rho_c, double* g_rho_u, double* g_rho_v, double* g_kappa,
                           double* g_zeta, double* g_eta, double* g_uadv, double* g_vadv) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, half = 0.5;
  const int i_u_max = g->per_x ? nx : nx - 1;
  const int j_v_max = g->per_y ? ny : ny - 1;
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
    for (int i = 1; i <= i_u_max; ++i) {
      const double a = s * U(Lu, i, j);
      if (g_rho_u) U(g_rho_u, i, j) += a * U(Xu, i, j) / dt;
      TXX(i + 1, j) -= U(Lu, i, j) * inv_dx; TXX(i, j) += U(Lu, i, j) * inv_dx;
      TXY(i, j) -= U(Lu, i, j) * inv_dy; TXY(i, j - 1) += U(Lu, i, j) * inv_dy;
      if (include_adv) {
        const double uE = half * (U(Xu, i, j) + U(Xu, i + 1, j));
        const double uW = half * (U(Xu, i - 1, j) + U(Xu, i, j));
        const double uN = half * (U(Xu, i, j) + U(Xu, i, j + 1));
        const double uS = half * (U(Xu, i, j - 1) + U(Xu, i, j));
        const double aE = half * (U(u_adv, i, j) + U(u_adv, i + 1, j));
        const double aW = half * (U(u_adv, i - 1, j) + U(u_adv, i, j));
        const double aN = half * (V(v_adv, i, j) + V(v_adv, i + 1, j));
        const double aS = half * (V(v_adv, i, j - 1) + V(v_adv, i + 1, j - 1));
        const double rN = half * (V(rho_v, i, j) + V(rho_v, i + 1, j));
        const double rS = half * (V(rho_v, i, j - 1) + V(rho_v, i + 1, j - 1));
        if (g_rho_c) {
          P(g_rho_c, i + 1, j) += a * aE * uE * inv_dx;
          P(g_rho_c, i, j) -= a * aW * uW * inv_dx;
        }
        if (g_rho_v) {
          V(g_rho_v, i, j) += a * half * aN * uN * inv_dy; V(g_rho_v, i + 1, j) += a * half * aN * uN * inv_dy;
          V(g_rho_v, i, j - 1) -= a * half * aS * uS * inv_dy; V(g_rho_v, i + 1, j - 1) -= a * half * aS * uS * inv_dy;
        }
        if (g_uadv) {
          const double e = a * P(rho_c, i + 1, j) * half * uE * inv_dx;
          const double w = a * P(rho_c, i, j) * half * uW * inv_dx;
          U(g_uadv, i, j) += e - w; U(g_uadv, i + 1, j) += e; U(g_uadv, i - 1, j) -= w;
        }
        if (g_vadv) {
          const double n = a * rN * half * uN * inv_dy;
          const double so = a * rS * half * uS * inv_dy;
          V(g_vadv, i, j) += n; V(g_vadv, i + 1, j) += n;
          V(g_vadv, i, j - 1) -= so; V(g_vadv, i + 1, j - 1) -= so;
        }
      }
    }
  for (int j = 1; j <= j_v_max; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double a = s * V(Lv, i, j);
      if (g_rho_v) V(g_rho_v, i, j) += a * V(Xv, i, j) / dt;
      TYX(i, j) -= V(Lv, i, j) * inv_dx; TYX(i - 1, j) += V(Lv, i, j) * inv_dx;
      TYY(i, j + 1) -= V(Lv, i, j) * inv_dy; TYY(i, j) += V(Lv, i, j) * inv_dy;
      if (include_adv) {
        const double vE = half * (V(Xv, i, j) + V(Xv, i + 1, j));
        const double vW = half * (V(Xv, i - 1, j) + V(Xv, i, j));
        const double vN = half * (V(Xv, i, j) + V(Xv, i, j + 1));
        const double vS = half * (V(Xv, i, j - 1) + V(Xv, i, j));
        const double aE = half * (U(u_adv, i, j) + U(u_adv, i, j + 1));
        const double aW = half * (U(u_adv, i - 1, j) + U(u_adv, i - 1, j + 1));
        const double aN = half * (V(v_adv, i, j) + V(v_adv, i, j + 1));
        const double aS = half * (V(v_adv, i, j - 1) + V(v_adv, i, j));
        const double rE = half * (U(rho_u, i, j) + U(rho_u, i, j + 1));
        const double rW = half * (U(rho_u, i - 1, j) + U(rho_u, i - 1, j + 1));
        if (g_rho_u) {
          U(g_rho_u, i, j) += a * half * aE * vE * inv_dx; U(g_rho_u, i, j + 1) += a * half * aE * vE * inv_dx;
          U(g_rho_u, i - 1, j) -= a * half * aW * vW * inv_dx; U(g_rho_u, i - 1, j + 1) -= a * half * aW * vW * inv_dx;
        }
        if (g_rho_c) {
          P(g_rho_c, i, j + 1) += a * aN * vN * inv_dy;
          P(g_rho_c, i, j) -= a * aS * vS * inv_dy;
        }
        if (g_uadv) {
          const double e = a * rE * half * vE * inv_dx;
          const double w = a * rW * half * vW * inv_dx;
          U(g_uadv, i, j) += e; U(g_uadv, i, j + 1) += e;
          U(g_uadv, i - 1, j) -= w; U(g_uadv, i - 1, j + 1) -= w;
        }
        if (g_vadv) {
          const double n = a * P(rho_c, i, j + 1) * half * vN * inv_dy;
          const double so = a * P(rho_c, i, j) * half * vS * inv_dy;
          V(g_vadv, i, j) += n - so; V(g_vadv, i, j + 1) += n; V(g_vadv, i, j - 1) -= so;
        }
      }
    }
  /* viscous coefficients */
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx + 1; ++i)
      if (g_kappa) P(g_kappa, i, j) += s * TXX(i, j) * (U(Xu, i, j) - U(Xu, i - 1, j)) * inv_dx;
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 1; i <= nx; ++i)
      if (g_kappa) P(g_kappa, i, j) += s * TYY(i, j) * (V(Xv, i, j) - V(Xv, i, j - 1)) * inv_dy;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double du = (U(Xu, i, j + 1) - U(Xu, i, j)) * inv_dy;
      const double dv = (V(Xv, i + 1, j) - V(Xv, i, j)) * inv_dx;
      if (g_zeta) W(g_zeta, i, j) += s * (TXY(i, j) * du + TYX(i, j) * dv);
      if (g_eta) W(g_eta, i, j) += s * (TXY(i, j) - TYX(i, j)) * (dv - du);
    }
#undef TXX
#undef TYY
#undef TXY
#undef TYX
  free(Txx); free(Tyy); free(Txy); free(Tyx);
}

/* ------------------------------------------------------------------------ */
/* Per-cell property path (configs with per-inclusion properties): mu_cell, rho_cell over the
 * padded cells -> the operator coefficients with the edge conventions of
 * src/mod_viscosity.f90:54-151.  The vertex viscosity is the harmonic mean of the four cell
 * viscosities, 4 / sum(1/mu_k), which equals harmonic_viscosity(mean C) when
 * mu_k = harmonic_viscosity(C_k). */
void og_cell_properties(const og_grid* g, const double* mu_cell, const double* rho_cell, double* rho_c,
                        double* rho_u, double* rho_v, double* kappa, double* zeta, double* eta) {
  DECL;
  for (int j = 1 - nh; j <= ny + nh; ++j)
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      P(rho_c, i, j) = P(rho_cell, i, j);
      P(kappa, i, j) = 2.0 * P(mu_cell, i, j);
    }
  for (int j = 1 - nh; j <= ny + nh; ++j) {
    for (int i = 1 - nh; i <= nx + nh - 1; ++i) U(rho_u, i, j) = 0.5 * (P(rho_c, i, j) + P(rho_c, i + 1, j));
    U(rho_u, -nh, j) = U(rho_u, 1 - nh, j);
    U(rho_u, nx + nh, j) = U(rho_u, nx + nh - 1, j);
  }
  for (int i = 1 - nh; i <= nx + nh; ++i) {
    for (int j = 1 - nh; j <= ny + nh - 1; ++j) V(rho_v, i, j) = 0.5 * (P(rho_c, i, j) + P(rho_c, i, j + 1));
    V(rho_v, i, -nh) = V(rho_v, i, 1 - nh);
    V(rho_v, i, ny + nh) = V(rho_v, i, ny + nh - 1);
  }
  for (int j = 1 - nh; j <= ny + nh - 1; ++j)
    for (int i = 1 - nh; i <= nx + nh - 1; ++i) {
      const double sinv = 1.0 / P(mu_cell, i, j) + 1.0 / P(mu_cell, i + 1, j) + 1.0 / P(mu_cell, i, j + 1) +
                          1.0 / P(mu_cell, i + 1, j + 1);
      const double mv = 4.0 / sinv;
      W(zeta, i, j) = 2.0 * mv;
      W(eta, i, j) = mv;
    }
  for (int j = 1 - nh; j <= ny + nh - 1; ++j) {
    W(zeta, -nh, j) = W(zeta, 1 - nh, j); W(zeta, nx + nh, j) = W(zeta, nx + nh - 1, j);
    W(eta, -nh, j) = W(eta, 1 - nh, j); W(eta, nx + nh, j) = W(eta, nx + nh - 1, j);
  }
  for (int i = -nh; i <= nx + nh; ++i) {
    W(zeta, i, -nh) = W(zeta, i, 1 - nh); W(zeta, i, ny + nh) = W(zeta, i, ny + nh - 1);
    W(eta, i, -nh) = W(eta, i, 1 - nh); W(eta, i, ny + nh) = W(eta, i, ny + nh - 1);
  }
}

/* Transpose of og_cell_properties' Jacobian at mu_cell: coefficient gradients (consumed:
 * the edge-copy transposes modify them) -> g_mu, g_rho (accumulated). */
void og_cell_properties_T(const og_grid* g, const double* mu_cell, double* g_rho_c, double* g_rho_u,
                          double* g_rho_v, double* g_kappa, double* g_zeta, double* g_eta,
                          double* g_mu, double* g_rho) {
  DECL;
  for (int i = -nh; i <= nx + nh; ++i) {
    W(g_zeta, i, 1 - nh) += W(g_zeta, i, -nh); W(g_zeta, i, -nh) = 0.0;
    W(g_zeta, i, ny + nh - 1) += W(g_zeta, i, ny + nh); W(g_zeta, i, ny + nh) = 0.0;
    W(g_eta, i, 1 - nh) += W(g_eta, i, -nh); W(g_eta, i, -nh) = 0.0;
    W(g_eta, i, ny + nh - 1) += W(g_eta, i, ny + nh); W(g_eta, i, ny + nh) = 0.0;
  }
  for (int j = 1 - nh; j <= ny + nh - 1; ++j) {
    W(g_zeta, 1 - nh, j) += W(g_zeta, -nh, j); W(g_zeta, -nh, j) = 0.0;
    W(g_zeta, nx + nh - 1, j) += W(g_zeta, nx + nh, j); W(g_zeta, nx + nh, j) = 0.0;
    W(g_eta, 1 - nh, j) += W(g_eta, -nh, j); W(g_eta, -nh, j) = 0.0;
    W(g_eta, nx + nh - 1, j) += W(g_eta, nx + nh, j); W(g_eta, nx + nh, j) = 0.0;
  }
  for (int j = 1 - nh; j <= ny + nh - 1; ++j)
    for (int i = 1 - nh; i <= nx + nh - 1; ++i) {
      const double m[4] = {P(mu_cell, i, j), P(mu_cell, i + 1, j), P(mu_cell, i, j + 1), P(mu_cell, i + 1, j + 1)};
      const double sinv = 1.0 / m[0] + 1.0 / m[1] + 1.0 / m[2] + 1.0 / m[3];
      const double mv = 4.0 / sinv;
      const double gmv = 2.0 * W(g_zeta, i, j) + W(g_eta, i, j);
      const double f = gmv * mv * mv / 4.0;
      P(g_mu, i, j) += f / (m[0] * m[0]);
      P(g_mu, i + 1, j) += f / (m[1] * m[1]);
      P(g_mu, i, j + 1) += f / (m[2] * m[2]);
      P(g_mu, i + 1, j + 1) += f / (m[3] * m[3]);
    }
  for (int i = 1 - nh; i <= nx + nh; ++i) {
    V(g_rho_v, i, 1 - nh) += V(g_rho_v, i, -nh); V(g_rho_v, i, -nh) = 0.0;
    V(g_rho_v, i, ny + nh - 1) += V(g_rho_v, i, ny + nh); V(g_rho_v, i, ny + nh) = 0.0;
    for (int j = 1 - nh; j <= ny + nh - 1; ++j) {
      P(g_rho_c, i, j) += 0.5 * V(g_rho_v, i, j);
      P(g_rho_c, i, j + 1) += 0.5 * V(g_rho_v, i, j);
    }
  }
  for (int j = 1 - nh; j <= ny + nh; ++j) {
    U(g_rho_u, 1 - nh, j) += U(g_rho_u, -nh, j); U(g_rho_u, -nh, j) = 0.0;
    U(g_rho_u, nx + nh - 1, j) += U(g_rho_u, nx + nh, j); U(g_rho_u, nx + nh, j) = 0.0;
    for (int i = 1 - nh; i <= nx + nh - 1; ++i) {
      P(g_rho_c, i, j) += 0.5 * U(g_rho_u, i, j);
      P(g_rho_c, i + 1, j) += 0.5 * U(g_rho_u, i, j);
    }
  }
  for (int j = 1 - nh; j <= ny + nh; ++j)
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      P(g_rho, i, j) += P(g_rho_c, i, j);
      P(g_mu, i, j) += 2.0 * P(g_kappa, i, j);
    }
}

/* E: owned X -> the padded operand the matvec sees (mask, velocity BCs with lid 0, halos). */
void og_extend(const og_grid* g, double* Xu, double* Xv, double* Xp) {
  og_mask_owned(g, Xu, Xv, Xp);
  og_velocity_bcs(g, Xu, Xv, 0.0);
  og_halo_u(g, Xu);
  og_halo_v(g, Xv);
  og_halo_p(g, Xp);
}

/* E^T: adjoint of og_extend (consumes the padded input, leaves owned values). */
void og_extend_T(const og_grid* g, double* Xu, double* Xv, double* Xp) {
  og_halo_p_T(g, Xp);
  og_halo_v_T(g, Xv);
  og_halo_u_T(g, Xu);
  og_velocity_bcs_T(g, Xu, Xv);
  og_mask_owned(g, Xu, Xv, Xp);
}

/* ------------------------------------------------------------------------ */
/* Transposes of the interface chain.  Each is the reverse mode of the forward routine at the
 * given state: the forward statements are revisited in reverse order and every
 * "dst = f(src)" becomes "src_bar += f'(src) dst_bar".  Kinks take one side: the clamp to [0, 1]
 * passes the derivative only strictly inside (0, 1), the limiters their one-sided slope at
 * r (van Leer: 2 / (1 + r)^2 for r > 0, else 0; Superbee: the slope of the branch the forward
 * select takes), the upwind switch the branch the forward takes (d|u|/du = 1 at u = 0), and a
 * face with |den| <= 1e-12 (r set to 0, psi(0) den = 0) contributes only its upwind cell. */

static double clamp01_d(double x) { return (x > 0.0 && x < 1.0) ? 1.0 : 0.0; }
static double clamp01_v(double x) {   /* Fortran max(0, min(1, x)), as og_vof_advection's clamp */
  const double m = (x < 1.0) ? x : 1.0;
  return (m > 0.0) ? m : 0.0;
}
static double lim_psi(double r, int lim) {
  if (lim == 1) {
    const double a = (1.0 < 2.0 * r) ? 1.0 : 2.0 * r, b = (2.0 < r) ? 2.0 : r, m = (b > a) ? b : a;
    return (m > 0.0) ? m : 0.0;
  }
  return (r <= 0.0) ? 0.0 : (2.0 * r) / (1.0 + r);
}
static double lim_dpsi(double r, int lim) {
  if (lim == 1) {
    const double a = (1.0 < 2.0 * r) ? 1.0 : 2.0 * r, b = (2.0 < r) ? 2.0 : r, m = (b > a) ? b : a;
    if (!(m > 0.0)) return 0.0;
    if (b > a) return (2.0 < r) ? 0.0 : 1.0;
    return (1.0 < 2.0 * r) ? 0.0 : 2.0;
  }
  return (r <= 0.0) ? 0.0 : 2.0 / ((1.0 + r) * (1.0 + r));
}

/* One face of og_vof_advection: c0, c1, c2 are the stencil (uf >= 0: c_{i-1}, c_i, c_{i+1};
 * uf < 0: c_i, c_{i+1}, c_{i+2}), dt and 1 / h (the Courant number formed as the forward forms
 * it).  Returns clamp01(cf); adds fbar dF/duf to *ubar and fbar dF/dstencil to d[0..2]. */
static double vof_face_T(double uf, double c0, double c1, double c2, double dt, double inv_h, int lim,
                         double fbar, double* ubar, double d[3]) {
  const double k = dt * inv_h;
  double num, den, up, sgn;
  if (uf >= 0.0) { num = c1 - c0; den = c2 - c1; up = c1; sgn = 1.0; }
  else { num = c2 - c1; den = c1 - c0; up = c1; sgn = -1.0; }
  const int big = fabs(den) > 1.0e-12;
  const double rr = big ? num / den : 0.0;
  const double cfl = fabs(uf) * dt * inv_h;
  const double A = 0.5 * (1.0 - cfl);
  const double psi = lim_psi(rr, lim);
  const double cf = up + sgn * A * psi * den;
  const double cl = clamp01_v(cf);
  const double dcl = clamp01_d(cf);
  /* F = uf clamp(cf); dcf/duf = -0.5 k psi den (both branches) */
  *ubar += fbar * (cl + uf * dcl * (-0.5 * k * psi * den));
  const double cfb = fbar * uf * dcl;
  const double pn = big ? A * lim_dpsi(rr, lim) : 0.0;              /* d(psi den)/d num */
  const double pd = big ? A * (psi - rr * lim_dpsi(rr, lim)) : 0.0; /* d(psi den)/d den */
  if (uf >= 0.0) {  /* cf = c1 + A psi den, num = c1 - c0, den = c2 - c1 */
    d[0] += cfb * (-pn);
    d[1] += cfb * (1.0 + pn - pd);
    d[2] += cfb * pd;
  } else {          /* cf = c1 - A psi den, num = c2 - c1, den = c1 - c0 */
    d[0] += cfb * pd;
    d[1] += cfb * (1.0 + pn - pd);
    d[2] += cfb * (-pn);
  }
  return cl;
}

/* Reverse mode of og_vof_advection (src/mod_vof.f90:38-140) at (c_in, u, v). */
void og_vof_advection_T(const og_grid* g, const double* c_in, const double* u, const double* v, double dt,
                        const double* cbar_out, double* cbar_in, double* ubar, double* vbar) {
  DECL;
  const int lim = og_get_vof_limiter();
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, kx = dt * inv_dx, ky = dt * inv_dy;
  const size_t sp = og_size_p(g);
  double* cb = zalloc(sp);
  double* tb = zalloc(sp);
  memcpy(cb, cbar_out, sp * sizeof(double));
  og_halo_p_T(g, cb);                              /* og_halo_p(c_out) */
  /* forward fluxes (for the unclamped updates) */
  const int lfx = nx + 1;
  double* Fx = zalloc((size_t)(nx + 1) * ny);
  double* Fy = zalloc((size_t)nx * (ny + 1));
#define FX(i, j) Fx[(size_t)(i) + (size_t)((j) - 1) * lfx]
#define FY(i, j) Fy[(size_t)((i) - 1) + (size_t)(j) * nx]
  double dummy, dd[3];
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double uf = U(u, i, j);
      const double cl = (uf >= 0.0) ? vof_face_T(uf, P(c_in, i - 1, j), P(c_in, i, j), P(c_in, i + 1, j), dt, inv_dx, lim, 0.0, &dummy, dd)
                                    : vof_face_T(uf, P(c_in, i, j), P(c_in, i + 1, j), P(c_in, i + 2, j), dt, inv_dx, lim, 0.0, &dummy, dd);
      FX(i, j) = uf * cl;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double vf = V(v, i, j);
      const double cl = (vf >= 0.0) ? vof_face_T(vf, P(c_in, i, j - 1), P(c_in, i, j), P(c_in, i, j + 1), dt, inv_dy, lim, 0.0, &dummy, dd)
                                    : vof_face_T(vf, P(c_in, i, j), P(c_in, i, j + 1), P(c_in, i, j + 2), dt, inv_dy, lim, 0.0, &dummy, dd);
      FY(i, j) = vf * cl;
    }
  /* c_out = c_in outside the interior (the copy); interior: clamp01(c_in - dt div F) */
  for (int j = 1 - nh; j <= ny + nh; ++j)
    for (int i = 1 - nh; i <= nx + nh; ++i) {
      if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
        const double pre = P(c_in, i, j) - dt * ((FX(i, j) - FX(i - 1, j)) * inv_dx + (FY(i, j) - FY(i, j - 1)) * inv_dy);
        P(tb, i, j) = P(cb, i, j) * clamp01_d(pre);
        P(cbar_in, i, j) += P(tb, i, j);
      } else {
        P(cbar_in, i, j) += P(cb, i, j);
      }
    }
  /* fluxes: Fbar_x(i, j) = kx (tb(i+1, j) - tb(i, j)); then each face's stencil */
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double fb = kx * (((i + 1 <= nx) ? P(tb, i + 1, j) : 0.0) - ((i >= 1) ? P(tb, i, j) : 0.0));
      const double uf = U(u, i, j);
      double d[3] = {0.0, 0.0, 0.0};
      double ub = 0.0;
      const int b = (uf >= 0.0) ? i - 1 : i;
      vof_face_T(uf, P(c_in, b, j), P(c_in, b + 1, j), P(c_in, b + 2, j), dt, inv_dx, lim, fb, &ub, d);
      U(ubar, i, j) += ub;
      for (int q = 0; q < 3; ++q) P(cbar_in, b + q, j) += d[q];
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double fb = ky * (((j + 1 <= ny) ? P(tb, i, j + 1) : 0.0) - ((j >= 1) ? P(tb, i, j) : 0.0));
      const double vf = V(v, i, j);
      double d[3] = {0.0, 0.0, 0.0};
      double vb = 0.0;
      const int b = (vf >= 0.0) ? j - 1 : j;
      vof_face_T(vf, P(c_in, i, b), P(c_in, i, b + 1), P(c_in, i, b + 2), dt, inv_dy, lim, fb, &vb, d);
      V(vbar, i, j) += vb;
      for (int q = 0; q < 3; ++q) P(cbar_in, i, b + q) += d[q];
    }
#undef FX
#undef FY
  free(Fx); free(Fy); free(cb); free(tb);
}

/* Transpose of the Neumann mirror of og_svof (the copies reversed: y pass, then x pass). */
static void mirror_p_T(const og_grid* g, double* a) {
  DECL;
  if (!g->per_y) {
    for (int i = nx + nh; i >= 1 - nh; --i)
      for (int j = nh; j >= 1; --j) { P(a, i, ny - j + 1) += P(a, i, ny + j); P(a, i, ny + j) = 0.0; }
    for (int i = nx + nh; i >= 1 - nh; --i)
      for (int j = nh; j >= 1; --j) { P(a, i, j) += P(a, i, 1 - j); P(a, i, 1 - j) = 0.0; }
  }
  if (!g->per_x) {
    for (int j = ny + nh; j >= 1 - nh; --j)
      for (int i = nh; i >= 1; --i) { P(a, nx - i + 1, j) += P(a, nx + i, j); P(a, nx + i, j) = 0.0; }
    for (int j = ny + nh; j >= 1 - nh; --j)
      for (int i = nh; i >= 1; --i) { P(a, i, j) += P(a, 1 - i, j); P(a, 1 - i, j) = 0.0; }
  }
}

/* Reverse mode of og_svof (src/mod_vof.f90:145-194) with its sweep count fixed at `iters` (the
 * count the forward call returned): the sweeps are linear in c_raw.  csbar: adjoint of the whole
 * padded output; crawbar accumulates on the interior (the forward overwrites every ghost). */
void og_svof_T(const og_grid* g, const double* csbar, double delta, int iters, double* crawbar) {
  DECL;
  const double cx = (delta / g->dx) * (delta / g->dx);
  const double cy = (delta / g->dy) * (delta / g->dy);
  const double diag = 1.0 + 2.0 * (cx + cy);
  const size_t sp = og_size_p(g);
  double* sb = zalloc(sp);
  double* ob = zalloc(sp);
  memcpy(sb, csbar, sp * sizeof(double));
  mirror_p_T(g, sb);   /* the final halo + mirror */
  og_halo_p_T(g, sb);
  for (int it = iters; it >= 1; --it) {
    /* c_smooth(int) = (c_raw + cx (old E + old W) + cy (old N + old S)) / diag */
    memset(ob, 0, sp * sizeof(double));
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double t = P(sb, i, j) / diag;
        P(crawbar, i, j) += t;
        P(ob, i + 1, j) += cx * t; P(ob, i - 1, j) += cx * t;
        P(ob, i, j + 1) += cy * t; P(ob, i, j - 1) += cy * t;
        P(sb, i, j) = 0.0;
      }
    /* c_old = c_smooth (whole array), after the halo exchange and the mirror */
    for (size_t q = 0; q < sp; ++q) sb[q] += ob[q];
    mirror_p_T(g, sb);
    og_halo_p_T(g, sb);
  }
  /* c_smooth = c_raw (whole array): its ghosts are overwritten before any use */
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) P(crawbar, i, j) += P(sb, i, j);
  free(sb); free(ob);
}

/* Reverse mode of og_curvature (src/mod_surface_tension.f90:25-72) at cs. */
void og_curvature_T(const og_grid* g, const double* cs, const double* kbar, double* csbar) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy, eps = 1.0e-12;
  const size_t sp = og_size_p(g);
  double* kb = zalloc(sp);
  memcpy(kb, kbar, sp * sizeof(double));
  og_halo_p_T(g, kb);   /* og_halo_p(kc); other ghosts stay zero: no dependence */
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double nb = ((i + 1 <= nx) ? P(kb, i + 1, j) : 0.0) * inv_dx - ((i >= 1) ? P(kb, i, j) : 0.0) * inv_dx;
      const double gx = (P(cs, i + 1, j) - P(cs, i, j)) * inv_dx;
      const double gy = 0.25 * (P(cs, i, j + 1) + P(cs, i + 1, j + 1) - P(cs, i, j - 1) - P(cs, i + 1, j - 1)) * inv_dy;
      const double n2 = gx * gx + gy * gy + eps, ng = sqrt(n2), n3 = n2 * ng;
      const double gxb = nb * (gy * gy + eps) / n3, gyb = -nb * gx * gy / n3;
      P(csbar, i + 1, j) += gxb * inv_dx; P(csbar, i, j) -= gxb * inv_dx;
      const double t = 0.25 * gyb * inv_dy;
      P(csbar, i, j + 1) += t; P(csbar, i + 1, j + 1) += t; P(csbar, i, j - 1) -= t; P(csbar, i + 1, j - 1) -= t;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double nb = ((j + 1 <= ny) ? P(kb, i, j + 1) : 0.0) * inv_dy - ((j >= 1) ? P(kb, i, j) : 0.0) * inv_dy;
      const double gx = 0.25 * (P(cs, i + 1, j) + P(cs, i + 1, j + 1) - P(cs, i - 1, j) - P(cs, i - 1, j + 1)) * inv_dx;
      const double gy = (P(cs, i, j + 1) - P(cs, i, j)) * inv_dy;
      const double n2 = gx * gx + gy * gy + eps, ng = sqrt(n2), n3 = n2 * ng;
      const double gyb = nb * (gx * gx + eps) / n3, gxb = -nb * gx * gy / n3;
      P(csbar, i, j + 1) += gyb * inv_dy; P(csbar, i, j) -= gyb * inv_dy;
      const double t = 0.25 * gxb * inv_dx;
      P(csbar, i + 1, j) += t; P(csbar, i + 1, j + 1) += t; P(csbar, i - 1, j) -= t; P(csbar, i - 1, j + 1) -= t;
    }
  free(kb);
}

/* Reverse mode of og_csf (src/mod_surface_tension.f90:78-109) at (c_raw, kc). */
void og_csf_T(const og_grid* g, const double* c_raw, const double* kc, double sigma, const double* fubar,
              const double* fvbar, double* cbar, double* kbar) {
  DECL;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const double fb = U(fubar, i, j);
      const double kf = 0.5 * (P(kc, i, j) + P(kc, i + 1, j));
      const double d = P(c_raw, i + 1, j) - P(c_raw, i, j);
      P(cbar, i + 1, j) += sigma * kf * inv_dx * fb; P(cbar, i, j) -= sigma * kf * inv_dx * fb;
      P(kbar, i, j) += 0.5 * sigma * d * inv_dx * fb; P(kbar, i + 1, j) += 0.5 * sigma * d * inv_dx * fb;
    }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double fb = V(fvbar, i, j);
      const double kf = 0.5 * (P(kc, i, j) + P(kc, i, j + 1));
      const double d = P(c_raw, i, j + 1) - P(c_raw, i, j);
      P(cbar, i, j + 1) += sigma * kf * inv_dy * fb; P(cbar, i, j) -= sigma * kf * inv_dy * fb;
      P(kbar, i, j) += 0.5 * sigma * d * inv_dy * fb; P(kbar, i, j + 1) += 0.5 * sigma * d * inv_dy * fb;
    }
}
