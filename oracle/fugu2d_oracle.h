/*
 * fugu2d_oracle.h -- CPU restatement of the FUGU-2D per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libfugu2d_hip.so) never calls into it.
 *
 * Every routine restates one Fortran procedure of the reference line by line,
 * in IEEE FP64 with the same operation order, single-rank MPI semantics
 * (coords = (0,0), dims = (1,1); periodic halo exchange = self wrap-copy,
 * non-periodic edges = MPI_PROC_NULL i.e. ghosts untouched).
 *
 * Array layout: Fortran column-major padded arrays, passed as a pointer to the
 * first element (lowest bounds), i fastest:
 *   u-type  (0-nh:nx+nh, 1-nh:ny+nh)   ld = nx+1+2nh, rows = ny+2nh
 *   v-type  (1-nh:nx+nh, 0-nh:ny+nh)   ld = nx+2nh,   rows = ny+1+2nh
 *   p-type  (1-nh:nx+nh, 1-nh:ny+nh)   ld = nx+2nh,   rows = ny+2nh
 *   w-type  (0-nh:nx+nh, 0-nh:ny+nh)   ld = nx+1+2nh, rows = ny+1+2nh (vertices)
 */
#ifndef FUGU2D_ORACLE_H
#define FUGU2D_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nh;
  double dx, dy;
  int per_x, per_y;
} og_grid;

typedef struct {
  double rho1, rho2, mu1, mu2, sigma, gx, gy;
  int time_order;
  double alpha_penalty;
  double cfl_target;
} og_phys;

/* Reduction summation: 1 = compensated (correctly rounded, default), 0 = plain sequential
 * in the reference's loop order. */
void og_set_sum_mode(int mode);
/* The PCD's inner PCG tolerance and cap (default 1e-5 / 150, the reference's). */
void og_set_pcd_pcg(double tol, int max_iter);
/* VOF face limiter: 0 van Leer (the reference, default), 1 Superbee (opt-in). */
void og_set_vof_limiter(int lim);
double og_superbee(double r);
int og_get_sum_mode(void);

size_t og_size_u(const og_grid* g);
size_t og_size_v(const og_grid* g);
size_t og_size_p(const og_grid* g);
size_t og_size_w(const og_grid* g);

/* src/mod_mpi.f90:113-245, single rank */
void og_halo_p(const og_grid* g, double* a);
void og_halo_u(const og_grid* g, double* a);
void og_halo_v(const og_grid* g, double* a);

/* src/mod_operators.f90 */
void og_divergence(const og_grid* g, const double* u, const double* v, double* div);
void og_grad_pressure(const og_grid* g, const double* p, double* gu, double* gv);
void og_viscous_stress(const og_grid* g, const double* u, const double* v,
                       const double* kappa, const double* zeta, const double* eta,
                       double* visc_u, double* visc_v);
void og_advection(const og_grid* g, const double* u_adv, const double* v_adv,
                  const double* rho_c, const double* rho_u, const double* rho_v,
                  const double* u, const double* v, double* adv_u, double* adv_v,
                  const double* mdot_u, const double* mdot_v);
void og_velocity_bcs(const og_grid* g, double* u, double* v, double u_lid);
void og_matvec_coupled(const og_grid* g, double* Xu, double* Xv, double* Xp,
                       double* Yu, double* Yv, double* Yp,
                       const double* u_adv, const double* v_adv,
                       const double* rho_c, const double* rho_u, const double* rho_v,
                       const double* kappa, const double* zeta, const double* eta,
                       double dt, int include_advection,
                       const double* pen_u, const double* pen_v,
                       const double* mdot_u, const double* mdot_v);

/* src/mod_bicgstab2.f90:32-72 */
double og_coupled_dot(const og_grid* g, const double* u1, const double* v1, const double* p1,
                      const double* u2, const double* v2, const double* p2);

/* src/mod_subsolvers.f90 */
void og_poisson_pcg(const og_grid* g, double* phi, const double* rhs,
                    const double* rho_u, const double* rho_v,
                    double tol, int max_iter, int* iter_out, double* res_out);
void og_helmholtz_u(const og_grid* g, double* u_sol, const double* rhs_u,
                    const double* rho_u, const double* kappa, const double* zeta,
                    const double* eta, double dt, int n_sweeps, const double* pen_u);
void og_helmholtz_v(const og_grid* g, double* v_sol, const double* rhs_v,
                    const double* rho_v, const double* kappa, const double* zeta,
                    const double* eta, double dt, int n_sweeps, const double* pen_v);

/* src/mod_pcd_precond.f90:14-104 */
void og_pcd(const og_grid* g, const double* zu, const double* zv, const double* zp,
            double* ru, double* rv, double* rp,
            const double* u_adv, const double* v_adv,
            const double* rho_u, const double* rho_v,
            const double* kappa, const double* zeta, const double* eta, double dt,
            const double* pen_u, const double* pen_v);

/* src/mod_bicgstab2.f90:77-264 and 270-570.  has_lid mirrors present(u_top_lid). */
void og_bicgstab(const og_grid* g, double* Xu, double* Xv, double* Xp,
                 const double* Bu, const double* Bv, const double* Bp,
                 const double* u_adv, const double* v_adv,
                 const double* rho_c, const double* rho_u, const double* rho_v,
                 const double* kappa, const double* zeta, const double* eta, double dt,
                 double tol, int max_iter, int* iter_out, double* res_out,
                 int has_lid, double u_lid,
                 const double* pen_u, const double* pen_v,
                 const double* mdot_u, const double* mdot_v);
void og_bicgstab2(const og_grid* g, double* Xu, double* Xv, double* Xp,
                  const double* Bu, const double* Bv, const double* Bp,
                  const double* u_adv, const double* v_adv,
                  const double* rho_c, const double* rho_u, const double* rho_v,
                  const double* kappa, const double* zeta, const double* eta, double dt,
                  double tol, int max_iter, int* iter_out, double* res_out,
                  int has_lid, double u_lid,
                  const double* pen_u, const double* pen_v,
                  const double* mdot_u, const double* mdot_v);

void og_bicgstab2_ex(const og_grid* g, double* Xu, double* Xv, double* Xp,
                     const double* Bu, const double* Bv, const double* Bp,
                     const double* u_adv, const double* v_adv,
                     const double* rho_c, const double* rho_u, const double* rho_v,
                     const double* kappa, const double* zeta, const double* eta, double dt,
                     double tol, int max_iter, int* iter_out, double* res_out,
                     int has_lid, double u_lid,
                     const double* pen_u, const double* pen_v,
                     const double* mdot_u, const double* mdot_v, int consistent);

/* src/mod_vof.f90 */
double og_van_leer(double r);
void og_vof_advection(const og_grid* g, const double* c_in, double* c_out,
                      const double* u, const double* v, double dt,
                      double* flux_x, double* flux_y);
void og_svof(const og_grid* g, const double* c_raw, double* c_smooth, double delta,
             int max_iter, double tol, int* iter_out);

/* src/mod_surface_tension.f90 */
void og_curvature(const og_grid* g, const double* c_smooth, double* kappa_curv);
void og_csf(const og_grid* g, const double* c_raw, const double* kappa_curv, double sigma,
            double* fu, double* fv);

/* src/mod_viscosity.f90 */
double og_harmonic_viscosity(double c, double mu1, double mu2);
void og_mixture_densities(const og_grid* g, const double* c, double rho1, double rho2,
                          double* rho_c, double* rho_u, double* rho_v);
void og_mixture_viscosities(const og_grid* g, const double* c, double mu1, double mu2,
                            double* kappa, double* zeta, double* eta);

/* src/mod_solver_coupled.f90:41-91 */
double og_adaptive_dt(const og_grid* g, const og_phys* ph, const double* u, const double* v,
                      double dt_max);

/* src/mod_solver_coupled.f90:137-380.  Optional arguments are nullable pointers /
 * flags: has_lid = present(u_top_lid); pen_* = present(phi_s_*). */
typedef struct {
  int use_bicgstab2;
  int has_lid;
  double u_lid;
  double tol;
  int max_iter;
  int use_amp;
  double u_solid, v_solid;
  int consistent;   /* BiCGStab(2) with BCs on r and s (not in the reference; 0 = the reference) */
} og_step_opts;

void og_advance_step(const og_grid* g, const og_phys* ph, const og_step_opts* o,
                     double* u, double* v, double* p, double* c,
                     double* u_prev, double* v_prev,
                     double* rho_c, double* rho_u, double* rho_v,
                     double* kappa, double* zeta, double* eta,
                     double* kappa_curv, double* f_u, double* f_v,
                     double dt, int step_idx,
                     const double* phi_s_u, const double* phi_s_v,
                     int* iter_out, double* res_out);

/* ---- adjoint building blocks (fugu2d_adjoint.c; no counterpart in the reference) ---- */
void og_mask_owned(const og_grid* g, double* u, double* v, double* p);
void og_halo_p_T(const og_grid* g, double* a);
void og_halo_u_T(const og_grid* g, double* a);
void og_halo_v_T(const og_grid* g, double* a);
void og_velocity_bcs_T(const og_grid* g, double* u, double* v);
void og_extend(const og_grid* g, double* Xu, double* Xv, double* Xp);
void og_extend_T(const og_grid* g, double* Xu, double* Xv, double* Xp);
void og_matvec_eff(const og_grid* g, const double* Xu, const double* Xv, const double* Xp,
                   double* Yu, double* Yv, double* Yp, const double* u_adv, const double* v_adv,
                   const double* rho_c, const double* rho_u, const double* rho_v,
                   const double* kappa, const double* zeta, const double* eta, double dt, int include_adv);
void og_matvec_T(const og_grid* g, const double* Lu, const double* Lv, const double* Lp,
                 double* Xu, double* Xv, double* Xp, const double* u_adv, const double* v_adv,
                 const double* rho_c, const double* rho_u, const double* rho_v,
                 const double* kappa, const double* zeta, const double* eta, double dt, int include_adv);
void og_pcd_lin(const og_grid* g, const double* zu, const double* zv, const double* zp,
                double* ru, double* rv, double* rp, const double* u_adv, const double* v_adv,
                const double* rho_u, const double* rho_v, const double* kappa, const double* zeta,
                const double* eta, double dt, double ptol, int pmax, double visc);
void og_pcd_lin_T(const og_grid* g, const double* yu, const double* yv, const double* yp,
                  double* zu, double* zv, double* zp, const double* u_adv, const double* v_adv,
                  const double* rho_u, const double* rho_v, const double* kappa, const double* zeta,
                  const double* eta, double dt, double ptol, int pmax, double visc);
void og_operator_coef_grad(const og_grid* g, const double* Lu, const double* Lv,
                           const double* Xu, const double* Xv,
                           const double* u_adv, const double* v_adv, const double* rho_c,
                           const double* rho_u, const double* rho_v, const double* kappa,
                           const double* zeta, const double* eta, double dt, int include_adv, double s,
                           double* g_rho_c, double* g_rho_u, double* g_rho_v, double* g_kappa,
                           double* g_zeta, double* g_eta, double* g_uadv, double* g_vadv);
void og_cell_properties(const og_grid* g, const double* mu_cell, const double* rho_cell, double* rho_c,
                        double* rho_u, double* rho_v, double* kappa, double* zeta, double* eta);
void og_cell_properties_T(const og_grid* g, const double* mu_cell, double* g_rho_c, double* g_rho_u,
                          double* g_rho_v, double* g_kappa, double* g_zeta, double* g_eta,
                          double* g_mu, double* g_rho);
/* Transposes of the interface chain (reverse mode of og_vof_advection, og_svof, og_curvature and
 * og_csf at a given state; every output accumulates: callers zero it first). */
int og_get_vof_limiter(void);
void og_vof_advection_T(const og_grid* g, const double* c_in, const double* u, const double* v, double dt,
                        const double* cbar_out, double* cbar_in, double* ubar, double* vbar);
void og_svof_T(const og_grid* g, const double* csbar, double delta, int iters, double* crawbar);
void og_curvature_T(const og_grid* g, const double* cs, const double* kbar, double* csbar);
void og_csf_T(const og_grid* g, const double* c_raw, const double* kc, double sigma, const double* fubar,
              const double* fvbar, double* cbar, double* kbar);

#ifdef __cplusplus
}
#endif
#endif
