/*
 * fugu2d.h -- C ABI of the MI355X (gfx950) FUGU-2D per-step hot path (libfugu2d_hip.so).
 *
 * Drop-in boundary for the reference's Fortran modules (mod_operators, mod_bicgstab2,
 * mod_pcd_precond, mod_subsolvers, mod_vof, mod_surface_tension, mod_viscosity,
 * mod_solver_coupled).  Plain pointers and sizes only.
 *
 * Arrays: every field argument is the reference's padded Fortran array, passed as a pointer
 * to its first element (lowest bounds), column-major with i fastest:
 *   u-kind (0-nh:nx+nh, 1-nh:ny+nh)   v-kind (1-nh:nx+nh, 0-nh:ny+nh)
 *   p-kind (1-nh:nx+nh, 1-nh:ny+nh)   w-kind (0-nh:nx+nh, 0-nh:ny+nh) (vertices)
 * Pointers may be host memory or device memory of the context's GPU (hipMemcpyDefault
 * staging): the "fine" entry points copy in, run on the device, copy out and synchronise.
 * The "device-resident" entry points (fugu_state_*, fugu_step, fugu_solve_*) keep all
 * fields in HBM between calls.
 *
 * Optional Fortran arguments map to NULL pointers; logical flags to int.
 * Return value: 0 ok; < 0 error (see fugu_last_error()); > 0 warning bits (FUGU_W_*).
 * A context is not shared between host threads.
 */
#ifndef FUGU2D_H
#define FUGU2D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fugu_ctx fugu_ctx;

/* Grid of one device-resident domain.  Replaces the module globals of src/mod_mpi.f90:10-26
 * (dx_grid, dy_grid, periods, coords, dims) at one rank; nh must be 2 (src/main.f90:22). */
typedef struct {
  int nx, ny, nh;
  double lx, ly;          /* dx = lx/nx, dy = ly/ny  (src/mod_mpi.f90:104-105) */
  int periodic_x, periodic_y;
} fugu_grid;

/* Physical parameters: the mutable globals of src/mod_params.f90:19-50. */
typedef struct {
  double rho1, rho2, mu1, mu2, sigma, gx, gy;
  int time_order;           /* 1 = BDF1, 2 = BDF2 from step 2 on */
  double alpha_penalty;
  double cfl_target;
} fugu_params;

/* Options of advance_coupled_step (src/mod_solver_coupled.f90:179-194). */
typedef struct {
  int use_bicgstab2;        /* use_bicgstab2 */
  int has_lid;              /* present(u_top_lid) */
  double u_lid;
  double tol;               /* tol_in (reference default tol_bicgstab2 = 1e-6) */
  int max_iter;             /* max_iter_in (reference default 500) */
  int use_amp;              /* consistent_amp */
  double u_solid, v_solid;  /* u_solid_in, v_solid_in */
  unsigned solver_flags;    /* FUGU_SOLVE_* / FUGU_PRECOND_* (0 = the reference solver exactly) */
  double inner_tol;         /* refined solve: inner Krylov tolerance (<= 0: 1e-3) */
  int inner_max_iter;       /* refined solve: inner Krylov iteration cap (0: 25; < 0: no inner Krylov,
                               the PCD alone preconditions the outer FGMRES) */
  int mg_cycles;            /* FUGU_PRECOND_MG: V-cycles per preconditioner application (<= 0: 2 from
                               2^22 cells, else 1) */
  int restart;              /* refined solve: outer FGMRES restart length (<= 0: 10 after a refined
                               solve of 16-30 FGMRES steps on the context, else 20 up to 2^21
                               cells and 30 above; max 47) */
} fugu_step_opts;

enum {
  /* BiCGStab(2) with velocity BCs applied to r and s before their matvecs (the reference
   * omits them, src/mod_bicgstab2.f90:430,489, which stalls it on walled domains). */
  FUGU_B2_CONSISTENT_BCS = 1u << 0,
  /* Step solve by restarted flexible GMRES in FP64 on the true residual (coupled_dot norm);
   * the flexible preconditioner is the PCD (inner_max_iter < 0) or an inner BiCGStab(2) + PCD. */
  FUGU_SOLVE_REFINED = 1u << 1,
  /* With FUGU_SOLVE_REFINED: preconditioner (and inner Krylov) stored in FP32, the matvec
   * arithmetic in FP64 (mixed precision); the outer FGMRES stays FP64. */
  FUGU_SOLVE_MIXED = 1u << 2,
  /* Pressure block of the PCD preconditioner by geometric multigrid V-cycles (a fixed linear
   * operator) instead of the reference's Jacobi-PCG to 1e-5 / 150 iterations. */
  FUGU_PRECOND_MG = 1u << 3,
  /* Schur approximation of the PCD also carries the viscous term: r_p -= kappa * z_p. */
  FUGU_PRECOND_VISC = 1u << 4,
  /* Velocity blocks of the PCD preconditioner by one geometric-multigrid V-cycle each (the
   * same Helmholtz operators) instead of the reference's three damped-Jacobi sweeps. */
  FUGU_PRECOND_VMG = 1u << 5,
  /* Initial guess of the step solve extrapolated in time, X0 = 2 X^n - X^(n-1) (u and v are the
   * AB2 advection velocities; p^(n-1) is kept by the context), instead of X^n.  From the second
   * consecutive step with the flag on; fugu_state_upload forgets p^(n-1). */
  FUGU_SOLVE_EXTRAPOLATE = 1u << 6
};

enum { FUGU_W_MAX_ITER = 1, FUGU_W_NONFINITE = 2 };
enum { FUGU_KIND_U = 0, FUGU_KIND_V = 1, FUGU_KIND_P = 2, FUGU_KIND_W = 3 };

/* ---- context -------------------------------------------------------------------- */
int fugu_ctx_create(const fugu_grid* grid, const fugu_params* params, fugu_ctx** out);
int fugu_ctx_destroy(fugu_ctx* ctx);
int fugu_set_params(fugu_ctx* ctx, const fugu_params* params);
const char* fugu_last_error(void);
/* Element count of a padded array of the given kind. */
long long fugu_field_size(const fugu_ctx* ctx, int kind);

/* Summation order of the reference-algorithm reductions: coupled_dot
 * (src/mod_bicgstab2.f90:32-72) and the PCG's means and dots (src/mod_subsolvers.f90:91-233).
 * FUGU_SUM_REFERENCE (the default of a new context): the reference's own order at one rank --
 * each sum adds its terms one by one from zero, j outer, i inner, u faces then v faces then
 * cells -- so results are bit-identical to the reference's arithmetic.  The terms are computed
 * in parallel; the additions are one dependent chain per sum (about one FP64 add latency per
 * owned point), so this order suits parity work and grids up to a few 10^5 cells.
 * FUGU_SUM_TREE: per-workgroup double-double partials folded in a fixed order -- fast, and within
 * a few units of the last place of the reference's sums, but not their exact roundings (the
 * Krylov trajectories then drift from the reference's by up to ~1e-7 relative over a step).
 * The chosen fast forms (FUGU_SOLVE_REFINED) and the adjoint always sum in the tree order.
 * No counterpart in the reference (it has one order). */
enum { FUGU_SUM_REFERENCE = 0, FUGU_SUM_TREE = 1 };
int fugu_set_sum_order(fugu_ctx* ctx, int order);
/* Face limiter of the VOF-TVD advection (fugu_vof_advection_tvd and the step).
 * FUGU_LIMITER_VAN_LEER (default): the reference's psi(r) = 2r/(1+r), r > 0 (src/mod_vof.f90:25-33).
 * FUGU_LIMITER_SUPERBEE: psi(r) = max(0, min(2r, 1), min(r, 2)), an opt-in sharper limiter (the
 * reference's README names Superbee; its code uses van Leer).  No counterpart in the reference. */
enum { FUGU_LIMITER_VAN_LEER = 0, FUGU_LIMITER_SUPERBEE = 1 };
int fugu_set_vof_limiter(fugu_ctx* ctx, int limiter);
int fugu_get_sum_order(const fugu_ctx* ctx);

/* ---- fine API: one entry per reference public procedure ----------------------------- */
/* src/mod_mpi.f90:113-245 at one rank (periodic self wrap-copy; walls untouched). */
int fugu_exchange_halos(fugu_ctx* ctx, double* a, int kind);
/* src/mod_operators.f90:34-53 */
int fugu_op_divergence(fugu_ctx* ctx, const double* u, const double* v, double* div);
/* src/mod_operators.f90:254-333; u_top_lid NULL = absent */
int fugu_apply_velocity_bcs(fugu_ctx* ctx, double* u, double* v, const double* u_top_lid);
/* src/mod_operators.f90:341-443.  X is in/out (periodic halo exchange writes its ghosts). */
int fugu_matvec_coupled(fugu_ctx* ctx, double* X_u, double* X_v, double* X_p,
                        double* Y_u, double* Y_v, double* Y_p,
                        const double* u_adv, const double* v_adv,
                        const double* rho_c, const double* rho_u, const double* rho_v,
                        const double* kappa, const double* zeta, const double* eta,
                        double dt_step, int include_advection,
                        const double* penalty_u, const double* penalty_v,
                        const double* mdot_u_cons, const double* mdot_v_cons);
/* src/mod_bicgstab2.f90:32-72 */
int fugu_coupled_dot(fugu_ctx* ctx, const double* u1, const double* v1, const double* p1,
                     const double* u2, const double* v2, const double* p2, double* dot_out);
/* coupled_dot without a context and without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The six arrays are device memory of the current device in the
 * padded layout of `grid`; the sum (the same double as fugu_coupled_dot in the same order) is
 * written to the device pointer dot_out, asynchronously on `stream` (a hipStream_t, NULL = the
 * null stream): the caller synchronises.  The scratch is a fixed device array of the library
 * (8.4 MB per GPU, held until the process ends), shared by every call on that GPU: calls that
 * overlap on one GPU race.  sum_order FUGU_SUM_REFERENCE: at most 2^20 owned u + v + p points
 * (e.g. 590 x 590 cells), larger grids return an error; FUGU_SUM_TREE: any size. */
int fugu_coupled_dot_noalloc(const fugu_grid* grid, const double* u1, const double* v1, const double* p1,
                             const double* u2, const double* v2, const double* p2, double* dot_out,
                             int sum_order, void* stream);
/* src/mod_subsolvers.f90:32-246 (tol <= 0 / max_iter <= 0 select the defaults 1e-5 / 150) */
int fugu_solve_poisson_pressure(fugu_ctx* ctx, double* phi, const double* rhs,
                                const double* rho_u, const double* rho_v,
                                double tol, int max_iter, int* iter_out, double* res_out);
/* No counterpart in the reference: the pressure operator of solve_poisson_pressure (same
 * coefficients, zero-mean projection with walls) solved by n_cycles multigrid V-cycles from
 * zero -- the pressure block of the FUGU_PRECOND_MG preconditioner.  phi: interior written,
 * ghosts zero. */
int fugu_solve_pressure_mg(fugu_ctx* ctx, double* phi, const double* rhs, const double* rho_u,
                           const double* rho_v, int n_cycles);
/* src/mod_subsolvers.f90:252-309 and 314-372 (n_sweeps <= 0 selects the default 5) */
int fugu_solve_helmholtz_u(fugu_ctx* ctx, double* u_sol, const double* rhs_u, const double* rho_u,
                           const double* kappa, const double* zeta, const double* eta,
                           double dt_step, int n_sweeps, const double* penalty_u);
int fugu_solve_helmholtz_v(fugu_ctx* ctx, double* v_sol, const double* rhs_v, const double* rho_v,
                           const double* kappa, const double* zeta, const double* eta,
                           double dt_step, int n_sweeps, const double* penalty_v);
/* src/mod_pcd_precond.f90:14-104 */
int fugu_apply_pcd_precond(fugu_ctx* ctx, const double* z_u, const double* z_v, const double* z_p,
                           double* r_u, double* r_v, double* r_p,
                           const double* u_adv, const double* v_adv,
                           const double* rho_u, const double* rho_v,
                           const double* kappa, const double* zeta, const double* eta,
                           double dt_step, const double* penalty_u, const double* penalty_v);
/* src/mod_bicgstab2.f90:77-264 (BiCGStab) and 270-570 (BiCGStab(2)).  tol <= 0 / max_iter
 * <= 0 select the defaults 1e-5 / 100.  flags: FUGU_B2_CONSISTENT_BCS (BiCGStab(2) only). */
int fugu_bicgstab_solve(fugu_ctx* ctx, double* X_u, double* X_v, double* X_p,
                        const double* B_u, const double* B_v, const double* B_p,
                        const double* u_adv, const double* v_adv,
                        const double* rho_c, const double* rho_u, const double* rho_v,
                        const double* kappa, const double* zeta, const double* eta,
                        double dt_step, double tol, int max_iter, int* iter_out, double* res_out,
                        const double* u_top_lid, const double* penalty_u, const double* penalty_v,
                        const double* mdot_u_cons, const double* mdot_v_cons);
int fugu_bicgstab2_solve(fugu_ctx* ctx, double* X_u, double* X_v, double* X_p,
                         const double* B_u, const double* B_v, const double* B_p,
                         const double* u_adv, const double* v_adv,
                         const double* rho_c, const double* rho_u, const double* rho_v,
                         const double* kappa, const double* zeta, const double* eta,
                         double dt_step, double tol, int max_iter, int* iter_out, double* res_out,
                         const double* u_top_lid, const double* penalty_u, const double* penalty_v,
                         const double* mdot_u_cons, const double* mdot_v_cons, unsigned flags);
/* src/mod_vof.f90:38-140 (flux outputs NULL = absent) */
int fugu_vof_advection_tvd(fugu_ctx* ctx, const double* c_in, double* c_out,
                           const double* u_vel, const double* v_vel, double dt_step,
                           double* flux_x_out, double* flux_y_out);
/* src/mod_vof.f90:145-194 (max_iter <= 0 / tol <= 0 select the defaults 100 / 1e-6) */
int fugu_svof_helmholtz_filter(fugu_ctx* ctx, const double* c_raw, double* c_smooth,
                               double delta_filter, int max_iter, double tol, int* iter_out);
/* src/mod_surface_tension.f90:25-72 and 78-109 (sigma from the context parameters) */
int fugu_compute_interface_curvature(fugu_ctx* ctx, const double* c_smooth, double* kappa_curv);
int fugu_compute_csf_force(fugu_ctx* ctx, const double* c_raw, const double* kappa_curv,
                           double* f_sigma_u, double* f_sigma_v);
/* src/mod_viscosity.f90:54-151 */
int fugu_compute_mixture_densities(fugu_ctx* ctx, const double* c, double* rho_c, double* rho_u,
                                   double* rho_v);
int fugu_compute_mixture_viscosities(fugu_ctx* ctx, const double* c, double* kappa, double* zeta,
                                     double* eta);
/* src/mod_solver_coupled.f90:41-91 */
int fugu_compute_adaptive_dt(fugu_ctx* ctx, const double* u, const double* v, double dt_max,
                             double* dt_out);
/* src/mod_solver_coupled.f90:137-380: every state array in/out (host or device). */
int fugu_advance_coupled_step(fugu_ctx* ctx, double* u, double* v, double* p, double* c,
                              double* u_prev, double* v_prev,
                              double* rho_c, double* rho_u, double* rho_v,
                              double* kappa, double* zeta, double* eta,
                              double* kappa_curv, double* f_sigma_u, double* f_sigma_v,
                              double dt_step, int step_idx, const fugu_step_opts* opts,
                              const double* phi_s_u, const double* phi_s_v,
                              int* iter_out, double* res_out);

/* ---- device-resident API ------------------------------------------------------------- */
/* The six restart fields of src/mod_io.f90:169-187 plus the solid mask (NULL = none). */
int fugu_state_upload(fugu_ctx* ctx, const double* u, const double* v, const double* u_prev,
                      const double* v_prev, const double* p, const double* c,
                      const double* phi_s_u, const double* phi_s_v);
int fugu_state_download(fugu_ctx* ctx, double* u, double* v, double* u_prev, double* v_prev,
                        double* p, double* c);
/* Download any device field by name ("rho_c", "kappa_curv", "f_sigma_u", ...). */
int fugu_field_download(fugu_ctx* ctx, const char* name, double* out);
/* Mixture properties from the resident C (src/main.f90:185-186). */
int fugu_state_init_properties(fugu_ctx* ctx);

/* ---- checkpoint / restart and VTK (src/mod_io.f90) at one rank, from the resident state -------- */
/* src/mod_io.f90:169-187 write_checkpoint: "<prefix>_r0000.bin", stream bytes: int32 step, real64 t,
 * int32 nx, ny, nh (24 bytes), then u, v, u_prev, v_prev, p, C as whole padded arrays, native byte
 * order -- the file the reference's read_checkpoint reads. */
int fugu_checkpoint_write(fugu_ctx* ctx, const char* prefix, int step, double t);
/* src/mod_io.f90:192-225 read_checkpoint: *ok = 0 when the file is missing or its nx, ny, nh differ
 * (the state is then untouched; step and t are set when the header was read); on success the six
 * arrays become the resident state and the mixture properties are recomputed from C, as the
 * reference driver does after a restart (src/main.f90:122-131, 185-186).  A restarted run continues
 * with step index step + 1. */
int fugu_checkpoint_read(fugu_ctx* ctx, const char* prefix, int* step, double* t, int* ok);
/* src/mod_io.f90:29-164 write_vtk_fields: legacy ASCII STRUCTURED_POINTS with VOF_C, Pressure,
 * Density, Curvature and the cell-centred velocity, every value in the reference's '(ES15.7)' /
 * '(3ES15.7)' editing. */
int fugu_write_vtk(fugu_ctx* ctx, const char* filename);

typedef struct {
  int iters;              /* Krylov iterations (inner ones summed for the refined solve) */
  int outer;              /* refinement steps (0 for the reference solver) */
  double res;             /* reference: recurrence residual; refined: true relative residual */
  double true_res;        /* ||B - A X|| / ||B|| in the coupled_dot norm, FP64 */
  double solve_seconds;   /* device time of the coupled solve (HIP events) */
  double step_seconds;    /* device time of the whole step */
  int pcg_calls;          /* preconditioner applications */
  /* Volume budget of the step's VOF update (src/mod_vof.f90:120-124), sums of C dx dy over the
   * interior cells: before the update, its change, and the part of the change the clamp of C to
   * [0, 1] made (mod_vof.f90:124).  Walls carry no flux and periodic faces cancel, so
   * vol_change - vol_clamp is the flux form's conservation error (rounding only). */
  double vol_before, vol_change, vol_clamp;
} fugu_step_stats;

/* Per-kernel device timing: when enabled, launches of the instrumented kernels ("pcg_iter1_f64",
 * "pcg_iter2_f64", "matvec_f64", their _f32 twins, "mg_smooth_*", "mg_resid_*", "mdot", "maxpy")
 * are bracketed by HIP events on the context stream (up to 1024 per kernel between
 * synchronisations); total_bytes sums each timed launch's algorithmic bytes (unique fields read
 * and written once).  Enabling resets the totals.
 * No counterpart in the reference (it has no timers); used to measure kernel rooflines. */
int fugu_kernel_timing(fugu_ctx* ctx, int enable);
int fugu_kernel_time(fugu_ctx* ctx, const char* kernel, double* total_ms, long long* count,
                     double* total_bytes);

/* One advance_coupled_step on the resident state.  stats may be NULL. */
int fugu_step(fugu_ctx* ctx, double dt_step, int step_idx, const fugu_step_opts* opts,
              fugu_step_stats* stats);

/* ---- adjoint (no counterpart in the reference: it has no adjoint) ----------------------
 * The step operator as the solvers see it maps the owned degrees of freedom (u: j=1..ny,
 * i=1..i_u_max; v: i=1..nx, j=1..j_v_max; p: interior) to the owned rows:
 *     A_eff X = R A_pad H B E0 X
 * (embed, velocity BCs with lid 0, halo exchange, the padded matvec of
 * src/mod_operators.f90:341-443, owned rows).  Inputs are read on the owned sets; outputs are
 * zero outside them. */
/* A_eff^T L (transpose of src/mod_operators.f90:341-443 with the BC extension folded in). */
int fugu_matvec_transpose(fugu_ctx* ctx, const double* L_u, const double* L_v, const double* L_p,
                          double* O_u, double* O_v, double* O_p,
                          const double* u_adv, const double* v_adv,
                          const double* rho_c, const double* rho_u, const double* rho_v,
                          const double* kappa, const double* zeta, const double* eta,
                          double dt_step, int include_advection);
/* Transpose of the linearised PCD (src/mod_pcd_precond.f90:14-104 with the truncated PCG of
 * src/mod_subsolvers.f90:32-246 replaced by the projected pressure solve phi = Pi A_p^+ Pi z_p):
 * ptol > 0 -> projected Jacobi-PCG to ptol, at most pmax iterations (pmax <= 0: 5000);
 * ptol <= 0 -> pmax multigrid V-cycles.  visc: coefficient of the viscous Schur term. */
int fugu_apply_pcd_transpose(fugu_ctx* ctx, const double* y_u, const double* y_v, const double* y_p,
                             double* z_u, double* z_v, double* z_p,
                             const double* u_adv, const double* v_adv,
                             const double* rho_u, const double* rho_v,
                             const double* kappa, const double* zeta, const double* eta,
                             double dt_step, double ptol, int pmax, double visc);
/* A_eff^T L = G by FP64 FGMRES(30) on the true residual with the transposed PCD (multigrid
 * pressure block, viscous Schur term) as flexible preconditioner; L in: initial guess, out:
 * solution.  tol <= 0: 1e-10; max_iter <= 0: 400.  Returns FUGU_W_MAX_ITER if not converged. */
int fugu_adjoint_solve(fugu_ctx* ctx, double* L_u, double* L_v, double* L_p,
                       const double* G_u, const double* G_v, const double* G_p,
                       const double* u_adv, const double* v_adv,
                       const double* rho_c, const double* rho_u, const double* rho_v,
                       const double* kappa, const double* zeta, const double* eta,
                       double dt_step, double tol, int max_iter, int* iter_out, double* res_out);
/* Transposes of the interface chain (reverse mode at the given state; the kinks take one side:
 * the clamp to [0, 1] passes the derivative only strictly inside, the limiters their one-sided
 * slope, the upwind switch the branch the forward takes, a face with |den| <= 1e-12 only its
 * upwind cell).  Outputs are whole padded arrays, overwritten.
 * (dVOF/dC_in)^T, (dVOF/du)^T, (dVOF/dv)^T of fugu_vof_advection_tvd (src/mod_vof.f90:38-140,
 * the context's limiter) applied to cbar_out. */
int fugu_vof_advection_transpose(fugu_ctx* ctx, const double* c_in, const double* u, const double* v,
                                 double dt_step, const double* cbar_out, double* cbar_in, double* ubar,
                                 double* vbar);
/* (dSVOF/dC_raw)^T csbar of fugu_svof_helmholtz_filter (src/mod_vof.f90:145-194) at the sweep count
 * its forward run on c_raw stops at (returned in iter_out): the sweeps are linear in C_raw.
 * max_iter <= 0: 100; tol <= 0: 1e-6. */
int fugu_svof_transpose(fugu_ctx* ctx, const double* c_raw, double delta, int max_iter, double tol,
                        const double* csbar, double* crawbar, int* iter_out);
/* (dkappa/dC_smooth)^T kbar of fugu_compute_interface_curvature (src/mod_surface_tension.f90:25-72). */
int fugu_curvature_transpose(fugu_ctx* ctx, const double* c_smooth, const double* kbar, double* csbar);
/* (dF/dC_raw)^T and (dF/dkappa)^T of fugu_compute_csf_force (src/mod_surface_tension.f90:78-109)
 * applied to (fubar, fvbar). */
int fugu_csf_transpose(fugu_ctx* ctx, const double* c_raw, const double* kappa_curv, double sigma,
                       const double* fubar, const double* fvbar, double* cbar, double* kbar);
/* Per-cell property path (the configurations with per-inclusion properties): mu_cell, rho_cell
 * over the padded cells -> the operator coefficients with the edge conventions of
 * src/mod_viscosity.f90:54-151; vertex viscosity 4 / sum(1/mu_k), which equals
 * harmonic_viscosity(mean C) (:28-46) for mixture-law inputs. */
int fugu_cell_properties(fugu_ctx* ctx, const double* mu_cell, const double* rho_cell,
                         double* rho_c, double* rho_u, double* rho_v,
                         double* kappa, double* zeta, double* eta);

typedef struct {
  int nsteps;           /* horizon N (BDF2 from step 2 when time_order is 2) */
  double dt;            /* fixed step */
  double y0;            /* objective band: J = 1/2 sum over owned faces with centre y >= y0 of
                           rho_face u^2 dx dy at step N */
  double tol;           /* true relative residual of every forward and adjoint solve (<= 0: 1e-12) */
  int max_iter;         /* FGMRES steps per solve (<= 0: 400) */
  int restart;          /* FGMRES restart (<= 0: 30) */
  int mg_cycles;        /* V-cycles per preconditioner application (<= 0: 1) */
  unsigned flags;       /* FUGU_ADJ_* */
  int checkpoint_every; /* > 1: keep the restart state of every checkpoint_every-th step only and
                           recompute each segment's forward states in the reverse sweep (bit-identical
                           to keeping them all); <= 1: the whole horizon stays resident */
} fugu_adjoint_opts;
enum {
  FUGU_ADJ_FORWARD_ONLY = 1u << 0,     /* J only (no reverse sweep) */
  /* use the interface (unit-sigma CSF forces) recorded by the previous call with the same
   * horizon instead of advecting C: the frozen-interface map the gradient differentiates, as
   * finite differences of the gradient need it */
  FUGU_ADJ_REUSE_INTERFACE = 1u << 1,
  /* velocity blocks of the forward and transposed preconditioners by multigrid V-cycles
   * (FUGU_PRECOND_VMG) */
  FUGU_ADJ_PRECOND_VMG = 1u << 2,
  /* forward solves by the mixed form (FP32 Arnoldi cycles, FP64 true residual and update at
   * every restart: FUGU_SOLVE_MIXED); the adjoint solves stay FP64 unless FUGU_ADJ_MIXED_ADJOINT */
  FUGU_ADJ_MIXED_FORWARD = 1u << 3,
  /* differentiate the interface motion too: the VOF advection (its limiter and clamps taken one
   * sided at their kinks), the SVOF filter at each step's sweep count, the curvature and the CSF
   * force are transposed in the reverse sweep, so the gradient sees how the parameters move the
   * interface and, through the surface force, the flow (otherwise the interface is held frozen);
   * not with FUGU_ADJ_REUSE_INTERFACE */
  FUGU_ADJ_MOVING_INTERFACE = 1u << 4,
  /* adjoint solves by the mixed form: FP32 Arnoldi cycles of the transposed operator and the
   * transposed PCD (FP32 storage, FP64 arithmetic in the kernels), FP64 true residual and update
   * at every restart, so every adjoint solve still meets opts->tol in FP64 */
  FUGU_ADJ_MIXED_ADJOINT = 1u << 5
};
typedef struct {
  double J;
  double forward_seconds, reverse_seconds;   /* device time (HIP events) */
  int forward_iters, adjoint_iters;           /* FGMRES steps summed over the horizon */
  double max_forward_res, max_adjoint_res;    /* worst true relative residual */
} fugu_adjoint_stats;
/* Gradient of J with respect to the per-cell viscosity and density (padded p-kind arrays)
 * and the surface tension, over a BDF2 horizon from (u0, v0, C0), properties held at the
 * per-cell fields, interface motion frozen in the adjoint (VOF, SVOF and CSF of the forward
 * pass recorded and held fixed) unless FUGU_ADJ_MOVING_INTERFACE, every solve to opts->tol in FP64 (the mixed
 * forms refine FP32 cycles to it).  g_mu, g_rho, g_sigma may be
 * NULL with FUGU_ADJ_FORWARD_ONLY; c0 may be NULL with FUGU_ADJ_REUSE_INTERFACE. */
int fugu_adjoint_gradient(fugu_ctx* ctx, const fugu_adjoint_opts* opts, const double* mu_cell,
                          const double* rho_cell, double sigma, const double* c0, const double* u0,
                          const double* v0, double* g_mu, double* g_rho, double* g_sigma,
                          fugu_adjoint_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
