! fugu2d_hip.f90 -- Fortran drop-in interface to libfugu2d_hip.so (MI355X / gfx950).
!
! The module exports the public procedures of the reference's hot-path modules under their
! own names and argument lists (explicit-shape padded dummies, the same optional
! arguments), implemented by the HIP library through iso_c_binding:
!
!   mod_operators       op_divergence, apply_velocity_bcs, matvec_coupled
!   mod_bicgstab2       coupled_dot, bicgstab_solve, bicgstab2_solve
!   mod_pcd_precond     apply_pcd_precond
!   mod_subsolvers      solve_poisson_pressure, solve_helmholtz_u, solve_helmholtz_v
!   mod_vof             vof_advection_tvd, svof_helmholtz_filter
!   mod_surface_tension compute_interface_curvature, compute_csf_force
!   mod_viscosity       compute_mixture_densities, compute_mixture_viscosities
!   mod_solver_coupled  compute_adaptive_dt, advance_coupled_step
!
! A driver swaps `use mod_operators` (etc.) for `use fugu2d_hip` and calls fugu_hip_init
! once after the grid and the namelist parameters are known (the reference reads them from
! module globals: dx_grid, dy_grid, periods of src/mod_mpi.f90:10-26 and rho1 ... of
! src/mod_params.f90:19-50).  The domain is one device-resident block (one rank).  A
! negative status from the library stops the program with the library's message.
module fugu2d_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter :: dp = c_double
  integer, parameter :: ip = c_int32_t

  type, bind(C), public :: fugu_grid_t
    integer(c_int) :: nx, ny, nh
    real(c_double) :: lx, ly
    integer(c_int) :: periodic_x, periodic_y
  end type fugu_grid_t

  type, bind(C), public :: fugu_params_t
    real(c_double) :: rho1, rho2, mu1, mu2, sigma, gx, gy
    integer(c_int) :: time_order
    real(c_double) :: alpha_penalty, cfl_target
  end type fugu_params_t

  type, bind(C), public :: fugu_step_opts_t
    integer(c_int) :: use_bicgstab2 = 1, has_lid = 0
    real(c_double) :: u_lid = 0.0_dp, tol = 1.0e-6_dp
    integer(c_int) :: max_iter = 500, use_amp = 0
    real(c_double) :: u_solid = 0.0_dp, v_solid = 0.0_dp
    integer(c_int) :: solver_flags = 0
    real(c_double) :: inner_tol = 0.0_dp
    integer(c_int) :: inner_max_iter = 0, mg_cycles = 0, restart = 0
  end type fugu_step_opts_t

  ! solver_flags bits of fugu_step_opts (include/fugu2d.h)
  integer(c_int), parameter, public :: FUGU_B2_CONSISTENT_BCS = 1, FUGU_SOLVE_REFINED = 2, &
                                       FUGU_SOLVE_MIXED = 4, FUGU_PRECOND_MG = 8, FUGU_PRECOND_VISC = 16, &
                                       FUGU_PRECOND_VMG = 32, FUGU_SOLVE_EXTRAPOLATE = 64
  integer(c_int), parameter, public :: FUGU_KIND_U = 0, FUGU_KIND_V = 1, FUGU_KIND_P = 2, FUGU_KIND_W = 3

  type(c_ptr), save :: g_ctx = c_null_ptr
  integer(ip), save :: g_nx = -1, g_ny = -1, g_nh = -1

  public :: fugu_hip_init, fugu_hip_set_params, fugu_hip_finalize, fugu_hip_exchange_halos, fugu_hip_step_opts
  public :: op_divergence, apply_velocity_bcs, matvec_coupled
  public :: coupled_dot, bicgstab_solve, bicgstab2_solve
  public :: apply_pcd_precond
  public :: solve_poisson_pressure, solve_helmholtz_u, solve_helmholtz_v
  public :: vof_advection_tvd, svof_helmholtz_filter
  public :: compute_interface_curvature, compute_csf_force
  public :: compute_mixture_densities, compute_mixture_viscosities
  public :: compute_adaptive_dt, advance_coupled_step

  interface
    integer(c_int) function c_ctx_create(grid, params, ctx) bind(C, name="fugu_ctx_create")
      import :: c_int, c_ptr, fugu_grid_t, fugu_params_t
      type(fugu_grid_t), intent(in) :: grid
      type(fugu_params_t), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function c_ctx_destroy(ctx) bind(C, name="fugu_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_set_params(ctx, params) bind(C, name="fugu_set_params")
      import :: c_int, c_ptr, fugu_params_t
      type(c_ptr), value :: ctx
      type(fugu_params_t), intent(in) :: params
    end function
    type(c_ptr) function c_last_error() bind(C, name="fugu_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c_exchange_halos(ctx, a, kind) bind(C, name="fugu_exchange_halos")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, a
      integer(c_int), value :: kind
    end function
    integer(c_int) function c_op_divergence(ctx, u, v, div) bind(C, name="fugu_op_divergence")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, v, div
    end function
    integer(c_int) function c_velocity_bcs(ctx, u, v, lid) bind(C, name="fugu_apply_velocity_bcs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, v, lid
    end function
    integer(c_int) function c_matvec(ctx, xu, xv, xp, yu, yv, yp, uadv, vadv, rho_c, rho_u, rho_v, &
                                     kappa, zeta, eta, dt, inc_adv, pu, pv, mu, mv) &
        bind(C, name="fugu_matvec_coupled")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, xu, xv, xp, yu, yv, yp, uadv, vadv, rho_c, rho_u, rho_v, kappa, zeta, eta
      real(c_double), value :: dt
      integer(c_int), value :: inc_adv
      type(c_ptr), value :: pu, pv, mu, mv
    end function
    integer(c_int) function c_coupled_dot(ctx, u1, v1, p1, u2, v2, p2, d) bind(C, name="fugu_coupled_dot")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, u1, v1, p1, u2, v2, p2
      real(c_double), intent(out) :: d
    end function
    integer(c_int) function c_poisson(ctx, phi, rhs, rho_u, rho_v, tol, max_iter, it, res) &
        bind(C, name="fugu_solve_poisson_pressure")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, phi, rhs, rho_u, rho_v
      real(c_double), value :: tol
      integer(c_int), value :: max_iter
      integer(c_int), intent(out) :: it
      real(c_double), intent(out) :: res
    end function
    integer(c_int) function c_helm_u(ctx, x, rhs, rho, kappa, zeta, eta, dt, n, pen) &
        bind(C, name="fugu_solve_helmholtz_u")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, rhs, rho, kappa, zeta, eta
      real(c_double), value :: dt
      integer(c_int), value :: n
      type(c_ptr), value :: pen
    end function
    integer(c_int) function c_helm_v(ctx, x, rhs, rho, kappa, zeta, eta, dt, n, pen) &
        bind(C, name="fugu_solve_helmholtz_v")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, x, rhs, rho, kappa, zeta, eta
      real(c_double), value :: dt
      integer(c_int), value :: n
      type(c_ptr), value :: pen
    end function
    integer(c_int) function c_pcd(ctx, zu, zv, zp, ru, rv, rp, uadv, vadv, rho_u, rho_v, kappa, zeta, eta, &
                                  dt, pu, pv) bind(C, name="fugu_apply_pcd_precond")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, zu, zv, zp, ru, rv, rp, uadv, vadv, rho_u, rho_v, kappa, zeta, eta
      real(c_double), value :: dt
      type(c_ptr), value :: pu, pv
    end function
    integer(c_int) function c_bicgstab(ctx, xu, xv, xp, bu, bv, bp, uadv, vadv, rho_c, rho_u, rho_v, &
                                       kappa, zeta, eta, dt, tol, max_iter, it, res, lid, pu, pv, mu, mv) &
        bind(C, name="fugu_bicgstab_solve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, xu, xv, xp, bu, bv, bp, uadv, vadv, rho_c, rho_u, rho_v, kappa, zeta, eta
      real(c_double), value :: dt, tol
      integer(c_int), value :: max_iter
      integer(c_int), intent(out) :: it
      real(c_double), intent(out) :: res
      type(c_ptr), value :: lid, pu, pv, mu, mv
    end function
    integer(c_int) function c_bicgstab2(ctx, xu, xv, xp, bu, bv, bp, uadv, vadv, rho_c, rho_u, rho_v, &
                                        kappa, zeta, eta, dt, tol, max_iter, it, res, lid, pu, pv, mu, mv, &
                                        flags) bind(C, name="fugu_bicgstab2_solve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, xu, xv, xp, bu, bv, bp, uadv, vadv, rho_c, rho_u, rho_v, kappa, zeta, eta
      real(c_double), value :: dt, tol
      integer(c_int), value :: max_iter
      integer(c_int), intent(out) :: it
      real(c_double), intent(out) :: res
      type(c_ptr), value :: lid, pu, pv, mu, mv
      integer(c_int), value :: flags
    end function
    integer(c_int) function c_vof(ctx, cin, cout, u, v, dt, fx, fy) bind(C, name="fugu_vof_advection_tvd")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, cin, cout, u, v
      real(c_double), value :: dt
      type(c_ptr), value :: fx, fy
    end function
    integer(c_int) function c_svof(ctx, craw, cs, delta, max_iter, tol, it) bind(C, name="fugu_svof_helmholtz_filter")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, craw, cs
      real(c_double), value :: delta
      integer(c_int), value :: max_iter
      real(c_double), value :: tol
      integer(c_int), intent(out) :: it
    end function
    integer(c_int) function c_curv(ctx, cs, kc) bind(C, name="fugu_compute_interface_curvature")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, cs, kc
    end function
    integer(c_int) function c_csf(ctx, craw, kc, fu, fv) bind(C, name="fugu_compute_csf_force")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, craw, kc, fu, fv
    end function
    integer(c_int) function c_dens(ctx, c, rc, ru, rv) bind(C, name="fugu_compute_mixture_densities")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, c, rc, ru, rv
    end function
    integer(c_int) function c_visc(ctx, c, kap, zeta, eta) bind(C, name="fugu_compute_mixture_viscosities")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, c, kap, zeta, eta
    end function
    integer(c_int) function c_adaptive_dt(ctx, u, v, dt_max, dt_out) bind(C, name="fugu_compute_adaptive_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, u, v
      real(c_double), value :: dt_max
      real(c_double), intent(out) :: dt_out
    end function
    integer(c_int) function c_advance(ctx, u, v, p, c, u_prev, v_prev, rho_c, rho_u, rho_v, kappa, zeta, eta, &
                                      kc, fu, fv, dt, step_idx, opts, phi_s_u, phi_s_v, it, res) &
        bind(C, name="fugu_advance_coupled_step")
      import :: c_int, c_ptr, c_double, fugu_step_opts_t
      type(c_ptr), value :: ctx, u, v, p, c, u_prev, v_prev, rho_c, rho_u, rho_v, kappa, zeta, eta, kc, fu, fv
      real(c_double), value :: dt
      integer(c_int), value :: step_idx
      type(fugu_step_opts_t), intent(in) :: opts
      type(c_ptr), value :: phi_s_u, phi_s_v
      integer(c_int), intent(out) :: it
      real(c_double), intent(out) :: res
    end function
  end interface

contains

  ! ---- context -------------------------------------------------------------------
  subroutine fugu_hip_init(nx, ny, nh, lx, ly, periodic_x, periodic_y, rho1, rho2, mu1, mu2, sigma, &
                           gx, gy, time_order, alpha_penalty, cfl_target)
    integer(ip), intent(in) :: nx, ny, nh, time_order
    real(dp), intent(in) :: lx, ly, rho1, rho2, mu1, mu2, sigma, gx, gy, alpha_penalty, cfl_target
    logical, intent(in) :: periodic_x, periodic_y
    type(fugu_grid_t) :: g
    type(fugu_params_t) :: p
    if (c_associated(g_ctx)) call fugu_hip_finalize()
    g = fugu_grid_t(nx, ny, nh, lx, ly, merge(1, 0, periodic_x), merge(1, 0, periodic_y))
    p = fugu_params_t(rho1, rho2, mu1, mu2, sigma, gx, gy, time_order, alpha_penalty, cfl_target)
    call check(c_ctx_create(g, p, g_ctx), "fugu_ctx_create")
    g_nx = nx; g_ny = ny; g_nh = nh
  end subroutine fugu_hip_init

  subroutine fugu_hip_set_params(rho1, rho2, mu1, mu2, sigma, gx, gy, time_order, alpha_penalty, cfl_target)
    integer(ip), intent(in) :: time_order
    real(dp), intent(in) :: rho1, rho2, mu1, mu2, sigma, gx, gy, alpha_penalty, cfl_target
    type(fugu_params_t) :: p
    p = fugu_params_t(rho1, rho2, mu1, mu2, sigma, gx, gy, time_order, alpha_penalty, cfl_target)
    call check(c_set_params(ctx(), p), "fugu_set_params")
  end subroutine fugu_hip_set_params

  subroutine fugu_hip_finalize()
    integer(c_int) :: rc
    if (c_associated(g_ctx)) rc = c_ctx_destroy(g_ctx)
    g_ctx = c_null_ptr
    g_nx = -1; g_ny = -1; g_nh = -1
  end subroutine fugu_hip_finalize

  ! Options of advance_coupled_step beyond the reference's optional arguments (the chosen
  ! fast solve forms); the defaults reproduce the reference solver.
  function fugu_hip_step_opts(solver_flags, inner_tol, inner_max_iter, mg_cycles, restart) result(o)
    integer(c_int), intent(in) :: solver_flags
    real(dp), intent(in), optional :: inner_tol
    integer(c_int), intent(in), optional :: inner_max_iter, mg_cycles, restart
    type(fugu_step_opts_t) :: o
    o%solver_flags = solver_flags
    if (present(inner_tol)) o%inner_tol = inner_tol
    if (present(inner_max_iter)) o%inner_max_iter = inner_max_iter
    if (present(mg_cycles)) o%mg_cycles = mg_cycles
    if (present(restart)) o%restart = restart
  end function fugu_hip_step_opts

  ! src/mod_mpi.f90:113-245 at one rank; kind = FUGU_KIND_U / V / P
  subroutine fugu_hip_exchange_halos(a, kind)
    real(dp), intent(inout), target :: a(*)
    integer(c_int), intent(in) :: kind
    call check(c_exchange_halos(ctx(), c_loc(a), kind), "fugu_exchange_halos")
  end subroutine fugu_hip_exchange_halos

  ! ---- mod_operators ---------------------------------------------------------------
  subroutine op_divergence(u_vel, v_vel, div_field, nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: u_vel
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: v_vel
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: div_field
    call dims(nx, ny, nh)
    call check(c_op_divergence(ctx(), c_loc(u_vel), c_loc(v_vel), c_loc(div_field)), "op_divergence")
  end subroutine op_divergence

  subroutine apply_velocity_bcs(u_vel, v_vel, nx, ny, nh, u_top_lid)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: u_vel
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: v_vel
    real(dp), optional, intent(in), target :: u_top_lid
    call dims(nx, ny, nh)
    call check(c_velocity_bcs(ctx(), c_loc(u_vel), c_loc(v_vel), opt0(u_top_lid)), "apply_velocity_bcs")
  end subroutine apply_velocity_bcs

  subroutine matvec_coupled(X_u, X_v, X_p, Y_u, Y_v, Y_p, u_adv, v_adv, rho_c, rho_u, rho_v, &
                            kappa, zeta, eta, dt_step, nx, ny, nh, include_advection, penalty_u, penalty_v, &
                            mdot_u_cons, mdot_v_cons)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: X_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: X_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: X_p
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: Y_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(out), target :: Y_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: Y_p
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: u_adv, rho_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: v_adv, rho_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: rho_c, kappa
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: zeta, eta
    real(dp), intent(in) :: dt_step
    logical, optional, intent(in) :: include_advection
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(in), target :: penalty_u, mdot_u_cons
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(in), target :: penalty_v, mdot_v_cons
    integer(c_int) :: adv
    call dims(nx, ny, nh)
    adv = 1
    if (present(include_advection)) adv = merge(1, 0, include_advection)
    call check(c_matvec(ctx(), c_loc(X_u), c_loc(X_v), c_loc(X_p), c_loc(Y_u), c_loc(Y_v), c_loc(Y_p), &
                        c_loc(u_adv), c_loc(v_adv), c_loc(rho_c), c_loc(rho_u), c_loc(rho_v), c_loc(kappa), &
                        c_loc(zeta), c_loc(eta), dt_step, adv, opt2(penalty_u), opt2(penalty_v), &
                        opt2(mdot_u_cons), opt2(mdot_v_cons)), "matvec_coupled")
  end subroutine matvec_coupled

  ! ---- mod_bicgstab2 ---------------------------------------------------------------
  function coupled_dot(u1, v1, p1, u2, v2, p2, nx, ny, nh) result(dot_val)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: u1, u2
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: v1, v2
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: p1, p2
    real(dp) :: dot_val
    call dims(nx, ny, nh)
    call check(c_coupled_dot(ctx(), c_loc(u1), c_loc(v1), c_loc(p1), c_loc(u2), c_loc(v2), c_loc(p2), dot_val), &
               "coupled_dot")
  end function coupled_dot

  subroutine bicgstab_solve(X_u, X_v, X_p, B_u, B_v, B_p, u_adv, v_adv, rho_c, rho_u, rho_v, &
                            kappa, zeta, eta, dt_step, nx, ny, nh, tol_in, max_iter_in, iter_out, res_out, &
                            u_top_lid, penalty_u, penalty_v, mdot_u_cons, mdot_v_cons)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: X_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: X_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: X_p
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: B_u, u_adv, rho_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: B_v, v_adv, rho_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: B_p, rho_c, kappa
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: zeta, eta
    real(dp), intent(in) :: dt_step
    real(dp), optional, intent(in) :: tol_in
    integer(ip), optional, intent(in) :: max_iter_in
    integer(ip), optional, intent(out) :: iter_out
    real(dp), optional, intent(out) :: res_out
    real(dp), optional, intent(in), target :: u_top_lid
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(in), target :: penalty_u, mdot_u_cons
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(in), target :: penalty_v, mdot_v_cons
    real(dp) :: tol, res
    integer(c_int) :: mi, it
    call dims(nx, ny, nh)
    tol = 1.0e-5_dp
    if (present(tol_in)) tol = tol_in
    mi = 100
    if (present(max_iter_in)) mi = max_iter_in
    call check(c_bicgstab(ctx(), c_loc(X_u), c_loc(X_v), c_loc(X_p), c_loc(B_u), c_loc(B_v), c_loc(B_p), &
                          c_loc(u_adv), c_loc(v_adv), c_loc(rho_c), c_loc(rho_u), c_loc(rho_v), c_loc(kappa), &
                          c_loc(zeta), c_loc(eta), dt_step, tol, mi, it, res, opt0(u_top_lid), &
                          opt2(penalty_u), opt2(penalty_v), opt2(mdot_u_cons), opt2(mdot_v_cons)), "bicgstab_solve")
    if (present(iter_out)) iter_out = it
    if (present(res_out)) res_out = res
  end subroutine bicgstab_solve

  subroutine bicgstab2_solve(X_u, X_v, X_p, B_u, B_v, B_p, u_adv, v_adv, rho_c, rho_u, rho_v, &
                             kappa, zeta, eta, dt_step, nx, ny, nh, tol_in, max_iter_in, iter_out, res_out, &
                             u_top_lid, penalty_u, penalty_v, mdot_u_cons, mdot_v_cons)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: X_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: X_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: X_p
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: B_u, u_adv, rho_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: B_v, v_adv, rho_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: B_p, rho_c, kappa
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: zeta, eta
    real(dp), intent(in) :: dt_step
    real(dp), optional, intent(in) :: tol_in
    integer(ip), optional, intent(in) :: max_iter_in
    integer(ip), optional, intent(out) :: iter_out
    real(dp), optional, intent(out) :: res_out
    real(dp), optional, intent(in), target :: u_top_lid
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(in), target :: penalty_u, mdot_u_cons
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(in), target :: penalty_v, mdot_v_cons
    real(dp) :: tol, res
    integer(c_int) :: mi, it
    call dims(nx, ny, nh)
    tol = 1.0e-5_dp
    if (present(tol_in)) tol = tol_in
    mi = 100
    if (present(max_iter_in)) mi = max_iter_in
    call check(c_bicgstab2(ctx(), c_loc(X_u), c_loc(X_v), c_loc(X_p), c_loc(B_u), c_loc(B_v), c_loc(B_p), &
                           c_loc(u_adv), c_loc(v_adv), c_loc(rho_c), c_loc(rho_u), c_loc(rho_v), c_loc(kappa), &
                           c_loc(zeta), c_loc(eta), dt_step, tol, mi, it, res, opt0(u_top_lid), &
                           opt2(penalty_u), opt2(penalty_v), opt2(mdot_u_cons), opt2(mdot_v_cons), 0), &
               "bicgstab2_solve")
    if (present(iter_out)) iter_out = it
    if (present(res_out)) res_out = res
  end subroutine bicgstab2_solve

  ! ---- mod_pcd_precond -------------------------------------------------------------
  subroutine apply_pcd_precond(z_u, z_v, z_p, r_u, r_v, r_p, u_adv, v_adv, rho_u, rho_v, &
                               kappa, zeta, eta, dt_step, nx, ny, nh, penalty_u, penalty_v)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: z_u, u_adv, rho_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: z_v, v_adv, rho_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: z_p, kappa
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: r_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(out), target :: r_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: r_p
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: zeta, eta
    real(dp), intent(in) :: dt_step
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(in), target :: penalty_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(in), target :: penalty_v
    call dims(nx, ny, nh)
    call check(c_pcd(ctx(), c_loc(z_u), c_loc(z_v), c_loc(z_p), c_loc(r_u), c_loc(r_v), c_loc(r_p), &
                     c_loc(u_adv), c_loc(v_adv), c_loc(rho_u), c_loc(rho_v), c_loc(kappa), c_loc(zeta), &
                     c_loc(eta), dt_step, opt2(penalty_u), opt2(penalty_v)), "apply_pcd_precond")
  end subroutine apply_pcd_precond

  ! ---- mod_subsolvers --------------------------------------------------------------
  subroutine solve_poisson_pressure(phi, rhs, rho_u, rho_v, nx, ny, nh, tol_in, max_iter_in, iter_out, res_out)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: phi
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: rhs
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: rho_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: rho_v
    real(dp), optional, intent(in) :: tol_in
    integer(ip), optional, intent(in) :: max_iter_in
    integer(ip), optional, intent(out) :: iter_out
    real(dp), optional, intent(out) :: res_out
    real(dp) :: tol, res
    integer(c_int) :: mi, it
    call dims(nx, ny, nh)
    tol = 1.0e-5_dp
    if (present(tol_in)) tol = tol_in
    mi = 150
    if (present(max_iter_in)) mi = max_iter_in
    call check(c_poisson(ctx(), c_loc(phi), c_loc(rhs), c_loc(rho_u), c_loc(rho_v), tol, mi, it, res), &
               "solve_poisson_pressure")
    if (present(iter_out)) iter_out = it
    if (present(res_out)) res_out = res
  end subroutine solve_poisson_pressure

  subroutine solve_helmholtz_u(u_sol, rhs_u, rho_u, kappa, zeta, eta, dt_step, nx, ny, nh, n_sweeps_in, penalty_u)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: u_sol
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: rhs_u, rho_u
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: kappa
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: zeta, eta
    real(dp), intent(in) :: dt_step
    integer(ip), optional, intent(in) :: n_sweeps_in
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(in), target :: penalty_u
    integer(c_int) :: n
    call dims(nx, ny, nh)
    n = 5
    if (present(n_sweeps_in)) n = n_sweeps_in
    call check(c_helm_u(ctx(), c_loc(u_sol), c_loc(rhs_u), c_loc(rho_u), c_loc(kappa), c_loc(zeta), c_loc(eta), &
                        dt_step, n, opt2(penalty_u)), "solve_helmholtz_u")
  end subroutine solve_helmholtz_u

  subroutine solve_helmholtz_v(v_sol, rhs_v, rho_v, kappa, zeta, eta, dt_step, nx, ny, nh, n_sweeps_in, penalty_v)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: v_sol
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: rhs_v, rho_v
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: kappa
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: zeta, eta
    real(dp), intent(in) :: dt_step
    integer(ip), optional, intent(in) :: n_sweeps_in
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(in), target :: penalty_v
    integer(c_int) :: n
    call dims(nx, ny, nh)
    n = 5
    if (present(n_sweeps_in)) n = n_sweeps_in
    call check(c_helm_v(ctx(), c_loc(v_sol), c_loc(rhs_v), c_loc(rho_v), c_loc(kappa), c_loc(zeta), c_loc(eta), &
                        dt_step, n, opt2(penalty_v)), "solve_helmholtz_v")
  end subroutine solve_helmholtz_v

  ! ---- mod_vof ---------------------------------------------------------------------
  subroutine vof_advection_tvd(C_in, C_out, u_vel, v_vel, dt_step, nx, ny, nh, flux_x_out, flux_y_out)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: C_in
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: C_out
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: u_vel
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: v_vel
    real(dp), intent(in) :: dt_step
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(out), target :: flux_x_out
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(out), target :: flux_y_out
    type(c_ptr) :: fx, fy
    call dims(nx, ny, nh)
    fx = c_null_ptr; fy = c_null_ptr
    if (present(flux_x_out)) fx = c_loc(flux_x_out)
    if (present(flux_y_out)) fy = c_loc(flux_y_out)
    call check(c_vof(ctx(), c_loc(C_in), c_loc(C_out), c_loc(u_vel), c_loc(v_vel), dt_step, fx, fy), &
               "vof_advection_tvd")
  end subroutine vof_advection_tvd

  subroutine svof_helmholtz_filter(C_raw, C_smooth, delta_filter, nx, ny, nh, max_iter, tol)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: C_raw
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: C_smooth
    real(dp), intent(in) :: delta_filter
    integer(ip), optional, intent(in) :: max_iter
    real(dp), optional, intent(in) :: tol
    integer(c_int) :: mi, it
    real(dp) :: t
    call dims(nx, ny, nh)
    mi = 100
    if (present(max_iter)) mi = max_iter
    t = 1.0e-6_dp
    if (present(tol)) t = tol
    call check(c_svof(ctx(), c_loc(C_raw), c_loc(C_smooth), delta_filter, mi, t, it), "svof_helmholtz_filter")
  end subroutine svof_helmholtz_filter

  ! ---- mod_surface_tension ---------------------------------------------------------
  subroutine compute_interface_curvature(C_smooth, kappa_curv, nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: C_smooth
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: kappa_curv
    call dims(nx, ny, nh)
    call check(c_curv(ctx(), c_loc(C_smooth), c_loc(kappa_curv)), "compute_interface_curvature")
  end subroutine compute_interface_curvature

  subroutine compute_csf_force(C_raw, kappa_curv, f_sigma_u, f_sigma_v, nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: C_raw, kappa_curv
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: f_sigma_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(out), target :: f_sigma_v
    call dims(nx, ny, nh)
    call check(c_csf(ctx(), c_loc(C_raw), c_loc(kappa_curv), c_loc(f_sigma_u), c_loc(f_sigma_v)), &
               "compute_csf_force")
  end subroutine compute_csf_force

  ! ---- mod_viscosity ---------------------------------------------------------------
  subroutine compute_mixture_densities(c_field, rho_c, rho_u, rho_v, nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: c_field
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: rho_c
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: rho_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(out), target :: rho_v
    call dims(nx, ny, nh)
    call check(c_dens(ctx(), c_loc(c_field), c_loc(rho_c), c_loc(rho_u), c_loc(rho_v)), &
               "compute_mixture_densities")
  end subroutine compute_mixture_densities

  subroutine compute_mixture_viscosities(c_field, kappa, zeta, eta, nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: c_field
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(out), target :: kappa
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(out), target :: zeta, eta
    call dims(nx, ny, nh)
    call check(c_visc(ctx(), c_loc(c_field), c_loc(kappa), c_loc(zeta), c_loc(eta)), "compute_mixture_viscosities")
  end subroutine compute_mixture_viscosities

  ! ---- mod_solver_coupled ----------------------------------------------------------
  subroutine compute_adaptive_dt(u_vel, v_vel, dt_max, dt_out, nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(in), target :: u_vel
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(in), target :: v_vel
    real(dp), intent(in) :: dt_max
    real(dp), intent(out) :: dt_out
    call dims(nx, ny, nh)
    call check(c_adaptive_dt(ctx(), c_loc(u_vel), c_loc(v_vel), dt_max, dt_out), "compute_adaptive_dt")
  end subroutine compute_adaptive_dt

  ! `fast_opts` (not in the reference) selects the chosen fast solve forms; absent, the
  ! reference solver runs with the reference's defaults (tol_bicgstab2 1e-6, 500 iterations).
  subroutine advance_coupled_step(u_vel, v_vel, p_field, c_field, u_prev, v_prev, &
                                  rho_c, rho_u, rho_v, kappa, zeta, eta, kappa_curv, f_sigma_u, f_sigma_v, &
                                  dt_step, step_idx, nx, ny, nh, iter_out, res_out, use_bicgstab2, u_top_lid, &
                                  tol_in, max_iter_in, consistent_amp, phi_s_u, phi_s_v, u_solid_in, v_solid_in, &
                                  fast_opts)
    integer(ip), intent(in) :: step_idx, nx, ny, nh
    real(dp), intent(in) :: dt_step
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: u_vel, u_prev
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: v_vel, v_prev
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: p_field, c_field
    real(dp), dimension(1-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: rho_c, kappa, kappa_curv
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), intent(inout), target :: rho_u, f_sigma_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: rho_v, f_sigma_v
    real(dp), dimension(0-nh:nx+nh, 0-nh:ny+nh), intent(inout), target :: zeta, eta
    integer(ip), optional, intent(out) :: iter_out
    real(dp), optional, intent(out) :: res_out
    logical, optional, intent(in) :: use_bicgstab2, consistent_amp
    real(dp), optional, intent(in) :: u_top_lid, tol_in, u_solid_in, v_solid_in
    integer(ip), optional, intent(in) :: max_iter_in
    real(dp), dimension(0-nh:nx+nh, 1-nh:ny+nh), optional, intent(in), target :: phi_s_u
    real(dp), dimension(1-nh:nx+nh, 0-nh:ny+nh), optional, intent(in), target :: phi_s_v
    type(fugu_step_opts_t), optional, intent(in) :: fast_opts
    type(fugu_step_opts_t) :: o
    type(c_ptr) :: psu, psv
    integer(c_int) :: it
    real(dp) :: res
    call dims(nx, ny, nh)
    if (present(fast_opts)) o = fast_opts
    o%use_bicgstab2 = 1
    if (present(use_bicgstab2)) o%use_bicgstab2 = merge(1, 0, use_bicgstab2)
    o%has_lid = 0
    o%u_lid = 0.0_dp
    if (present(u_top_lid)) then
      o%has_lid = 1
      o%u_lid = u_top_lid
    end if
    o%tol = 1.0e-6_dp
    if (present(tol_in)) o%tol = tol_in
    o%max_iter = 500
    if (present(max_iter_in)) o%max_iter = max_iter_in
    o%use_amp = 0
    if (present(consistent_amp)) o%use_amp = merge(1, 0, consistent_amp)
    o%u_solid = 0.0_dp
    if (present(u_solid_in)) o%u_solid = u_solid_in
    o%v_solid = 0.0_dp
    if (present(v_solid_in)) o%v_solid = v_solid_in
    psu = c_null_ptr; psv = c_null_ptr
    if (present(phi_s_u) .and. present(phi_s_v)) then
      psu = c_loc(phi_s_u)
      psv = c_loc(phi_s_v)
    end if
    call check(c_advance(ctx(), c_loc(u_vel), c_loc(v_vel), c_loc(p_field), c_loc(c_field), c_loc(u_prev), &
                         c_loc(v_prev), c_loc(rho_c), c_loc(rho_u), c_loc(rho_v), c_loc(kappa), c_loc(zeta), &
                         c_loc(eta), c_loc(kappa_curv), c_loc(f_sigma_u), c_loc(f_sigma_v), dt_step, step_idx, o, &
                         psu, psv, it, res), "advance_coupled_step")
    if (present(iter_out)) iter_out = it
    if (present(res_out)) res_out = res
  end subroutine advance_coupled_step

  ! ---- helpers ---------------------------------------------------------------------
  function ctx() result(p)
    type(c_ptr) :: p
    if (.not. c_associated(g_ctx)) error stop "fugu2d_hip: fugu_hip_init has not been called"
    p = g_ctx
  end function ctx

  subroutine dims(nx, ny, nh)
    integer(ip), intent(in) :: nx, ny, nh
    if (nx /= g_nx .or. ny /= g_ny .or. nh /= g_nh) error stop "fugu2d_hip: array extents differ from fugu_hip_init"
  end subroutine dims

  function opt0(x) result(p)
    real(dp), optional, intent(in), target :: x
    type(c_ptr) :: p
    p = c_null_ptr
    if (present(x)) p = c_loc(x)
  end function opt0

  function opt2(a) result(p)
    real(dp), optional, intent(in), target :: a(:, :)
    type(c_ptr) :: p
    p = c_null_ptr
    if (present(a)) p = c_loc(a)
  end function opt2

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: k
    if (rc >= 0) return
    call c_f_pointer(c_last_error(), msg, [256])
    text = ""
    do k = 1, 256
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write (0, '(a, a, a, i0, a, a)') "fugu2d_hip: ", what, " failed (", rc, "): ", trim(text)
    error stop 1
  end subroutine check

end module fugu2d_hip
