! dropin_check.f90 -- a reference-style driver fragment on the Fortran drop-in module:
! a 32x32 static droplet (src/main.f90:137-151 initialiser: C = 1 inside r <= 0.25), the
! mixture properties, then two advance_coupled_step calls with the reference's BiCGStab
! (use_bicgstab2 = .false.), BDF2, as in the reference driver loop (src/main.f90:197-228).
! Writes the iteration counts, residuals and the state after each step to a raw stream
! file (argument 1) for comparison with the oracle.
program dropin_check
  use, intrinsic :: iso_fortran_env, only: real64, int32
  use fugu2d_hip
  implicit none
  integer, parameter :: dp = real64, ip = int32
  integer(ip), parameter :: nx = 32, ny = 32, nh = 2
  real(dp), parameter :: lx = 1.0_dp, ly = 1.0_dp
  real(dp), allocatable :: u(:, :), v(:, :), p(:, :), c(:, :), u_prev(:, :), v_prev(:, :)
  real(dp), allocatable :: rho_c(:, :), rho_u(:, :), rho_v(:, :), kappa(:, :), zeta(:, :), eta(:, :)
  real(dp), allocatable :: kc(:, :), fu(:, :), fv(:, :)
  real(dp) :: dx, dy, x, y, res, dot
  integer(ip) :: i, j, step, it, unit
  character(len=512) :: path

  call get_command_argument(1, path)
  dx = lx / nx
  dy = ly / ny
  allocate(u(0-nh:nx+nh, 1-nh:ny+nh), u_prev(0-nh:nx+nh, 1-nh:ny+nh), rho_u(0-nh:nx+nh, 1-nh:ny+nh))
  allocate(fu(0-nh:nx+nh, 1-nh:ny+nh))
  allocate(v(1-nh:nx+nh, 0-nh:ny+nh), v_prev(1-nh:nx+nh, 0-nh:ny+nh), rho_v(1-nh:nx+nh, 0-nh:ny+nh))
  allocate(fv(1-nh:nx+nh, 0-nh:ny+nh))
  allocate(p(1-nh:nx+nh, 1-nh:ny+nh), c(1-nh:nx+nh, 1-nh:ny+nh), rho_c(1-nh:nx+nh, 1-nh:ny+nh))
  allocate(kappa(1-nh:nx+nh, 1-nh:ny+nh), kc(1-nh:nx+nh, 1-nh:ny+nh))
  allocate(zeta(0-nh:nx+nh, 0-nh:ny+nh), eta(0-nh:nx+nh, 0-nh:ny+nh))
  u = 0; v = 0; p = 0; u_prev = 0; v_prev = 0; kc = 0; fu = 0; fv = 0
  do j = 1 - nh, ny + nh
    do i = 1 - nh, nx + nh
      x = (real(i, dp) - 0.5_dp) * dx
      y = (real(j, dp) - 0.5_dp) * dy
      c(i, j) = merge(1.0_dp, 0.0_dp, hypot(x - 0.5_dp, y - 0.5_dp) <= 0.25_dp)
    end do
  end do

  call fugu_hip_init(nx, ny, nh, lx, ly, .false., .false., 1000.0_dp, 1.0_dp, 1.0_dp, 0.01_dp, 1.0_dp, &
                     0.0_dp, 0.0_dp, 2_ip, 1.0e12_dp, 0.5_dp)
  call compute_mixture_densities(c, rho_c, rho_u, rho_v, nx, ny, nh)
  call compute_mixture_viscosities(c, kappa, zeta, eta, nx, ny, nh)

  open(newunit=unit, file=trim(path), access="stream", form="unformatted", status="replace")
  do step = 1, 2
    call advance_coupled_step(u, v, p, c, u_prev, v_prev, rho_c, rho_u, rho_v, kappa, zeta, eta, kc, fu, fv, &
                              1.0e-4_dp, step, nx, ny, nh, iter_out=it, res_out=res, use_bicgstab2=.false., &
                              tol_in=1.0e-8_dp, max_iter_in=40_ip)
    dot = coupled_dot(u, v, p, u, v, p, nx, ny, nh)
    write (unit) it, res, dot
    write (unit) u, v, p, c
  end do
  close(unit)
  call fugu_hip_finalize()
  print '(a)', "dropin_check: 2 steps through the Fortran drop-in module"
end program dropin_check
