"""Writes the fixtures that pin the checkpoint and VTK restatements to a real Fortran runtime.

A small Fortran program of this build (not the reference's source) is compiled with flang and run:
it reads the values in fortran_format_values.bin (native float64) and writes them with the edit
descriptors the reference's VTK writer uses (src/mod_io.f90:109-160: '(ES15.7)', '(3ES15.7)',
'(A,3I8)', '(A,3ES15.7)', '(A,I10)') into fortran_format_out.txt, and writes a 4 x 3 checkpoint
with the reference's stream layout (src/mod_io.f90:169-187: int32 step, real64 t, int32 nx, ny,
nh, then u, v, u_prev, v_prev, p, C as whole padded arrays) into fortran_checkpoint_4x3.bin.
Run from this directory: python make_io_fixtures.py (needs flang; the fixtures are committed).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PROGRAM = r"""
program io_fixtures
  use, intrinsic :: iso_fortran_env, only: real64, int32
  implicit none
  integer, parameter :: nx = 4, ny = 3, nh = 2
  real(real64), allocatable :: vals(:)
  real(real64) :: u(0-nh:nx+nh, 1-nh:ny+nh), v(1-nh:nx+nh, 0-nh:ny+nh)
  real(real64) :: up(0-nh:nx+nh, 1-nh:ny+nh), vp(1-nh:nx+nh, 0-nh:ny+nh)
  real(real64) :: p(1-nh:nx+nh, 1-nh:ny+nh), c(1-nh:nx+nh, 1-nh:ny+nh)
  integer(int32) :: n, k, unit_in, unit_out, step
  real(real64) :: t
  character(len=*), parameter :: dir = '@DIR@'
  open(newunit=unit_in, file=dir//'/fortran_format_values.bin', form='unformatted', access='stream', &
       status='old', action='read')
  read(unit_in) n
  allocate(vals(n))
  read(unit_in) vals
  close(unit_in)
  open(newunit=unit_out, file=dir//'/fortran_format_out.txt', status='replace', action='write')
  do k = 1, n
    write(unit_out, '(ES15.7)') vals(k)
  end do
  do k = 1, n - 2, 3
    write(unit_out, '(3ES15.7)') vals(k), vals(k + 1), vals(k + 2)
  end do
  write(unit_out, '(A,3I8)') 'DIMENSIONS ', 128, 2048, 1
  write(unit_out, '(A,3ES15.7)') 'ORIGIN ', 0.5d0 / 128, 0.5d0 / 96, 0.0d0
  write(unit_out, '(A,3ES15.7)') 'SPACING ', 1.0d0 / 128, 1.0d0 / 96, 1.0d0 / 128
  write(unit_out, '(A,I10)') 'POINT_DATA ', 4194304
  write(unit_out, '(A,I10)') 'POINT_DATA ', 12
  close(unit_out)
  k = 0
  call fill(u, k); call fill(v, k); call fill(up, k); call fill(vp, k); call fill(p, k); call fill(c, k)
  step = 10
  t = 0.1d0
  open(newunit=unit_out, file=dir//'/fortran_checkpoint_4x3.bin', form='unformatted', access='stream', &
       status='replace', action='write')
  write(unit_out) step, t, nx, ny, nh
  write(unit_out) u, v, up, vp, p, c
  close(unit_out)
contains
  subroutine fill(a, k)
    real(real64), intent(inout) :: a(:,:)
    integer(int32), intent(inout) :: k
    integer :: i, j
    do j = 1, size(a, 2)
      do i = 1, size(a, 1)
        k = k + 1
        a(i, j) = sin(0.37d0 * k) * 10.0d0 ** mod(k, 7) - 0.5d0 * k
      end do
    end do
  end subroutine fill
end program io_fixtures
"""


def values():
    rng = np.random.default_rng(2024)
    v = [0.0, -0.0, 1.0, -1.0, 0.5, 0.1, 1.0 / 3.0, 2.0 / 3.0, 9.99999995e-3, 9.9999999e-3, 9.99999994999e-3,
         1.23456785, 1.234567850001, 1e-100, -1e-100, 1.5e100, 9.9999999e99, 1e-310, -4.9e-324, 1.7976931348623157e308,
         123456789.0, 0.00390625, 0.0103412, 4.0, 1e-8, -2.5e-7, 6.02214076e23, np.inf, -np.inf, np.nan]
    v += list(rng.standard_normal(24) * 10.0 ** rng.integers(-12, 12, 24))
    return np.array(v, dtype=np.float64)


def main():
    fc = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    vals = values()
    with open(os.path.join(HERE, "fortran_format_values.bin"), "wb") as f:
        f.write(np.int32(len(vals)).tobytes())
        f.write(vals.tobytes())
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "io_fixtures.f90")
        open(src, "w").write(PROGRAM.replace("@DIR@", HERE))
        exe = os.path.join(td, "io_fixtures")
        subprocess.run([fc, "-O0", src, "-o", exe], check=True, cwd=td)
        subprocess.run([exe], check=True, cwd=td)
    print("wrote fortran_format_values.bin, fortran_format_out.txt, fortran_checkpoint_4x3.bin")


if __name__ == "__main__":
    main()
