! Test and baseline infrastructure (not part of the product): times the
! reference's own FP64 matrix-diagonalisation solver and writes its solution.
! Linked by oracle/build_ref.sh against the reference's sources
! (src/precision/precision.f90, src/class_solver/class_solver_coefficient_3d.f90,
! src/class_solver/class_solver_3d.f90), compiled where they lie.  Calls
! solver_coeffs_init_3d (class_solver_3d.f90:380) once and solve_3d
! (class_solver_3d.f90:83) nrep times, each on a fresh copy of the right-hand
! side (solve_3d folds the boundary values into it in place).
!
! usage: ref_solver_driver IN OUT
!   IN  (stream, native byte order): nx ny nz so nrep bct(6) (int32), sigma,
!       gx(nx), gy(ny), gz(nz), rhs(nx-2,ny-2,nz-2), bcx(ny-2,nz-2,2),
!       bcy(nx-2,nz-2,2), bcz(nx-2,ny-2,2) (real64, Fortran order)
!   OUT (stream): seconds per solve_3d call, seconds of solver_coeffs_init_3d,
!       sol(0:nx-1,0:ny-1,0:nz-1) after the last call
program ref_solver_driver
  use precision
  use class_solver_3d
  implicit none
  type(solver_coeffs_3d) :: sc
  integer(ik) :: nx, ny, nz, so, nrep, bct(6), n1, n2, n3, r
  real(rk) :: sigma, tsetup, tsolve
  real(rk), allocatable :: gx(:), gy(:), gz(:), rhs(:,:,:), rhs0(:,:,:), sol(:,:,:)
  real(rk), allocatable :: bcx(:,:,:), bcy(:,:,:), bcz(:,:,:)
  integer(int8) :: c0, c1, crate, ticks
  character(len=4096) :: fin, fout
  integer :: iu

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(iu) nx, ny, nz, so, nrep, bct
  read(iu) sigma
  n1 = nx - 2
  n2 = ny - 2
  n3 = nz - 2
  allocate(gx(nx), gy(ny), gz(nz))
  allocate(rhs(n1, n2, n3), rhs0(n1, n2, n3), sol(0:n1+1, 0:n2+1, 0:n3+1))
  allocate(bcx(n2, n3, 2), bcy(n1, n3, 2), bcz(n1, n2, 2))
  read(iu) gx, gy, gz, rhs0, bcx, bcy, bcz
  close(iu)

  call system_clock(c0, crate)
  call solver_coeffs_init_3d(gx, gy, gz, sc, nx, ny, nz, bct, so)
  call system_clock(c1)
  tsetup = real(c1 - c0, rk) / real(crate, rk)

  sol = 0._rk
  ticks = 0
  do r = 1, nrep
     rhs = rhs0
     call system_clock(c0)
     call solve_3d(sc, rhs, sigma, sol, bcx, bcy, bcz, n1, n2, n3)
     call system_clock(c1)
     ticks = ticks + (c1 - c0)
  end do
  tsolve = real(ticks, rk) / real(crate, rk) / real(max(nrep, 1), rk)

  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace', &
       action='write')
  write(iu) tsolve, tsetup, sol
  close(iu)
end program ref_solver_driver
