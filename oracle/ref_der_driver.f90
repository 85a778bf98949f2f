! Test infrastructure (not part of the product): the reference's own compact
! derivatives of a set of lines.  Linked by oracle/build_ref.sh against the
! reference's sources (src/precision/precision.f90,
! src/class_derivatives/class_derivatives_coefficient.f90,
! src/class_derivatives/class_derivatives.f90), compiled where they lie.
! Calls der_coeffs_init (class_derivatives.f90:111) once, then der, dder
! (:186, :224) and, for a solver-mode set-up, der_s, dder_s (:205, :243) on
! every line.
!
! usage: ref_der_driver IN OUT
!   IN  (stream, native byte order): n so solver nlines (int32), grid(n),
!       x(n, nlines) (real64, Fortran order)
!   OUT (stream): der(n, nlines), dder(n, nlines) and, when solver /= 0,
!       der_s(n, nlines), dder_s(n, nlines) (their first and last points 0)
program ref_der_driver
  use precision
  use class_derivatives
  implicit none
  type(derivatives_coefficients) :: dc
  integer(ik) :: n, so, solver, nl, l
  real(rk), allocatable :: grid(:), x(:,:), d1(:,:), d2(:,:), s1(:,:), s2(:,:)
  character(len=4096) :: fin, fout
  integer :: iu

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(iu) n, so, solver, nl
  allocate(grid(n), x(n, nl), d1(n, nl), d2(n, nl), s1(n, nl), s2(n, nl))
  read(iu) grid, x
  close(iu)

  call der_coeffs_init(grid, dc, n, solver /= 0, so)
  d1 = 0._rk
  d2 = 0._rk
  s1 = 0._rk
  s2 = 0._rk
  do l = 1, nl
     call der(dc, x(:, l), d1(:, l))
     call dder(dc, x(:, l), d2(:, l))
     if (solver /= 0) then
        call der_s(dc, x(:, l), s1(:, l))
        call dder_s(dc, x(:, l), s2(:, l))
     end if
  end do

  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace', &
       action='write')
  write(iu) d1, d2
  if (solver /= 0) write(iu) s1, s2
  close(iu)
end program ref_der_driver
