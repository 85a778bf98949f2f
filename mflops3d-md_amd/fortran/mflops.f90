! ISO_C_BINDING front end of libmflops for Fortran callers.
!
! Drop-in counterparts of the reference's module procedures (paths relative
! to the reference root):
!   module mflops_derivatives  <- class_derivatives   src/class_derivatives/class_derivatives.f90
!     der_coeffs_init(grid,dc,n,solve,so)          :111
!     der, dder, der_s, dder_s(dc,x,dx)            :186-260
!     derx_array .. dderz_array(dc,f,df,nx,ny,nz)  <- derx..dderz, src/class_field/class_field.f90:629-978
!   module mflops_solver_3d    <- class_solver_3d     src/class_solver/class_solver_3d.f90
!     solver_coeffs_init_3d(grid1,grid2,grid3,sc,n1,n2,n3,bct,so)   :380
!     solve_3d(sc,rhs,sigma,solf,bcx,bcy,bcz,n1,n2,n3)              :83
!     solver_3d_array(sc,g,u,bcx,bcy,bcz,sigma[,precision])         <- solver_3d, class_field.f90:1072
! Arrays are the caller's host arrays (the library stages them through HBM).
! Errors stop the program with the library's message, as the reference's
! own setup stops on a complex spectrum (class_solver_3d.f90:537-540).

module mflops_c
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: MF_OK = 0, MF_FP64 = 0, MF_MIXED = 1, MF_MIXED_BF16X3 = 2, &
                                MF_MIXED_FP16X2 = 3

  type, bind(c) :: mf_solve_report
     integer(c_int) :: iterations
     real(c_double) :: residual
     real(c_double) :: residual0
  end type mf_solve_report

  interface
     function mf_last_error() bind(c, name="mf_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function mf_last_error
     function mf_der_create(grid, n, solver_mode, so, h) bind(c, name="mf_der_create") result(rc)
       import :: c_double, c_int, c_ptr
       real(c_double), intent(in) :: grid(*)
       integer(c_int), value :: n, solver_mode, so
       type(c_ptr), intent(out) :: h
       integer(c_int) :: rc
     end function mf_der_create
     function mf_der_destroy(h) bind(c, name="mf_der_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int) :: rc
     end function mf_der_destroy
     function mf_der3d_host(h, axis, order, x, out, nx, ny, nz) bind(c, name="mf_der3d_host") result(rc)
       import :: c_double, c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: axis, order, nx, ny, nz
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(inout) :: out(*)
       integer(c_int) :: rc
     end function mf_der3d_host
     function mf_solver3d_create(gx, nx, gy, ny, gz, nz, bct, so, h) &
          bind(c, name="mf_solver3d_create") result(rc)
       import :: c_double, c_int, c_ptr
       real(c_double), intent(in) :: gx(*), gy(*), gz(*)
       integer(c_int), value :: nx, ny, nz, so
       integer(c_int), intent(in) :: bct(6)
       type(c_ptr), intent(out) :: h
       integer(c_int) :: rc
     end function mf_solver3d_create
     function mf_solver3d_destroy(h) bind(c, name="mf_solver3d_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int) :: rc
     end function mf_solver3d_destroy
     function mf_solver3d_solve_host(h, g, sigma, u, bcx, bcy, bcz, precision, tol, max_iter, rep) &
          bind(c, name="mf_solver3d_solve_host") result(rc)
       import :: c_double, c_int, c_ptr, mf_solve_report
       type(c_ptr), value :: h
       real(c_double), intent(in) :: g(*), bcx(*), bcy(*), bcz(*)
       real(c_double), value :: sigma, tol
       real(c_double), intent(inout) :: u(*)
       integer(c_int), value :: precision, max_iter
       type(mf_solve_report), intent(out) :: rep
       integer(c_int) :: rc
     end function mf_solver3d_solve_host
     function c_strlen(s) bind(c, name="strlen") result(n)
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: n
     end function c_strlen
  end interface

contains

  subroutine mf_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == MF_OK) return
    p = mf_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    print '(a,a,i0,a,*(a))', where, ': mflops error ', rc, ': ', msg
    stop 1
  end subroutine mf_check

end module mflops_c


module mflops_derivatives
  use iso_c_binding
  use mflops_c
  implicit none
  private

  ! type(derivatives_coefficients) (class_derivatives.f90:30-45)
  type, public :: derivatives_coefficients
     type(c_ptr) :: h = c_null_ptr
     integer :: n = 0
     logical :: solver = .false.
  end type derivatives_coefficients

  public :: der_coeffs_init, der, dder, der_s, dder_s, der_coeffs_destroy
  public :: derx_array, dderx_array, dery_array, ddery_array, derz_array, dderz_array

contains

  subroutine der_coeffs_init(grid, dc, n, solve, so)
    integer, intent(in) :: n, so
    real(c_double), intent(in) :: grid(n)
    type(derivatives_coefficients), intent(inout) :: dc
    logical, intent(in) :: solve
    integer(c_int) :: mode
    if (c_associated(dc%h)) call der_coeffs_destroy(dc)
    mode = 0
    if (solve) mode = 1
    call mf_check(mf_der_create(grid, int(n, c_int), mode, int(so, c_int), dc%h), 'der_coeffs_init')
    dc%n = n
    dc%solver = solve
  end subroutine der_coeffs_init

  subroutine der_coeffs_destroy(dc)
    type(derivatives_coefficients), intent(inout) :: dc
    integer(c_int) :: rc
    if (c_associated(dc%h)) rc = mf_der_destroy(dc%h)
    dc%h = c_null_ptr
  end subroutine der_coeffs_destroy

  ! One line through the 3-D x-driver: in a (n,3,3) stack the middle line
  ! (2,2) is interior and gets the full scheme; line (1,2) lies in one
  ! boundary plane and gets the sub-line scheme of a solver-mode handle.
  subroutine line(dc, x, dx, order, sub)
    type(derivatives_coefficients), intent(in) :: dc
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: dx(:)
    integer, intent(in) :: order
    logical, intent(in) :: sub
    real(c_double), allocatable :: f(:,:,:), d(:,:,:)
    integer :: n
    n = dc%n
    if (sub .and. .not. dc%solver) then
       print '(a)', 'der_s/dder_s need coefficients built with solve=.true.'
       stop 1
    end if
    allocate(f(n,3,3), d(n,3,3))
    f = 0.0_c_double
    if (sub) then
       f(:,1,2) = x(1:n)
    else
       f(:,2,2) = x(1:n)
    end if
    call mf_check(mf_der3d_host(dc%h, 0_c_int, int(order, c_int), f, d, int(n, c_int), 3_c_int, &
         3_c_int), 'der')
    if (sub) then
       dx(2:n-1) = d(2:n-1,1,2)
    else
       dx(1:n) = d(:,2,2)
    end if
  end subroutine line

  subroutine der(dc, x, dx)
    type(derivatives_coefficients), intent(in) :: dc
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: dx(:)
    call line(dc, x, dx, 1, .false.)
  end subroutine der

  subroutine dder(dc, x, dx)
    type(derivatives_coefficients), intent(in) :: dc
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: dx(:)
    call line(dc, x, dx, 2, .false.)
  end subroutine dder

  subroutine der_s(dc, x, dx)
    type(derivatives_coefficients), intent(in) :: dc
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: dx(:)
    call line(dc, x, dx, 1, .true.)
  end subroutine der_s

  subroutine dder_s(dc, x, dx)
    type(derivatives_coefficients), intent(in) :: dc
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: dx(:)
    call line(dc, x, dx, 2, .true.)
  end subroutine dder_s

  subroutine axis_der(dc, f, df, nx, ny, nz, axis, order)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz, axis, order
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call mf_check(mf_der3d_host(dc%h, int(axis, c_int), int(order, c_int), f, df, &
         int(nx, c_int), int(ny, c_int), int(nz, c_int)), 'derivative')
  end subroutine axis_der

  subroutine derx_array(dc, f, df, nx, ny, nz)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call axis_der(dc, f, df, nx, ny, nz, 0, 1)
  end subroutine derx_array

  subroutine dderx_array(dc, f, df, nx, ny, nz)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call axis_der(dc, f, df, nx, ny, nz, 0, 2)
  end subroutine dderx_array

  subroutine dery_array(dc, f, df, nx, ny, nz)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call axis_der(dc, f, df, nx, ny, nz, 1, 1)
  end subroutine dery_array

  subroutine ddery_array(dc, f, df, nx, ny, nz)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call axis_der(dc, f, df, nx, ny, nz, 1, 2)
  end subroutine ddery_array

  subroutine derz_array(dc, f, df, nx, ny, nz)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call axis_der(dc, f, df, nx, ny, nz, 2, 1)
  end subroutine derz_array

  subroutine dderz_array(dc, f, df, nx, ny, nz)
    type(derivatives_coefficients), intent(in) :: dc
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in) :: f(nx,ny,nz)
    real(c_double), intent(inout) :: df(nx,ny,nz)
    call axis_der(dc, f, df, nx, ny, nz, 2, 2)
  end subroutine dderz_array

end module mflops_derivatives


module mflops_solver_3d
  use iso_c_binding
  use mflops_c
  implicit none
  private

  ! type(solver_coeffs_3d) (class_solver_3d.f90:47-50)
  type, public :: solver_coeffs_3d
     type(c_ptr) :: h = c_null_ptr
     integer :: nx = 0, ny = 0, nz = 0
  end type solver_coeffs_3d

  public :: solver_coeffs_init_3d, solve_3d, solver_3d_array, solver_coeffs_destroy_3d
  public :: MF_FP64, MF_MIXED, MF_MIXED_BF16X3, MF_MIXED_FP16X2, mf_solve_report

contains

  ! grids are full (boundary points included), as in the reference
  subroutine solver_coeffs_init_3d(grid1, grid2, grid3, sc, n1, n2, n3, bct, so)
    integer, intent(in) :: n1, n2, n3, so
    real(c_double), intent(in) :: grid1(n1), grid2(n2), grid3(n3)
    type(solver_coeffs_3d), intent(inout) :: sc
    integer, intent(in) :: bct(6)
    integer(c_int) :: b(6)
    if (c_associated(sc%h)) call solver_coeffs_destroy_3d(sc)
    b = int(bct, c_int)
    call mf_check(mf_solver3d_create(grid1, int(n1, c_int), grid2, int(n2, c_int), grid3, &
         int(n3, c_int), b, int(so, c_int), sc%h), 'solver_coeffs_init_3d')
    sc%nx = n1
    sc%ny = n2
    sc%nz = n3
  end subroutine solver_coeffs_init_3d

  subroutine solver_coeffs_destroy_3d(sc)
    type(solver_coeffs_3d), intent(inout) :: sc
    integer(c_int) :: rc
    if (c_associated(sc%h)) rc = mf_solver3d_destroy(sc%h)
    sc%h = c_null_ptr
  end subroutine solver_coeffs_destroy_3d

  ! solve_3d(sc,rhs,sigma,solf,bcx,bcy,bcz,n1,n2,n3) with interior sizes
  subroutine solve_3d(sc, rhs, sigma, solf, bcx, bcy, bcz, n1, n2, n3)
    type(solver_coeffs_3d), intent(in) :: sc
    integer, intent(in) :: n1, n2, n3
    real(c_double), intent(inout) :: rhs(n1,n2,n3)
    real(c_double), intent(in) :: sigma
    real(c_double), intent(inout) :: solf(0:n1+1,0:n2+1,0:n3+1)
    real(c_double), intent(in) :: bcx(n2,n3,2), bcy(n1,n3,2), bcz(n1,n2,2)
    real(c_double), allocatable :: g(:,:,:)
    type(mf_solve_report) :: rep
    allocate(g(0:n1+1,0:n2+1,0:n3+1))
    g = 0.0_c_double
    g(1:n1,1:n2,1:n3) = rhs
    call mf_check(mf_solver3d_solve_host(sc%h, g, sigma, solf, bcx, bcy, bcz, MF_FP64, &
         0.0_c_double, 0_c_int, rep), 'solve_3d')
  end subroutine solve_3d

  ! solver_3d on full-size arrays (class_field.f90:1072-1109)
  subroutine solver_3d_array(sc, g, u, bcx, bcy, bcz, sigma, precision, rep)
    type(solver_coeffs_3d), intent(in) :: sc
    real(c_double), intent(in) :: g(sc%nx,sc%ny,sc%nz)
    real(c_double), intent(inout) :: u(sc%nx,sc%ny,sc%nz)
    real(c_double), intent(in) :: bcx(sc%ny-2,sc%nz-2,2), bcy(sc%nx-2,sc%nz-2,2), bcz(sc%nx-2,sc%ny-2,2)
    real(c_double), intent(in) :: sigma
    integer, intent(in), optional :: precision
    type(mf_solve_report), intent(out), optional :: rep
    type(mf_solve_report) :: r
    integer(c_int) :: p
    p = MF_FP64
    if (present(precision)) p = int(precision, c_int)
    call mf_check(mf_solver3d_solve_host(sc%h, g, sigma, u, bcx, bcy, bcz, p, 0.0_c_double, &
         0_c_int, r), 'solver_3d')
    if (present(rep)) rep = r
  end subroutine solver_3d_array

end module mflops_solver_3d
