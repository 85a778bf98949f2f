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
!   module mflops_io           <- class_io            src/class_io/class_io.f90
!     write_var3d(file_name,dim_name,dim_len,var_name,var[,dbl])    :23-124
!     read_var3d(file_name,var_name,var)                           :126-189
!     get_var3d_info(file_name,var_name,dim_name,dim_len)          :191-223
!   module mflops_navier       <- class_navier_3D     src/class_navier/class_navier_3D.f90
!     navier_initialization(nav,gx,gy,gz,re,dt[,tou,top,nlt,so,pt,nsubite]) :2084-2345
!     navier_set_bc / navier_set_state / navier_set_flow_rate / navier_set_les,
!     navier_step(nav,nsteps[,precision]) (navier3d.f90:48-286), navier_get,
!     restart_write(nav[,dirname]) :186-257, restart_read(nav[,dirname]) :259-340
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
     function mf_der_lines_host(h, order, sub, x, out, nlines, ld) bind(c, name="mf_der_lines_host") &
          result(rc)
       import :: c_double, c_int, c_ptr, c_long_long
       type(c_ptr), value :: h
       integer(c_int), value :: order, sub, nlines
       integer(c_long_long), value :: ld
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(inout) :: out(*)
       integer(c_int) :: rc
     end function mf_der_lines_host
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

  ! One line (der/dder and the sub-line der_s/dder_s, which leaves dx(1) and
  ! dx(n) untouched) through mf_der_lines_host: the handle's staging buffer
  ! is reused, so a loop over the lines of a field allocates nothing.
  subroutine line(dc, x, dx, order, sub)
    type(derivatives_coefficients), intent(in) :: dc
    real(c_double), intent(in) :: x(:)
    real(c_double), intent(inout) :: dx(:)
    integer, intent(in) :: order
    logical, intent(in) :: sub
    integer(c_int) :: s
    if (sub .and. .not. dc%solver) then
       print '(a)', 'der_s/dder_s need coefficients built with solve=.true.'
       stop 1
    end if
    s = 0
    if (sub) s = 1
    call mf_check(mf_der_lines_host(dc%h, int(order, c_int), s, x, dx, 1_c_int, &
         int(dc%n, c_long_long)), 'der')
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


module mflops_io
  use iso_c_binding
  use mflops_c
  implicit none
  private
  public :: write_var3d, read_var3d, get_var3d_info

  interface
     function mf_io_write_field(path, var_name, dim_names, nx, ny, nz, f, dbl) &
          bind(c, name="mf_io_write_field") result(rc)
       import :: c_char, c_int, c_double, c_ptr
       character(kind=c_char), intent(in) :: path(*), var_name(*)
       type(c_ptr), intent(in) :: dim_names(3)
       integer(c_int), value :: nx, ny, nz, dbl
       real(c_double), intent(in) :: f(*)
       integer(c_int) :: rc
     end function mf_io_write_field
     function mf_io_field_info(path, var_name, dims, dim_names, name_len, is_double) &
          bind(c, name="mf_io_field_info") result(rc)
       import :: c_char, c_int
       character(kind=c_char), intent(in) :: path(*), var_name(*)
       integer(c_int), intent(out) :: dims(3)
       character(kind=c_char), intent(out) :: dim_names(*)
       integer(c_int), value :: name_len
       integer(c_int), intent(out) :: is_double
       integer(c_int) :: rc
     end function mf_io_field_info
     function mf_io_read_field(path, var_name, nx, ny, nz, f) bind(c, name="mf_io_read_field") result(rc)
       import :: c_char, c_int, c_double
       character(kind=c_char), intent(in) :: path(*), var_name(*)
       integer(c_int), value :: nx, ny, nz
       real(c_double), intent(out) :: f(*)
       integer(c_int) :: rc
     end function mf_io_read_field
  end interface

contains

  ! write_var3d(file_name,dim_name,dim_len,var_name,var,dbl): FP32 unless dbl
  ! is present, as the reference; an existing file gains the variable
  subroutine write_var3d(file_name, dim_name, dim_len, var_name, var, dbl)
    character(len=*), intent(in) :: file_name, var_name, dim_name(3)
    integer, intent(in) :: dim_len(3)
    real(c_double), intent(in) :: var(dim_len(1), dim_len(2), dim_len(3))
    character(*), intent(in), optional :: dbl
    character(kind=c_char, len=513), target :: dn(3)
    type(c_ptr) :: dp(3)
    integer :: a
    integer(c_int) :: d
    do a = 1, 3
       dn(a) = trim(dim_name(a)) // c_null_char
       dp(a) = c_loc(dn(a))
    end do
    d = 0
    if (present(dbl)) d = 1
    call mf_check(mf_io_write_field(trim(file_name) // c_null_char, trim(var_name) // c_null_char, &
         dp, int(dim_len(1), c_int), int(dim_len(2), c_int), int(dim_len(3), c_int), var, d), &
         'write_var3d')
  end subroutine write_var3d

  subroutine get_var3d_info(file_name, var_name, dim_name, dim_len)
    character(len=*), intent(in) :: file_name, var_name
    character(len=*), intent(out) :: dim_name(3)
    integer, intent(out) :: dim_len(3)
    integer(c_int) :: dims(3), dbl
    character(kind=c_char) :: names(3 * 512)
    integer :: a, q
    call mf_check(mf_io_field_info(trim(file_name) // c_null_char, trim(var_name) // c_null_char, &
         dims, names, 512_c_int, dbl), 'get_var3d_info')
    dim_len = int(dims)
    do a = 1, 3
       dim_name(a) = ''
       do q = 1, min(512, len(dim_name(a)))
          if (names((a - 1) * 512 + q) == c_null_char) exit
          dim_name(a)(q:q) = names((a - 1) * 512 + q)
       end do
    end do
  end subroutine get_var3d_info

  subroutine read_var3d(file_name, var_name, var)
    character(len=*), intent(in) :: file_name, var_name
    real(c_double), intent(out) :: var(:,:,:)
    call mf_check(mf_io_read_field(trim(file_name) // c_null_char, trim(var_name) // c_null_char, &
         int(size(var, 1), c_int), int(size(var, 2), c_int), int(size(var, 3), c_int), var), &
         'read_var3d')
  end subroutine read_var3d

end module mflops_io


module mflops_navier
  use iso_c_binding
  use mflops_c
  implicit none
  private

  ! the NS state lives in HBM; host arrays are staged through `stage`
  type, public :: navier3d
     type(c_ptr) :: h = c_null_ptr
     type(c_ptr) :: stage(5) = c_null_ptr
     integer :: nx = 0, ny = 0, nz = 0
  end type navier3d

  public :: navier_initialization, navier_finalize, navier_set_bc, navier_set_state, &
       navier_set_flow_rate, navier_set_les, navier_step, navier_get, navier_time, &
       restart_write, restart_read

  integer(c_int), parameter :: H2D = 1, D2H = 2

  interface
     function hipMalloc(p, bytes) bind(c, name="hipMalloc") result(rc)
       import :: c_ptr, c_size_t, c_int
       type(c_ptr) :: p
       integer(c_size_t), value :: bytes
       integer(c_int) :: rc
     end function hipMalloc
     function hipFree(p) bind(c, name="hipFree") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: p
       integer(c_int) :: rc
     end function hipFree
     function hipMemcpy(dst, src, bytes, kind) bind(c, name="hipMemcpy") result(rc)
       import :: c_ptr, c_size_t, c_int
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
       integer(c_int), value :: kind
       integer(c_int) :: rc
     end function hipMemcpy
     function mf_ns_create(gx, nx, gy, ny, gz, nz, re, dt, tou, top, nlt, so, h) &
          bind(c, name="mf_ns_create") result(rc)
       import :: c_double, c_int, c_ptr
       real(c_double), intent(in) :: gx(*), gy(*), gz(*)
       integer(c_int), value :: nx, ny, nz, tou, top, nlt, so
       real(c_double), value :: re, dt
       type(c_ptr) :: h
       integer(c_int) :: rc
     end function mf_ns_create
     function mf_ns_destroy(h) bind(c, name="mf_ns_destroy") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: h
       integer(c_int) :: rc
     end function mf_ns_destroy
     function mf_ns_set_bc(h, c, bcx, bcy, bcz, st) bind(c, name="mf_ns_set_bc") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: h, bcx, bcy, bcz, st
       integer(c_int), value :: c
       integer(c_int) :: rc
     end function mf_ns_set_bc
     function mf_ns_set_state(h, u, v, w, st) bind(c, name="mf_ns_set_state") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: h, u, v, w, st
       integer(c_int) :: rc
     end function mf_ns_set_state
     function mf_ns_set_flow_rate(h, enable, g0, st) bind(c, name="mf_ns_set_flow_rate") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, st
       integer(c_int), value :: enable
       real(c_double), value :: g0
       integer(c_int) :: rc
     end function mf_ns_set_flow_rate
     function mf_ns_set_les(h, les_type, c, st) bind(c, name="mf_ns_set_les") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, st
       integer(c_int), value :: les_type
       real(c_double), value :: c
       integer(c_int) :: rc
     end function mf_ns_set_les
     function mf_ns_set_scheme(h, pt, nsubite, st) bind(c, name="mf_ns_set_scheme") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: h, st
       integer(c_int), value :: pt, nsubite
       integer(c_int) :: rc
     end function mf_ns_set_scheme
     function mf_ns_step(h, nsteps, precision, tol, st) bind(c, name="mf_ns_step") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, st
       integer(c_int), value :: nsteps, precision
       real(c_double), value :: tol
       integer(c_int) :: rc
     end function mf_ns_step
     function mf_ns_get(h, u, v, w, p, phi, st) bind(c, name="mf_ns_get") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: h, u, v, w, p, phi, st
       integer(c_int) :: rc
     end function mf_ns_get
     function mf_ns_info(h, time, flow_force, st) bind(c, name="mf_ns_info") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, st
       real(c_double), intent(out) :: time, flow_force
       integer(c_int) :: rc
     end function mf_ns_info
     function mf_ns_restart_write(h, dirname, dir_out, dir_out_len, st) &
          bind(c, name="mf_ns_restart_write") result(rc)
       import :: c_ptr, c_int, c_char
       type(c_ptr), value :: h, dirname, st
       character(kind=c_char), intent(out) :: dir_out(*)
       integer(c_int), value :: dir_out_len
       integer(c_int) :: rc
     end function mf_ns_restart_write
     function mf_ns_restart_read(h, dirname, st) bind(c, name="mf_ns_restart_read") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: h, dirname, st
       integer(c_int) :: rc
     end function mf_ns_restart_read
  end interface

contains

  subroutine hip_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
       print '(a,a,i0)', where, ': HIP error ', rc
       stop 1
    end if
  end subroutine hip_check

  ! navier_initialization: grids (full), Re, dt, BDF order tou, pressure
  ! extrapolation top, nonlinear form nlt, scheme order so, projection type
  ! pt and sub-iterations nsubite (defaults 2, 2, 2, 6, 1, 1)
  subroutine navier_initialization(nav, gx, gy, gz, re, dt, tou, top, nlt, so, pt, nsubite)
    type(navier3d), intent(inout) :: nav
    real(c_double), intent(in) :: gx(:), gy(:), gz(:), re, dt
    integer, intent(in), optional :: tou, top, nlt, so, pt, nsubite
    integer(c_int) :: p1, p2
    integer(c_int) :: t1, t2, t3, t4
    integer :: q
    integer(c_size_t) :: bytes
    t1 = 2; t2 = 2; t3 = 2; t4 = 6
    if (present(tou)) t1 = int(tou, c_int)
    if (present(top)) t2 = int(top, c_int)
    if (present(nlt)) t3 = int(nlt, c_int)
    if (present(so)) t4 = int(so, c_int)
    call navier_finalize(nav)
    nav%nx = size(gx); nav%ny = size(gy); nav%nz = size(gz)
    call mf_check(mf_ns_create(gx, int(nav%nx, c_int), gy, int(nav%ny, c_int), gz, int(nav%nz, c_int), &
         re, dt, t1, t2, t3, t4, nav%h), 'navier_initialization')
    p1 = 1; p2 = 1
    if (present(pt)) p1 = int(pt, c_int)
    if (present(nsubite)) p2 = int(nsubite, c_int)
    if (p1 /= 1 .or. p2 /= 1) &
         call mf_check(mf_ns_set_scheme(nav%h, p1, p2, c_null_ptr), 'navier_initialization')
    bytes = int(nav%nx, c_size_t) * nav%ny * nav%nz * 8
    do q = 1, 5
       call hip_check(hipMalloc(nav%stage(q), bytes), 'navier_initialization')
    end do
  end subroutine navier_initialization

  subroutine navier_finalize(nav)
    type(navier3d), intent(inout) :: nav
    integer(c_int) :: rc
    integer :: q
    if (c_associated(nav%h)) rc = mf_ns_destroy(nav%h)
    nav%h = c_null_ptr
    do q = 1, 5
       if (c_associated(nav%stage(q))) rc = hipFree(nav%stage(q))
       nav%stage(q) = c_null_ptr
    end do
  end subroutine navier_finalize

  subroutine put(nav, q, a, n)
    type(navier3d), intent(inout) :: nav
    integer, intent(in) :: q, n
    real(c_double), intent(in), target :: a(n)
    call hip_check(hipMemcpy(nav%stage(q), c_loc(a), int(n, c_size_t) * 8, H2D), 'navier (upload)')
  end subroutine put

  subroutine fetch(nav, q, a, n)
    type(navier3d), intent(inout) :: nav
    integer, intent(in) :: q, n
    real(c_double), intent(inout), target :: a(n)
    call hip_check(hipMemcpy(c_loc(a), nav%stage(q), int(n, c_size_t) * 8, D2H), 'navier (download)')
  end subroutine fetch

  ! velocity boundary values of component c (1, 2, 3 = u, v, w), as
  ! type(boundary_condition): bcx(ny-2,nz-2,2), bcy(nx-2,nz-2,2), bcz(nx-2,ny-2,2)
  subroutine navier_set_bc(nav, c, bcx, bcy, bcz)
    type(navier3d), intent(inout) :: nav
    integer, intent(in) :: c
    real(c_double), intent(in) :: bcx(:,:,:), bcy(:,:,:), bcz(:,:,:)
    call put(nav, 1, bcx, size(bcx))
    call put(nav, 2, bcy, size(bcy))
    call put(nav, 3, bcz, size(bcz))
    call mf_check(mf_ns_set_bc(nav%h, int(c - 1, c_int), nav%stage(1), nav%stage(2), nav%stage(3), &
         c_null_ptr), 'navier_set_bc')
  end subroutine navier_set_bc

  subroutine navier_set_state(nav, u, v, w)
    type(navier3d), intent(inout) :: nav
    real(c_double), intent(in) :: u(:,:,:), v(:,:,:), w(:,:,:)
    call put(nav, 1, u, size(u))
    call put(nav, 2, v, size(v))
    call put(nav, 3, w, size(w))
    call mf_check(mf_ns_set_state(nav%h, nav%stage(1), nav%stage(2), nav%stage(3), c_null_ptr), &
         'navier_set_state')
  end subroutine navier_set_state

  subroutine navier_set_flow_rate(nav, enable, g0)
    type(navier3d), intent(inout) :: nav
    logical, intent(in) :: enable
    real(c_double), intent(in) :: g0
    integer(c_int) :: e
    e = 0
    if (enable) e = 1
    call mf_check(mf_ns_set_flow_rate(nav%h, e, g0, c_null_ptr), 'navier_set_flow_rate')
  end subroutine navier_set_flow_rate

  ! les_type 0..3 and the command-line constant les_c (squared inside)
  subroutine navier_set_les(nav, les_type, les_c)
    type(navier3d), intent(inout) :: nav
    integer, intent(in) :: les_type
    real(c_double), intent(in) :: les_c
    call mf_check(mf_ns_set_les(nav%h, int(les_type, c_int), les_c, c_null_ptr), 'navier_set_les')
  end subroutine navier_set_les

  subroutine navier_step(nav, nsteps, precision, tol)
    type(navier3d), intent(inout) :: nav
    integer, intent(in) :: nsteps
    integer, intent(in), optional :: precision
    real(c_double), intent(in), optional :: tol
    integer(c_int) :: p
    real(c_double) :: t
    p = MF_FP64
    t = 1.0e-12_c_double
    if (present(precision)) p = int(precision, c_int)
    if (present(tol)) t = tol
    call mf_check(mf_ns_step(nav%h, int(nsteps, c_int), p, t, c_null_ptr), 'navier_step')
  end subroutine navier_step

  ! the latest time level
  subroutine navier_get(nav, u, v, w, p, phi)
    type(navier3d), intent(inout) :: nav
    real(c_double), intent(inout), optional :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), phi(:,:,:)
    call mf_check(mf_ns_get(nav%h, nav%stage(1), nav%stage(2), nav%stage(3), nav%stage(4), &
         nav%stage(5), c_null_ptr), 'navier_get')
    if (present(u)) call fetch(nav, 1, u, size(u))
    if (present(v)) call fetch(nav, 2, v, size(v))
    if (present(w)) call fetch(nav, 3, w, size(w))
    if (present(p)) call fetch(nav, 4, p, size(p))
    if (present(phi)) call fetch(nav, 5, phi, size(phi))
  end subroutine navier_get

  function navier_time(nav) result(t)
    type(navier3d), intent(in) :: nav
    real(c_double) :: t, g
    call mf_check(mf_ns_info(nav%h, t, g, c_null_ptr), 'navier_time')
  end function navier_time

  ! restart_write: directory dirname, or restart_<tn(time)> when absent
  subroutine restart_write(nav, dirname, written)
    type(navier3d), intent(inout) :: nav
    character(len=*), intent(in), optional :: dirname
    character(len=*), intent(out), optional :: written
    character(kind=c_char, len=1024), target :: d
    character(kind=c_char) :: out(1024)
    type(c_ptr) :: dp
    integer :: q
    dp = c_null_ptr
    if (present(dirname)) then
       d = trim(dirname) // c_null_char
       dp = c_loc(d)
    end if
    call mf_check(mf_ns_restart_write(nav%h, dp, out, 1024_c_int, c_null_ptr), 'restart_write')
    if (present(written)) then
       written = ''
       do q = 1, min(1024, len(written))
          if (out(q) == c_null_char) exit
          written(q:q) = out(q)
       end do
    end if
  end subroutine restart_write

  ! restart_read: from dirname, or ./restart when absent (as the reference)
  subroutine restart_read(nav, dirname)
    type(navier3d), intent(inout) :: nav
    character(len=*), intent(in), optional :: dirname
    character(kind=c_char, len=1024), target :: d
    type(c_ptr) :: dp
    dp = c_null_ptr
    if (present(dirname)) then
       d = trim(dirname) // c_null_char
       dp = c_loc(d)
    end if
    call mf_check(mf_ns_restart_read(nav%h, dp, c_null_ptr), 'restart_read')
  end subroutine restart_read

end module mflops_navier
