! Fortran caller of libmflops through the ISO_C_BINDING modules, following
! the reference's own drivers:
!   src/main/test_solver_3d.f90        u = x^9 + y^9 + xy, sigma = 2, all
!                                      Dirichlet, relative error err() :310-361
!   src/main/test_derivatives_omp.f90  sin(2 pi a x) + cos(2 pi b x) per axis,
!                                      a = 1, b = 2, solver-type coefficients
! Grids use the reference's asin/cos stretching (class_mesh.f90:181-200) on
! x,y in [-0.5,0.5], z in [-0.25,0.25].  Prints one "name value" line per
! result for the calling test to check.
program test_solver_3d_mf
  use iso_c_binding
  use mflops_solver_3d
  use mflops_derivatives
  implicit none
  integer :: n, nx, ny, nz, i, j, k, bct(6), axis, order
  real(c_double), allocatable :: gx(:), gy(:), gz(:)
  real(c_double), allocatable :: u(:,:,:), uex(:,:,:), g(:,:,:), um(:,:,:)
  real(c_double), allocatable :: bcx(:,:,:), bcy(:,:,:), bcz(:,:,:)
  real(c_double), allocatable :: f(:,:,:), df(:,:,:), dex(:,:,:)
  real(c_double) :: sigma, err64, errmx, diff, pi, e
  type(solver_coeffs_3d) :: sc
  type(derivatives_coefficients) :: dc
  type(mf_solve_report) :: rep
  character(len=32) :: arg

  n = 33
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     read(arg, *) n
  end if
  nx = n ; ny = n ; nz = n
  pi = 4.0_c_double * atan(1.0_c_double)
  allocate(gx(nx), gy(ny), gz(nz))
  call grid_asin(gx, nx, -0.5_c_double, 0.5_c_double)
  call grid_asin(gy, ny, -0.5_c_double, 0.5_c_double)
  call grid_asin(gz, nz, -0.25_c_double, 0.25_c_double)

  ! ---- Helmholtz solve (test_solver_3d.f90)
  sigma = 2.0_c_double
  bct = (/1,1,1,1,1,1/)
  allocate(u(nx,ny,nz), uex(nx,ny,nz), g(nx,ny,nz), um(nx,ny,nz))
  allocate(bcx(ny-2,nz-2,2), bcy(nx-2,nz-2,2), bcz(nx-2,ny-2,2))
  do k = 1, nz
     do j = 1, ny
        do i = 1, nx
           uex(i,j,k) = gx(i)**9 + gy(j)**9 + gx(i)*gy(j)
           g(i,j,k) = 72.0_c_double*gx(i)**7 + 72.0_c_double*gy(j)**7 + sigma*uex(i,j,k)
        end do
     end do
  end do
  bcx(:,:,1) = uex(1,2:ny-1,2:nz-1)  ; bcx(:,:,2) = uex(nx,2:ny-1,2:nz-1)
  bcy(:,:,1) = uex(2:nx-1,1,2:nz-1)  ; bcy(:,:,2) = uex(2:nx-1,ny,2:nz-1)
  bcz(:,:,1) = uex(2:nx-1,2:ny-1,1)  ; bcz(:,:,2) = uex(2:nx-1,2:ny-1,nz)
  call solver_coeffs_init_3d(gx, gy, gz, sc, nx, ny, nz, bct, 6)
  u = 0.0_c_double
  call solver_3d_array(sc, g, u, bcx, bcy, bcz, sigma)
  err64 = rel_err(u, uex)
  um = 0.0_c_double
  call solver_3d_array(sc, g, um, bcx, bcy, bcz, sigma, MF_MIXED, rep)
  errmx = rel_err(um, uex)
  diff = maxval(abs(um(2:nx-1,2:ny-1,2:nz-1) - u(2:nx-1,2:ny-1,2:nz-1))) / &
       maxval(abs(u(2:nx-1,2:ny-1,2:nz-1)))
  print '(a,es24.16)', 'solver_fp64_error ', err64
  print '(a,es24.16)', 'solver_mixed_error ', errmx
  print '(a,es24.16)', 'solver_mixed_vs_fp64 ', diff
  print '(a,es24.16)', 'solver_mixed_residual ', rep%residual
  call solver_coeffs_destroy_3d(sc)

  ! ---- derivatives (test_derivatives_omp.f90, alpha = 1, beta = 2)
  allocate(f(nx,ny,nz), df(nx,ny,nz), dex(nx,ny,nz))
  do k = 1, nz
     do j = 1, ny
        do i = 1, nx
           f(i,j,k) = tf(gx(i)) + tf(gy(j)) + tf(gz(k))
        end do
     end do
  end do
  do axis = 0, 2
     do order = 1, 2
        do k = 1, nz
           do j = 1, ny
              do i = 1, nx
                 if (axis == 0) dex(i,j,k) = td(gx(i), order)
                 if (axis == 1) dex(i,j,k) = td(gy(j), order)
                 if (axis == 2) dex(i,j,k) = td(gz(k), order)
              end do
           end do
        end do
        if (axis == 0) call der_coeffs_init(gx, dc, nx, .true., 6)
        if (axis == 1) call der_coeffs_init(gy, dc, ny, .true., 6)
        if (axis == 2) call der_coeffs_init(gz, dc, nz, .true., 6)
        if (axis == 0 .and. order == 1) call derx_array(dc, f, df, nx, ny, nz)
        if (axis == 0 .and. order == 2) call dderx_array(dc, f, df, nx, ny, nz)
        if (axis == 1 .and. order == 1) call dery_array(dc, f, df, nx, ny, nz)
        if (axis == 1 .and. order == 2) call ddery_array(dc, f, df, nx, ny, nz)
        if (axis == 2 .and. order == 1) call derz_array(dc, f, df, nx, ny, nz)
        if (axis == 2 .and. order == 2) call dderz_array(dc, f, df, nx, ny, nz)
        ! error over points off the box edges (solver-type edges are undefined)
        e = sqrt(sum((df(2:nx-1,2:ny-1,2:nz-1) - dex(2:nx-1,2:ny-1,2:nz-1))**2) / &
             sum(dex(2:nx-1,2:ny-1,2:nz-1)**2))
        print '(a,i1,a,i1,a,es24.16)', 'derivative_axis', axis, '_order', order, ' ', e
        call der_coeffs_destroy(dc)
     end do
  end do

contains

  subroutine grid_asin(x, m, a, b)
    integer, intent(in) :: m
    real(c_double), intent(out) :: x(m)
    real(c_double), intent(in) :: a, b
    real(c_double) :: alpha, xi, p
    integer :: q
    alpha = 1.0_c_double - 33.3_c_double / real(m*m, c_double)
    p = (b - a) / real(m - 1, c_double)
    do q = 1, m
       xi = a + real(q - 1, c_double) * p
       x(q) = a + (b - a) * (asin(-alpha*cos(pi*(xi - a)/(b - a))) / asin(alpha) + 1.0_c_double) * 0.5_c_double
    end do
  end subroutine grid_asin

  ! relative error over interior and faces (test_solver_3d.f90:310-361)
  function rel_err(v, w) result(r)
    real(c_double), intent(in) :: v(nx,ny,nz), w(nx,ny,nz)
    real(c_double) :: r, s1, s2
    s1 = sum((v(2:nx-1,2:ny-1,2:nz-1) - w(2:nx-1,2:ny-1,2:nz-1))**2)
    s2 = sum(w(2:nx-1,2:ny-1,2:nz-1)**2)
    s1 = s1 + sum((v(1,2:ny-1,2:nz-1) - w(1,2:ny-1,2:nz-1))**2) + sum((v(nx,2:ny-1,2:nz-1) - w(nx,2:ny-1,2:nz-1))**2)
    s1 = s1 + sum((v(2:nx-1,1,2:nz-1) - w(2:nx-1,1,2:nz-1))**2) + sum((v(2:nx-1,ny,2:nz-1) - w(2:nx-1,ny,2:nz-1))**2)
    s1 = s1 + sum((v(2:nx-1,2:ny-1,1) - w(2:nx-1,2:ny-1,1))**2) + sum((v(2:nx-1,2:ny-1,nz) - w(2:nx-1,2:ny-1,nz))**2)
    r = sqrt(s1) / sqrt(s2)
  end function rel_err

  function tf(x) result(v)
    real(c_double), intent(in) :: x
    real(c_double) :: v
    v = sin(2.0_c_double*pi*x) + cos(4.0_c_double*pi*x)
  end function tf

  function td(x, o) result(v)
    real(c_double), intent(in) :: x
    integer, intent(in) :: o
    real(c_double) :: v
    if (o == 1) then
       v = 2.0_c_double*pi*(cos(2.0_c_double*pi*x) - 2.0_c_double*sin(4.0_c_double*pi*x))
    else
       v = 4.0_c_double*pi*pi*(-sin(2.0_c_double*pi*x) - 4.0_c_double*cos(4.0_c_double*pi*x))
    end if
  end function td

end program test_solver_3d_mf
