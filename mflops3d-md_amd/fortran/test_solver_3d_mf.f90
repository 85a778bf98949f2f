! Fortran caller of libmflops through the ISO_C_BINDING modules, following
! the reference's own drivers:
!   src/main/test_solver_3d.f90        u = x^9 + y^9 + xy, sigma = 2, all
!                                      Dirichlet, relative error err() :310-361
!   src/main/test_derivatives_omp.f90  sin(2 pi a x) + cos(2 pi b x) per axis,
!                                      a = 1, b = 2, solver-type coefficients
! Grids use the reference's asin/cos stretching (class_mesh.f90:181-200) on
! x,y in [-0.5,0.5], z in [-0.25,0.25].  Then a field file round trip
! (write_var3d / get_var3d_info / read_var3d) and a laminar channel through
! the NS step (the parabola with the flow-rate force is a discrete steady
! state), written to a restart directory and continued from it.  Prints one
! "name value" line per result for the calling test to check.
program test_solver_3d_mf
  use iso_c_binding
  use mflops_solver_3d
  use mflops_derivatives
  use mflops_io
  use mflops_navier
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
        if (axis == 0) then
           ! the reference's line loop (class_field.f90:649-684): der on every
           ! interior x-line, der_s on the lines lying in one boundary plane
           block
             real(c_double), allocatable :: dl(:,:,:)
             integer(8) :: c0, c1, rate
             integer :: nl
             allocate(dl(nx,ny,nz))
             dl = 0.0_c_double
             nl = 0
             call system_clock(c0, rate)
             do k = 1, nz
                do j = 1, ny
                   if ((j == 1 .or. j == ny) .and. (k == 1 .or. k == nz)) cycle
                   if (j == 1 .or. j == ny .or. k == 1 .or. k == nz) then
                      if (order == 1) call der_s(dc, f(:,j,k), dl(:,j,k))
                      if (order == 2) call dder_s(dc, f(:,j,k), dl(:,j,k))
                   else
                      if (order == 1) call der(dc, f(:,j,k), dl(:,j,k))
                      if (order == 2) call dder(dc, f(:,j,k), dl(:,j,k))
                   end if
                   nl = nl + 1
                end do
             end do
             call system_clock(c1)
             e = 0.0_c_double
             do k = 1, nz
                do j = 1, ny
                   if ((j == 1 .or. j == ny) .and. (k == 1 .or. k == nz)) cycle
                   if (j == 1 .or. j == ny .or. k == 1 .or. k == nz) then
                      e = max(e, maxval(abs(dl(2:nx-1,j,k) - df(2:nx-1,j,k))))
                   else
                      e = max(e, maxval(abs(dl(:,j,k) - df(:,j,k))))
                   end if
                end do
             end do
             print '(a,i1,a,es24.16)', 'der_lines_order', order, '_vs_3d ', e / maxval(abs(df))
             print '(a,i1,a,es24.16)', 'der_lines_order', order, '_us_per_line ', &
                  1.0d6 * real(c1 - c0, 8) / real(rate, 8) / real(nl, 8)
           end block
        end if
        call der_coeffs_destroy(dc)
     end do
  end do


  ! ---- field file round trip (class_io.f90)
  block
    character(len=512) :: dn(3)
    integer :: dl(3)
    real(c_double), allocatable :: back(:,:,:)
    call write_var3d('mf_io_test.nc', [character(len=512) :: 'resolution_x', 'resolution_y', &
         'resolution_z'], [nx, ny, nz], 'U', u, dbl='y')
    call get_var3d_info('mf_io_test.nc', 'U', dn, dl)
    allocate(back(dl(1), dl(2), dl(3)))
    call read_var3d('mf_io_test.nc', 'U', back)
    print '(a,1x,es24.16)', 'io_roundtrip_maxdiff', maxval(abs(back - u))
    if (dl(1) == nx .and. dl(2) == ny .and. dl(3) == nz .and. trim(dn(2)) == 'resolution_y') then
       print '(a,1x,i0)', 'io_dims_ok', 1
    else
       print '(a,1x,i0)', 'io_dims_ok', 0
    end if
  end block

  ! ---- laminar channel through the NS step, restart and continue
  block
    type(navier3d) :: nav, nav2
    integer, parameter :: cx = 16, cy = 24, cz = 12
    real(c_double) :: cgx(cx), cgy(cy), cgz(cz), re, dt
    real(c_double), allocatable :: cu(:,:,:), cv(:,:,:), par(:,:,:), a1(:,:,:), a2(:,:,:)
    real(c_double), allocatable :: bx(:,:,:), by(:,:,:), bz(:,:,:), zx(:,:,:), zy(:,:,:), zz(:,:,:)
    re = 3000.0_c_double
    dt = 1.0e-3_c_double
    do i = 1, cx
       cgx(i) = 2.0_c_double * pi * (i - 1) / (cx - 1)
    end do
    do j = 1, cy
       cgy(j) = -1.0_c_double + (1.0_c_double + tanh(2.5_c_double * (2.0_c_double * (j - 1) / (cy - 1) - &
            1.0_c_double)) / tanh(2.5_c_double))
    end do
    do k = 1, cz
       cgz(k) = pi * (k - 1) / (cz - 1)
    end do
    allocate(cu(cx,cy,cz), cv(cx,cy,cz), par(cx,cy,cz), a1(cx,cy,cz), a2(cx,cy,cz))
    do k = 1, cz
       do j = 1, cy
          par(:, j, k) = 1.0_c_double - cgy(j)**2
       end do
    end do
    cu = par
    cv = 0.0_c_double
    allocate(bx(cy-2,cz-2,2), by(cx-2,cz-2,2), bz(cx-2,cy-2,2))
    allocate(zx(cy-2,cz-2,2), zy(cx-2,cz-2,2), zz(cx-2,cy-2,2))
    bx(:,:,1) = cu(1,2:cy-1,2:cz-1) ; bx(:,:,2) = cu(cx,2:cy-1,2:cz-1)
    by(:,:,1) = cu(2:cx-1,1,2:cz-1) ; by(:,:,2) = cu(2:cx-1,cy,2:cz-1)
    bz(:,:,1) = cu(2:cx-1,2:cy-1,1) ; bz(:,:,2) = cu(2:cx-1,2:cy-1,cz)
    zx = 0.0_c_double ; zy = 0.0_c_double ; zz = 0.0_c_double
    call navier_initialization(nav, cgx, cgy, cgz, re, dt)
    call navier_set_bc(nav, 1, bx, by, bz)
    call navier_set_bc(nav, 2, zx, zy, zz)
    call navier_set_bc(nav, 3, zx, zy, zz)
    call navier_set_flow_rate(nav, .true., 2.0_c_double / re)
    call navier_set_state(nav, cu, cv, cv)
    call navier_step(nav, 5)
    call navier_get(nav, u=a1)
    print '(a,1x,es24.16)', 'ns_channel_steady', &
         maxval(abs(a1(2:cx-1,2:cy-1,2:cz-1) - par(2:cx-1,2:cy-1,2:cz-1)))
    print '(a,1x,es24.16)', 'ns_time', navier_time(nav)
    call restart_write(nav, 'restart_f')
    call navier_initialization(nav2, cgx, cgy, cgz, re, dt)
    call navier_set_bc(nav2, 1, bx, by, bz)
    call navier_set_bc(nav2, 2, zx, zy, zz)
    call navier_set_bc(nav2, 3, zx, zy, zz)
    call navier_set_flow_rate(nav2, .true., 2.0_c_double / re)
    call restart_read(nav2, 'restart_f')
    ! the restarted run re-initialises its flow-rate force, as the reference;
    ! so does the continued one here, to compare like with like
    call navier_set_flow_rate(nav, .true., 2.0_c_double / re)
    call navier_step(nav, 3)
    call navier_step(nav2, 3)
    call navier_get(nav, u=a1)
    call navier_get(nav2, u=a2)
    print '(a,1x,es24.16)', 'ns_restart_maxdiff', maxval(abs(a1 - a2))
    call navier_finalize(nav)
    call navier_finalize(nav2)
  end block

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
