!> Fortran host program driving the solver through the ISO_C_BINDING layer,
!! as HICAR would: the reference unit test's set-up
!! (tests/unit/test_wind_iterative.F90:51-209): flat 250 m plane, 20 levels,
!! alpha = 1, u = 8 cos(2 pi (i-ids)/Lx), v = 6 sin(2 pi (j-jds)/Ly), w = 0,
!! rho = 1, calc_divergence with advect_density = .false. before and after one
!! calc_iter_winds. Passes when the result is finite and max|div| falls.
program test_drop_in
    use iso_c_binding
    use windhip_interface, only : windhip_solve_info
    use wind_iterative_hip
    implicit none
    integer, parameter :: nx = 40, ny = 36, nz = 20
    real, parameter :: pi = 3.14159265358979
    type(wind_tile_t) :: tile
    type(wind_solver_options_t) :: opts
    type(windhip_solve_info) :: info
    real, allocatable, target :: dzdx(:,:,:), dzdy(:,:,:), jac(:,:,:), jac_u(:,:,:), jac_v(:,:,:), jac_w(:,:,:)
    real, allocatable, target :: dzdx_u(:,:,:), dzdy_v(:,:,:), adz(:,:,:), terrain(:,:), rho(:,:,:)
    real, allocatable, target :: u(:,:,:), v(:,:,:), w(:,:,:), alpha(:,:,:), div0(:,:,:), div1(:,:,:)
    real, allocatable, target :: div_h(:,:,:), wr(:,:,:)
    real :: dz(nz), lx, ly, d0, d1, db
    integer :: i, j, k

    call auto_dz_icon(nz, 20.0, 8000.0, 0.7, dz)
    allocate(dzdx(nx,nz,ny), dzdy(nx,nz,ny), jac(nx,nz,ny), jac_w(nx,nz,ny), adz(nx,nz,ny), rho(nx,nz,ny))
    allocate(jac_u(nx+1,nz,ny), dzdx_u(nx+1,nz,ny), jac_v(nx,nz,ny+1), dzdy_v(nx,nz,ny+1), terrain(nx,ny))
    allocate(u(nx+1,nz,ny), v(nx,nz,ny+1), w(nx,nz,ny), alpha(nx,nz,ny), div0(nx,nz,ny), div1(nx,nz,ny))
    dzdx = 0; dzdy = 0; dzdx_u = 0; dzdy_v = 0; terrain = 0
    jac = 1; jac_u = 1; jac_v = 1; jac_w = 1; rho = 1; alpha = 1
    do k = 1, nz
        adz(:,k,:) = dz(k)
    end do
    lx = real(nx - 1); ly = real(ny - 1)
    do i = 1, nx + 1
        u(i,:,:) = 8.0 * cos(2.0 * pi * real(i - 1) / lx)
    end do
    do j = 1, ny + 1
        v(:,:,j) = 6.0 * sin(2.0 * pi * real(j - 1) / ly)
    end do
    w = 0

    tile%nx = nx; tile%ny = ny; tile%nz = nz; tile%dx = 250.0
    tile%ims = 1; tile%ime = nx; tile%kms = 1; tile%kme = nz; tile%jms = 1; tile%jme = ny
    tile%dzdx => dzdx; tile%dzdy => dzdy; tile%jacobian => jac; tile%jacobian_u => jac_u
    tile%jacobian_v => jac_v; tile%jacobian_w => jac_w; tile%dzdx_u => dzdx_u; tile%dzdy_v => dzdy_v
    tile%advection_dz => adz; tile%terrain => terrain; tile%density => rho
    tile%u => u; tile%v => v; tile%w => w

    call init_iter_winds(tile, opts)
    call calc_divergence(tile, div0, .false., .false., .true.)
    call calc_iter_winds(tile, alpha, div0, .false.)
    info = last_solve_info()
    call calc_divergence(tile, div1, .false., .false., .true.)
    d0 = maxval(abs(div0)); d1 = maxval(abs(div1))

    ! balance_uvw: w from the horizontal divergence of (u, v); the full
    ! density-weighted divergence then vanishes to FP32 rounding
    allocate(div_h(nx,nz,ny), wr(nx,nz,ny))
    do k = 1, nz
        rho(:,k,:) = 1.2 - 0.01 * real(k)
    end do
    u = u * 1.25
    call balance_uvw(tile, .true.)
    call calc_divergence(tile, div_h, .true., .true., .false.)
    call calc_divergence(tile, div1, .true., .false., .false.)
    db = maxval(abs(div1)) / maxval(abs(div_h))
    ! calc_w_real then calc_idealized_wgrid: the top level of w is 0
    call calc_w_real(tile, wr, w)
    call calc_idealized_wgrid(tile, wr)
    call finalize_iter_winds()
    write(*,'(A,ES10.3,A,ES10.3)') 'balanced: max|div| / max|div_h| ', db, ' top w ', maxval(abs(w(:,nz,:)))
    if (db /= db .or. db > 1.0e-4 .or. maxval(abs(w(:,nz,:))) /= 0.0 .or. any(wr /= wr)) then
        write(*,*) 'FAIL'
        error stop 1
    end if

    write(*,'(A,ES12.4,A,ES12.4,A,I0,A,I0,A,ES10.3)') 'max|div| before ', d0, ' after ', d1, &
        ' status ', info%status, ' iters ', info%iters, ' true_rel_res ', info%true_rel_res
    if (d1 /= d1 .or. .not. (d1 < d0) .or. info%status /= 0) then
        write(*,*) 'FAIL'
        error stop 1
    end if
    write(*,*) 'PASS'

contains

    !> auto_dz case 1 (src/objects/domain_obj.F90:1532-1544)
    subroutine auto_dz_icon(n, lowest, top, stretch, dzo)
        integer, intent(in) :: n
        real, intent(in) :: lowest, top, stretch
        real, intent(out) :: dzo(n)
        real :: x1, a, b, c, vct(0:n)
        integer :: jk
        x1 = (2.0 * stretch - 1.0) * lowest
        b = (top - (x1 / 6.0) * n**3 - (lowest - x1 / 6.0) * n) / (n**2 - 1.0 / 3.0 * n**3 - 2.0 / 3.0 * n)
        a = (x1 - 2.0 * b) / 6.0
        c = lowest - (a + b)
        do jk = 0, n
            vct(jk) = a * jk**3 + b * jk**2 + c * jk
        end do
        dzo = vct(1:n) - vct(0:n-1)
    end subroutine auto_dz_icon

end program test_drop_in
