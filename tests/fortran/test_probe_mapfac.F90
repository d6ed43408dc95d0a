!> The `wind_iterative` adapter's probing API and map factors, driven through
!! domain_t as HICAR drives the reference module (stub modules of
!! stub_hicar.F90), on a terrain nest with non-unit map factors and density:
!!   1. calc_iter_winds on nest 1 (the stability weight alpha enters the
!!      correction), then calibrate_projection_operator's 27-colour loop
!!      (wind.F90:1010-1099) through probe_lambda_pattern,
!!      probe_zero_corrections, probe_apply_corrections, probe_record and
!!      probe_finalize, with the library's calc_divergence standing in for
!!      HICAR's: the off-stencil leakage is exactly 0, u, v, w are restored,
!!      and the recorded operator equals the one probe_operator builds on the
!!      device for nest 2 (same domain) to FP32 rounding of the responses;
!!   2. a solve with the probed operator leaves the winds divergence free to
!!      the tolerance;
!!   3. nest 3 (the same terrain, map factors, alpha and winds): calc_divergence
!!      and calc_iter_winds with the analytic operator, and every input and
!!      output written to the file named by the first argument, which
!!      tests/test_fortran_dropin.py checks against the CPU oracle.
program test_probe_mapfac
    use iso_c_binding
    use icar_constants, only : kVARS
    use domain_interface, only : domain_t
    use options_interface, only : options_t
    use windhip_interface, only : windhip_solve_info, windhip_apply_operator, WINDHIP_PLAIN_F64, WINDHIP_START_WARM
    use wind_iterative_hip, only : wind_tile_t, calc_divergence, nest_handle
    ! the use-list of HICAR's wind.F90:14-16, which must compile against this module
    use wind_iterative,    only : calc_iter_winds, init_iter_winds, &
                                  probe_lambda_pattern, probe_zero_corrections, &
                                  probe_apply_corrections, probe_record, probe_finalize
    use wind_iterative,    only : finalize_iter_winds, probe_operator, set_wind_solver_choice, wind_tile_of, &
                                  last_solve_info
    implicit none
    integer, parameter :: nx = 30, ny = 26, nz = 10
    real, parameter :: dx = 200.0
    type(domain_t), target :: d1, d2, d3
    type(options_t) :: opts
    type(windhip_solve_info) :: info
    real, allocatable, target :: alpha(:,:,:), div(:,:,:), us(:,:,:), vs(:,:,:), ws(:,:,:)
    real(c_double), allocatable, target :: x(:), y1(:), y2(:)
    real :: max_leak, d0, da, rel
    real(c_double) :: leak2
    integer :: ca, cb, cc, nrows, q, un
    character(len=512) :: dump
    logical :: ok
    type(wind_tile_t) :: tile

    ok = .true.
    call get_command_argument(1, dump)
    opts%wind%wind_solver_iterations = 5000
    call set_wind_solver_choice(WINDHIP_PLAIN_F64, WINDHIP_START_WARM, tol_rel=1.0e-9_c_double)
    call make_domain(d1, 1)
    call make_domain(d2, 2)
    call make_domain(d3, 3)
    allocate(alpha(nx, nz, ny), div(nx, nz, ny))
    call fill_alpha(alpha)

    ! 1. nest 1: a solve sets alpha, then the reference's calibration loop
    call init_iter_winds(d1, opts)
    call wind_tile_of(d1, tile)
    call calc_divergence(tile, div, .true., .false., .true.)
    call calc_iter_winds(d1, alpha, div, .true.)
    associate(u_dqdt => d1%vars_3d(d1%var_indx(kVARS%u)%v)%dqdt_3d, &
              v_dqdt => d1%vars_3d(d1%var_indx(kVARS%v)%v)%dqdt_3d, &
              w_dqdt => d1%vars_3d(d1%var_indx(kVARS%w)%v)%dqdt_3d)
        us = u_dqdt; vs = v_dqdt; ws = w_dqdt
        max_leak = 0.0
        do cc = 0, 2
            do cb = 0, 2
                do ca = 0, 2
                    call probe_lambda_pattern(ca, cb, cc)
                    call probe_zero_corrections(d1)
                    call probe_apply_corrections(d1, .true.)
                    call calc_divergence(tile, div, .true., .false., .true.)
                    call probe_record(d1, div, ca, cb, cc, max_leak)
                enddo
            enddo
        enddo
        call probe_finalize(max_leak)
        u_dqdt = us; v_dqdt = vs; w_dqdt = ws
    end associate
    write(*,'(A,ES12.4)') 'five-call probing: max leak ', max_leak
    if (max_leak /= 0.0) call fail("leakage of the five-call probing")

    ! nest 2: the same domain, probed in one call on the device
    call init_iter_winds(d2, opts)
    call probe_operator(d2, alpha, .true., leak2)
    if (leak2 /= 0.0_c_double) call fail("leakage of probe_operator")

    ! the two operators on the same random rows
    nrows = (nx + 2) * (nz + 2) * (ny + 2)
    allocate(x(nrows), y1(nrows), y2(nrows))
    do q = 1, nrows
        x(q) = sin(0.37d0 * q) + cos(0.11d0 * q * q)
    end do
    if (windhip_apply_operator(nest_handle(1), WINDHIP_PLAIN_F64, c_loc(x), c_loc(y1)) /= 0) call fail("apply 1")
    if (windhip_apply_operator(nest_handle(2), WINDHIP_PLAIN_F64, c_loc(x), c_loc(y2)) /= 0) call fail("apply 2")
    rel = real(maxval(abs(y1 - y2)) / maxval(abs(y2)))
    write(*,'(A,ES12.4)') 'five-call vs one-call probed operator: max |dy| / max |y| = ', rel
    if (.not. (rel <= 1.0e-5)) call fail("probed operators differ")

    ! 2. a solve with the five-call probed operator
    call init_iter_winds(d1, opts)
    call calc_divergence(tile, div, .true., .false., .true.)
    d0 = maxval(abs(div))
    call calc_iter_winds(d1, alpha, div, .true.)
    call calc_divergence(tile, div, .true., .false., .true.)
    da = maxval(abs(div))
    info = last_solve_info()
    write(*,'(A,ES12.4,A,ES12.4,A,I0)') 'probed solve: max|div| ', d0, ' -> ', da, ' iterations ', info%iters
    if (info%status /= 0 .or. .not. (da < 1.0e-3 * d0)) call fail("probed solve")

    ! 3. nest 3: analytic operator with map factors, dumped for the oracle
    call init_iter_winds(d3, opts)
    call wind_tile_of(d3, tile)
    call calc_divergence(tile, div, .true., .false., .true.)
    open(newunit=un, file=trim(dump), access='stream', form='unformatted', status='replace')
    write(un) nx, ny, nz, dx
    associate(v => d3%vars_3d, xi => d3%var_indx)
        write(un) v(xi(kVARS%dzdx)%v)%data_3d, v(xi(kVARS%dzdy)%v)%data_3d, v(xi(kVARS%jacobian)%v)%data_3d, &
                  v(xi(kVARS%jacobian_u)%v)%data_3d, v(xi(kVARS%jacobian_v)%v)%data_3d, &
                  v(xi(kVARS%jacobian_w)%v)%data_3d, v(xi(kVARS%dzdx_u)%v)%data_3d, v(xi(kVARS%dzdy_v)%v)%data_3d, &
                  v(xi(kVARS%advection_dz)%v)%data_3d, d3%vars_2d(1)%data_2d
        write(un) d3%mapfac_mx_u, d3%mapfac_my_u, d3%mapfac_my_v, d3%mapfac_mx_v, d3%mapfac_mxy
        write(un) alpha, v(xi(kVARS%density)%v)%data_3d, v(xi(kVARS%u)%v)%dqdt_3d, v(xi(kVARS%v)%v)%dqdt_3d, &
                  v(xi(kVARS%w)%v)%dqdt_3d, div
        call calc_iter_winds(d3, alpha, div, .true.)
        info = last_solve_info()
        write(un) info%status, info%iters
        write(un) v(xi(kVARS%u)%v)%dqdt_3d, v(xi(kVARS%v)%v)%dqdt_3d, v(xi(kVARS%w)%v)%dqdt_3d
    end associate
    close(un)
    write(*,'(A,I0,A)') 'map-factor solve: ', info%iters, ' iterations, written for the oracle'

    call finalize_iter_winds()
    if (.not. ok) error stop 1
    write(*,*) 'PASS'

contains

    subroutine fail(what)
        character(len=*), intent(in) :: what
        write(*,*) 'FAIL: ', what
        ok = .false.
    end subroutine fail

    subroutine fill_alpha(a)
        real, intent(out) :: a(:,:,:)
        integer :: i, j, k
        do j = 1, ny
            do k = 1, nz
                do i = 1, nx
                    a(i, k, j) = 0.65 + 0.3 * sin(0.21 * i + 0.13 * j) * cos(0.4 * k)
                end do
            end do
        end do
    end subroutine fill_alpha

    !> A terrain nest: smooth analytic metrics, stretched levels, map factors
    !! that vary in x and y, density that varies in all three directions, and
    !! winds with a sheared profile (the same in every nest).
    subroutine make_domain(d, nest)
        type(domain_t), intent(out) :: d
        integer, intent(in) :: nest
        integer :: q, i, j, k
        real :: xx, yy
        integer, parameter :: order(13) = [13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1]
        d%nest_indx = nest
        d%dx = dx
        d%grid%ids = 1; d%grid%ide = nx; d%grid%jds = 1; d%grid%jde = ny; d%grid%kds = 1; d%grid%kde = nz
        d%ids = 1; d%ide = nx; d%jds = 1; d%jde = ny; d%kds = 1; d%kde = nz
        d%ims = 1; d%ime = nx; d%jms = 1; d%jme = ny; d%kms = 1; d%kme = nz
        allocate(d%vars_3d(13), d%vars_2d(1))
        do q = 1, 13
            d%var_indx(q)%v = order(q)
        end do
        d%var_indx(kVARS%neighbor_terrain)%v = 1
        allocate(d%vars_2d(1)%data_2d(nx, ny))
        allocate(d%mapfac_mx_u(nx + 1, ny), d%mapfac_my_u(nx + 1, ny), d%mapfac_my_v(nx, ny + 1), &
                 d%mapfac_mx_v(nx, ny + 1), d%mapfac_mxy(nx, ny))
        associate(v => d%vars_3d, x => d%var_indx)
            allocate(v(x(kVARS%u)%v)%dqdt_3d(nx + 1, nz, ny), v(x(kVARS%v)%v)%dqdt_3d(nx, nz, ny + 1))
            allocate(v(x(kVARS%w)%v)%dqdt_3d(nx, nz, ny))
            allocate(v(x(kVARS%jacobian_u)%v)%data_3d(nx + 1, nz, ny), v(x(kVARS%dzdx_u)%v)%data_3d(nx + 1, nz, ny))
            allocate(v(x(kVARS%jacobian_v)%v)%data_3d(nx, nz, ny + 1), v(x(kVARS%dzdy_v)%v)%data_3d(nx, nz, ny + 1))
            allocate(v(x(kVARS%dzdx)%v)%data_3d(nx, nz, ny), v(x(kVARS%dzdy)%v)%data_3d(nx, nz, ny))
            allocate(v(x(kVARS%jacobian)%v)%data_3d(nx, nz, ny), v(x(kVARS%jacobian_w)%v)%data_3d(nx, nz, ny))
            allocate(v(x(kVARS%advection_dz)%v)%data_3d(nx, nz, ny), v(x(kVARS%density)%v)%data_3d(nx, nz, ny))
            do j = 1, ny + 1
                do k = 1, nz
                    do i = 1, nx + 1
                        xx = real(i) / nx; yy = real(j) / ny
                        if (j <= ny) then
                            v(x(kVARS%dzdx_u)%v)%data_3d(i, k, j) = 0.12 * cos(6.0 * xx) * sin(4.0 * yy) / k
                            v(x(kVARS%jacobian_u)%v)%data_3d(i, k, j) = 1.0 + 0.08 * sin(5.0 * xx + yy) / k
                            v(x(kVARS%u)%v)%dqdt_3d(i, k, j) = 6.0 + 0.3 * k + sin(3.0 * xx + 2.0 * yy)
                        end if
                        if (i <= nx) then
                            v(x(kVARS%dzdy_v)%v)%data_3d(i, k, j) = 0.1 * sin(5.0 * xx) * cos(3.0 * yy) / k
                            v(x(kVARS%jacobian_v)%v)%data_3d(i, k, j) = 1.0 + 0.07 * cos(4.0 * yy - xx) / k
                            v(x(kVARS%v)%v)%dqdt_3d(i, k, j) = 2.0 + 0.1 * k + cos(2.0 * xx - 3.0 * yy)
                        end if
                        if (i <= nx .and. j <= ny) then
                            v(x(kVARS%dzdx)%v)%data_3d(i, k, j) = 0.12 * cos(6.0 * xx) * sin(4.0 * yy) / k + 0.01
                            v(x(kVARS%dzdy)%v)%data_3d(i, k, j) = 0.1 * sin(5.0 * xx) * cos(3.0 * yy) / k - 0.01
                            v(x(kVARS%jacobian)%v)%data_3d(i, k, j) = 1.0 + 0.08 * sin(5.0 * xx + yy) / k
                            v(x(kVARS%jacobian_w)%v)%data_3d(i, k, j) = 1.0 + 0.06 * cos(3.0 * xx - yy) / k
                            v(x(kVARS%advection_dz)%v)%data_3d(i, k, j) = 15.0 * 1.3**(k - 1) * &
                                (1.0 + 0.05 * sin(xx + yy))
                            v(x(kVARS%density)%v)%data_3d(i, k, j) = 1.2 - 0.03 * k + 0.02 * sin(7.0 * xx * yy)
                            v(x(kVARS%w)%v)%dqdt_3d(i, k, j) = 0.0
                        end if
                    end do
                end do
            end do
        end associate
        do j = 1, ny
            do i = 1, nx
                d%vars_2d(1)%data_2d(i, j) = 300.0 + 40.0 * sin(3.0 * i / real(nx)) * cos(2.0 * j / real(ny))
                d%mapfac_mxy(i, j) = (1.0 + 0.04 * sin(2.0 * i / real(nx))) * (1.0 + 0.03 * cos(3.0 * j / real(ny)))
            end do
        end do
        do j = 1, ny
            do i = 1, nx + 1
                d%mapfac_mx_u(i, j) = 1.0 + 0.04 * sin(2.0 * i / real(nx)) + 0.01 * j / real(ny)
                d%mapfac_my_u(i, j) = 1.0 + 0.03 * cos(3.0 * j / real(ny)) - 0.01 * i / real(nx)
            end do
        end do
        do j = 1, ny + 1
            do i = 1, nx
                d%mapfac_my_v(i, j) = 1.0 + 0.03 * cos(3.0 * j / real(ny)) + 0.02 * i / real(nx)
                d%mapfac_mx_v(i, j) = 1.0 + 0.04 * sin(2.0 * i / real(nx)) - 0.01 * j / real(ny)
            end do
        end do
    end subroutine make_domain

end program test_probe_mapfac
