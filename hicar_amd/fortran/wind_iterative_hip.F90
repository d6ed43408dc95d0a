!>----------------------------------------------------------------------------
!! Drop-in for HICAR's `wind_iterative` module (src/physics/wind_iterative.F90)
!! on one MI355X, through the C-ABI of libwindhip.so.
!!
!! Same public operations and argument meaning as the reference module
!! (:39-41): init_iter_winds, calc_iter_winds, finalize_iter_winds, and the
!! probing entry point used by calibrate_projection_operator. The reference
!! takes `type(domain_t)`; this module takes `wind_tile_t`, which holds exactly
!! the domain fields the solve reads (pointers into the caller's arrays, no
!! copies). A HICAR adapter fills it from domain_t (INTEGRATION.md shows how).
!!
!! Fields are HICAR's `real` (FP32), host memory, Fortran (i,k,j) with the
!! tile's memory bounds; the library promotes them to FP64 on the GPU.
!! Not thread-safe (module state per nest, as in the reference, :134-181).
!!----------------------------------------------------------------------------
module wind_iterative_hip
    use iso_c_binding
    use windhip_interface
    implicit none
    private

    public :: wind_tile_t, wind_solver_options_t
    public :: init_iter_winds, calc_iter_winds, finalize_iter_winds, probe_operator, last_solve_info
    public :: probe_lambda_pattern, probe_zero_corrections, probe_apply_corrections, probe_record, probe_finalize
    public :: nest_handle
    public :: calc_divergence, load_learned_guess, balance_uvw, calc_w_real, calc_idealized_wgrid

    integer, parameter :: MAX_NESTS = 4   ! wind_iterative.F90:139

    !> The parts of domain_t the wind solve reads (one tile, one GPU).
    type :: wind_tile_t
        integer :: nest_indx = 1
        integer :: nx, ny, nz
        real    :: dx
        integer :: ims, ime, kms, kme, jms, jme
        ! metrics (ims:ime[+1], kms:kme, jms:jme[+1]) and terrain (ims:ime, jms:jme)
        real, pointer, contiguous :: dzdx(:,:,:) => null(), dzdy(:,:,:) => null()
        real, pointer, contiguous :: jacobian(:,:,:) => null(), jacobian_u(:,:,:) => null()
        real, pointer, contiguous :: jacobian_v(:,:,:) => null(), jacobian_w(:,:,:) => null()
        real, pointer, contiguous :: dzdx_u(:,:,:) => null(), dzdy_v(:,:,:) => null()
        real, pointer, contiguous :: advection_dz(:,:,:) => null(), terrain(:,:) => null()
        real, pointer, contiguous :: density(:,:,:) => null()
        ! the forcing-tendency winds the solve corrects in place (u%dqdt_3d, ...)
        real, pointer, contiguous :: u(:,:,:) => null(), v(:,:,:) => null(), w(:,:,:) => null()
        ! map factors (domain_h.F90:91-93; unassociated = 1): m_x and m_y on the
        ! u grid (ims:ime+1, jms:jme), on the v grid (ims:ime, jms:jme+1), and
        ! m_x m_y on the mass grid (ims:ime, jms:jme)
        real, pointer, contiguous :: mapfac_mx_u(:,:) => null(), mapfac_my_u(:,:) => null()
        real, pointer, contiguous :: mapfac_my_v(:,:) => null(), mapfac_mx_v(:,:) => null()
        real, pointer, contiguous :: mapfac_mxy(:,:) => null()
    end type wind_tile_t

    !> options%wind%wind_solver_iterations and the solver choices of this build.
    type :: wind_solver_options_t
        integer :: wind_solver_iterations = 1500     ! namelist_utilities.F90:3871-3876
        real(c_double) :: tol_rel = 1.0e-5_c_double   ! wind_iterative.F90:60
        real(c_double) :: tol_abs = 1.0e-10_c_double  ! wind_iterative.F90:59
        integer :: precision = WINDHIP_PLAIN_F64      ! or WINDHIP_MIXED, WINDHIP_MIXED_LINE, WINDHIP_MIXED_MG
        integer :: start = WINDHIP_START_WARM         ! WINDHIP_START_LEARNED once load_learned_guess ran
        logical :: verbose = .false.
        ! WINDHIP_HOST: the tile's arrays are host memory (staged by the library);
        ! WINDHIP_DEVICE: their addresses (c_loc of the tile's pointers) are
        ! device addresses, e.g. pointers set up from device arrays
        integer :: field_mem = WINDHIP_HOST
    end type wind_solver_options_t

    type :: nest_slot_t
        type(c_ptr) :: handle = c_null_ptr
        type(c_ptr) :: net = c_null_ptr
        type(windhip_solve_opts) :: opts
        logical :: verbose = .false.
    end type nest_slot_t

    type(nest_slot_t), save :: slots(MAX_NESTS)
    type(windhip_solve_info), save :: info_last

contains

    !> init_iter_winds(domain, options) (wind_iterative.F90:215-291): creates the
    !! nest's device state on first use and uploads the terrain metrics
    !! (init_module_vars, :2274-2421); later calls for a known nest are O(1).
    subroutine init_iter_winds(tile, options)
        type(wind_tile_t), intent(in) :: tile
        type(wind_solver_options_t), intent(in) :: options
        type(windhip_desc) :: d
        integer(c_int) :: rc
        integer :: n

        n = tile%nest_indx
        if (n < 1 .or. n > MAX_NESTS) error stop "wind_iterative_hip: nest index out of range"
        if (c_associated(slots(n)%handle)) return

        d%nx = tile%nx; d%ny = tile%ny; d%nz = tile%nz
        d%dx = real(tile%dx, c_double)
        d%ims = tile%ims; d%ime = tile%ime; d%kms = tile%kms; d%kme = tile%kme
        d%jms = tile%jms; d%jme = tile%jme
        d%field_dtype = WINDHIP_F32
        d%field_mem = options%field_mem
        rc = windhip_create(d, slots(n)%handle)
        call check(rc, "windhip_create")
        rc = windhip_set_metrics(slots(n)%handle, c_loc(tile%dzdx), c_loc(tile%dzdy), c_loc(tile%jacobian), &
                                 c_loc(tile%jacobian_u), c_loc(tile%jacobian_v), c_loc(tile%jacobian_w), &
                                 c_loc(tile%dzdx_u), c_loc(tile%dzdy_v), c_loc(tile%advection_dz), &
                                 c_loc(tile%terrain), addr2(tile%mapfac_mx_u), addr2(tile%mapfac_my_u), &
                                 addr2(tile%mapfac_my_v), addr2(tile%mapfac_mx_v), addr2(tile%mapfac_mxy))
        call check(rc, "windhip_set_metrics")
        call windhip_default_opts(slots(n)%opts)
        slots(n)%opts%max_iters = options%wind_solver_iterations
        slots(n)%opts%tol_rel = options%tol_rel
        slots(n)%opts%tol_abs = options%tol_abs
        slots(n)%opts%precision = options%precision
        slots(n)%opts%start = options%start
        slots(n)%verbose = options%verbose
        ! the verbose report's per-region times come from the per-kernel device times
        slots(n)%opts%profile = merge(1_c_int, 0_c_int, options%verbose)
    end subroutine init_iter_winds

    !> calc_iter_winds(domain, alpha_in, div_in, adv_den) (wind_iterative.F90:295-532):
    !! solve for lambda and correct tile%u, tile%v, tile%w in place.
    subroutine calc_iter_winds(tile, alpha_in, div_in, adv_den)
        type(wind_tile_t), intent(inout) :: tile
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(in), target :: &
            alpha_in, div_in
        logical, intent(in) :: adv_den
        integer(c_int) :: rc, ad
        type(c_ptr) :: rho, h

        h = slots(tile%nest_indx)%handle
        if (.not. c_associated(h)) error stop "calc_iter_winds: init_iter_winds has not been called"
        ad = merge(1_c_int, 0_c_int, adv_den)
        rho = c_null_ptr
        if (adv_den .and. associated(tile%density)) rho = c_loc(tile%density)
        rc = windhip_calc_iter_winds(h, c_loc(alpha_in), c_loc(div_in), c_loc(tile%u), c_loc(tile%v), &
                                     c_loc(tile%w), rho, ad, slots(tile%nest_indx)%opts, info_last)
        if (slots(tile%nest_indx)%verbose) then
            write(*,*) ' windhip BiCGStab status=', info_last%status, ' iterations=', info_last%iters, &
                       ' precond_n_sweeps=', info_last%nsweeps
            write(*,*) '   Residual at iter 0:    ', info_last%res0
            write(*,*) '   Residual at final iter:', info_last%res_final
            call report_timing(h, info_last)
        end if
        if (rc == WINDHIP_ERR_DIVERGED) then
            ! the reference stops the model here (wind_iterative.F90:506-528)
            write(*,*) '---------------windhip BiCGStab DIVERGED: ', windhip_error_message()
            error stop
        end if
        call check(rc, "windhip_calc_iter_winds")
    end subroutine calc_iter_winds

    !> The per-solve timing breakdown of the reference's verbose report
    !! (wind_iterative.F90:447-471), from the device times of the solve's kernels
    !! (windhip_profile, read and reset): SpMV = every kernel that applies the
    !! operator (the last Jacobi-Richardson sweep is fused into the two matvec
    !! kernels of an iteration and counts here); preconditioner = the lone sweeps
    !! of the retry ladder, line solves, inner residuals and the multigrid
    !! levels; vector ops = the updates, the FGMRES set-up, Gram-Schmidt steps and
    !! lambda accumulation. One device, one process: no halo exchange, and the
    !! dot products are reduced inside their kernels (no allreduce).
    subroutine report_timing(h, info)
        type(c_ptr), intent(in) :: h
        type(windhip_solve_info), intent(in) :: info
        real(c_double) :: calls(WINDHIP_PROFILE_KINDS), ms(WINDHIP_PROFILE_KINDS)
        real(c_double) :: total, spmv, prec, vec, other
        integer :: k
        integer, parameter :: spmv_kinds(*) = [1, 3, 6, 8, 11, 13, 15, 19, 20]
        integer, parameter :: prec_kinds(*) = [0, 5, 17, 18, 21, 22, 23, 24, 25]
        integer, parameter :: vec_kinds(*) = [2, 4, 7, 9, 12, 14, 16]
        if (windhip_profile(h, 1_c_int, calls, ms) /= WINDHIP_OK) return
        total = info%ms * 1.0e-3_c_double
        spmv = 0; prec = 0; vec = 0
        do k = 1, size(spmv_kinds)
            spmv = spmv + ms(spmv_kinds(k) + 1) * 1.0e-3_c_double
        end do
        do k = 1, size(prec_kinds)
            prec = prec + ms(prec_kinds(k) + 1) * 1.0e-3_c_double
        end do
        do k = 1, size(vec_kinds)
            vec = vec + ms(vec_kinds(k) + 1) * 1.0e-3_c_double
        end do
        if (total <= 0) return
        other = total - (spmv + prec + vec)
        write(*,'(A)')                '   --- per-solve timing breakdown (device, seconds) ---'
        write(*,'(A,F10.4,A)')        '     total           : ', total, ''
        write(*,'(A,F10.4,A,F6.1,A)') '     SpMV            : ', spmv, '   (', 100 * spmv / total, ' %)'
        write(*,'(A,F10.4,A,F6.1,A)') '     halo exchange   : ', 0.0_c_double, '   (', 0.0_c_double, ' %)'
        write(*,'(A,F10.4,A,F6.1,A)') '     preconditioner  : ', prec, '   (', 100 * prec / total, ' %)'
        write(*,'(A,F10.4,A,F6.1,A)') '     allreduce       : ', 0.0_c_double, '   (', 0.0_c_double, ' %)'
        write(*,'(A,F10.4,A,F6.1,A)') '     vector ops      : ', vec, '   (', 100 * vec / total, ' %)'
        write(*,'(A,F10.4,A,F6.1,A)') '     other           : ', other, '   (', 100 * other / total, ' %)'
        if (info%iters > 0) write(*,'(A,F10.6,A)') '     per-iter total  : ', total / real(info%iters, c_double), ' s'
    end subroutine report_timing

    !> calibrate_projection_operator (wind.F90:1010-1099) on the device: switch the
    !! nest to the probed operator A = 2 D(G(.)).
    subroutine probe_operator(tile, alpha_in, adv_den, max_leak)
        type(wind_tile_t), intent(inout) :: tile
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(in), target :: alpha_in
        logical, intent(in) :: adv_den
        real(c_double), intent(out) :: max_leak
        integer(c_int) :: rc
        type(c_ptr) :: rho
        rho = c_null_ptr
        if (adv_den .and. associated(tile%density)) rho = c_loc(tile%density)
        rc = windhip_probe_operator(slots(tile%nest_indx)%handle, c_loc(alpha_in), rho, &
                                    merge(1_c_int, 0_c_int, adv_den), max_leak)
        call check(rc, "windhip_probe_operator")
    end subroutine probe_operator

    !> probe_lambda_pattern(ca, cb, cc) (wind_iterative.F90:1096-1116) on nest
    !! nest_indx: lambda = 1 on the physical cells of colour (ca, cb, cc).
    subroutine probe_lambda_pattern(nest_indx, ca, cb, cc)
        integer, intent(in) :: nest_indx, ca, cb, cc
        call check(windhip_probe_lambda_pattern(handle_of(nest_indx), int(ca, c_int), int(cb, c_int), &
                                                int(cc, c_int)), "windhip_probe_lambda_pattern")
    end subroutine probe_lambda_pattern

    !> probe_zero_corrections(domain) (:1118-1153): the tile's u, v, w over
    !! their memory bounds set to zero, so the correction leaves G(lambda) there.
    subroutine probe_zero_corrections(tile)
        type(wind_tile_t), intent(inout) :: tile
        tile%u = 0.0
        tile%v = 0.0
        tile%w = 0.0
    end subroutine probe_zero_corrections

    !> probe_apply_corrections(domain, adv_den) (:1155-1161): the correction
    !! operator G (calc_updated_winds) from the current lambda.
    subroutine probe_apply_corrections(tile, adv_den)
        type(wind_tile_t), intent(inout) :: tile
        logical, intent(in) :: adv_den
        type(c_ptr) :: rho
        rho = c_null_ptr
        if (adv_den .and. associated(tile%density)) rho = c_loc(tile%density)
        call check(windhip_apply_correction(handle_of(tile%nest_indx), c_loc(tile%u), c_loc(tile%v), &
                                            c_loc(tile%w), rho, merge(1_c_int, 0_c_int, adv_den)), &
                   "windhip_apply_correction")
    end subroutine probe_apply_corrections

    !> probe_record(domain, Tfield, ca, cb, cc, max_leak) (:1163-1230): 2 T into
    !! the coefficients of colour (ca, cb, cc); max_leak raised to the largest
    !! off-stencil |2 T|.
    subroutine probe_record(tile, Tfield, ca, cb, cc, max_leak)
        type(wind_tile_t), intent(in) :: tile
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(in), target :: Tfield
        integer, intent(in) :: ca, cb, cc
        real, intent(inout) :: max_leak
        real(c_double) :: lk
        lk = real(max_leak, c_double)
        call check(windhip_probe_record(handle_of(tile%nest_indx), c_loc(Tfield), int(ca, c_int), int(cb, c_int), &
                                        int(cc, c_int), lk), "windhip_probe_record")
        max_leak = max(max_leak, real(lk))
    end subroutine probe_record

    !> probe_finalize(max_leak) (:1232-1248): nest nest_indx solves with the
    !! recorded operator from now on; lambda is reset.
    subroutine probe_finalize(nest_indx, max_leak)
        integer, intent(in) :: nest_indx
        real, intent(in) :: max_leak
        call check(windhip_probe_finalize(handle_of(nest_indx)), "windhip_probe_finalize")
        write(*,*) "Poisson projection operator calibrated by probing (A = 2*D o G) on the GPU."
        write(*,*) "  off-stencil leakage max |T| = ", max_leak, " (should be ~0)"
    end subroutine probe_finalize

    !> calc_divergence(div, domain, advect_density, horz_only, use_dqdt)
    !! (wind.F90:250-523) of the tile's u, v, w on the device.
    subroutine calc_divergence(tile, div, advect_density, horz_only, use_dqdt)
        type(wind_tile_t), intent(in) :: tile
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(out), target :: div
        logical, intent(in) :: advect_density, horz_only, use_dqdt
        integer(c_int) :: rc
        type(c_ptr) :: rho
        rho = c_null_ptr
        if (advect_density .and. associated(tile%density)) rho = c_loc(tile%density)
        rc = windhip_divergence(slots(tile%nest_indx)%handle, c_loc(tile%u), c_loc(tile%v), c_loc(tile%w), rho, &
                                merge(1_c_int, 0_c_int, advect_density), merge(1_c_int, 0_c_int, horz_only), &
                                merge(1_c_int, 0_c_int, use_dqdt), c_loc(div))
        call check(rc, "windhip_divergence")
    end subroutine calc_divergence

    !> balance_uvw(domain, adv_den, update_in) (wind.F90:123-191): the tile's w
    !! from the horizontal divergence of its u, v (calc_w, :193-234). The tile
    !! holds the arrays update_in selects (data_3d or dqdt_3d).
    subroutine balance_uvw(tile, adv_den)
        type(wind_tile_t), intent(inout) :: tile
        logical, intent(in) :: adv_den
        integer(c_int) :: rc
        type(c_ptr) :: rho
        rho = c_null_ptr
        if (associated(tile%density)) rho = c_loc(tile%density)
        rc = windhip_balance_uvw(slots(tile%nest_indx)%handle, c_loc(tile%u), c_loc(tile%v), c_loc(tile%w), rho, &
                                 merge(1_c_int, 0_c_int, adv_den), 1_c_int)
        call check(rc, "windhip_balance_uvw")
    end subroutine balance_uvw

    !> calc_w_real (wind.F90:1175-1227) of the tile's u, v and a grid-relative
    !! w_grid (absent: zero).
    subroutine calc_w_real(tile, w_real, w_grid)
        type(wind_tile_t), intent(in) :: tile
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(out), target :: w_real
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(in), target, &
            optional :: w_grid
        integer(c_int) :: rc
        type(c_ptr) :: wg
        wg = c_null_ptr
        if (present(w_grid)) wg = c_loc(w_grid)
        rc = windhip_calc_w_real(slots(tile%nest_indx)%handle, c_loc(tile%u), c_loc(tile%v), wg, c_loc(w_real))
        call check(rc, "windhip_calc_w_real")
    end subroutine calc_w_real

    !> calc_idealized_wgrid(domain) (wind.F90:940-998): the tile's w from a
    !! prescribed w_real and its u, v.
    subroutine calc_idealized_wgrid(tile, w_real)
        type(wind_tile_t), intent(inout) :: tile
        real, dimension(tile%ims:tile%ime, tile%kms:tile%kme, tile%jms:tile%jme), intent(in), target :: w_real
        integer(c_int) :: rc
        rc = windhip_idealized_wgrid(slots(tile%nest_indx)%handle, c_loc(tile%u), c_loc(tile%v), c_loc(w_real), &
                                     c_loc(tile%w))
        call check(rc, "windhip_idealized_wgrid")
    end subroutine calc_idealized_wgrid

    !> Load the trained parameters of the learned initial guess (the flat FP32
    !! array of windhip.h, windhip_net_param_count() values) for nest
    !! nest_indx, after its init_iter_winds; solves with options%start =
    !! WINDHIP_START_LEARNED then start from the network's guess when its
    !! residual beats the previous lambda's. No counterpart in the reference.
    subroutine load_learned_guess(nest_indx, params)
        integer, intent(in) :: nest_indx
        real(c_float), intent(in), target :: params(:)
        integer(c_int) :: rc
        if (.not. c_associated(slots(nest_indx)%handle)) error stop "load_learned_guess: call init_iter_winds first"
        if (size(params, kind=c_long) /= windhip_net_param_count()) error stop "load_learned_guess: wrong size"
        if (.not. c_associated(slots(nest_indx)%net)) then
            rc = windhip_net_create(slots(nest_indx)%net)
            call check(rc, "windhip_net_create")
        end if
        rc = windhip_net_set_params(slots(nest_indx)%net, c_loc(params))
        call check(rc, "windhip_net_set_params")
        rc = windhip_set_net(slots(nest_indx)%handle, slots(nest_indx)%net)
        call check(rc, "windhip_set_net")
    end subroutine load_learned_guess

    !> finalize_iter_winds (wind_iterative.F90:2142-2267): release every nest.
    subroutine finalize_iter_winds()
        integer :: n
        integer(c_int) :: rc
        do n = 1, MAX_NESTS
            if (c_associated(slots(n)%handle)) then
                rc = windhip_destroy(slots(n)%handle)
                slots(n)%handle = c_null_ptr
            end if
            if (c_associated(slots(n)%net)) then
                rc = windhip_net_destroy(slots(n)%net)
                slots(n)%net = c_null_ptr
            end if
        end do
    end subroutine finalize_iter_winds

    function last_solve_info() result(info)
        type(windhip_solve_info) :: info
        info = info_last
    end function last_solve_info

    !> The C handle of nest nest_indx (for direct C-ABI calls on it).
    type(c_ptr) function nest_handle(nest_indx)
        integer, intent(in) :: nest_indx
        nest_handle = handle_of(nest_indx)
    end function nest_handle

    type(c_ptr) function handle_of(nest_indx)
        integer, intent(in) :: nest_indx
        if (nest_indx < 1 .or. nest_indx > MAX_NESTS) error stop "wind_iterative_hip: nest index out of range"
        handle_of = slots(nest_indx)%handle
        if (.not. c_associated(handle_of)) error stop "wind_iterative_hip: init_iter_winds has not been called"
    end function handle_of

    !> The address of an optional 2-D field (null when not associated).
    type(c_ptr) function addr2(a)
        real, pointer, contiguous, intent(in) :: a(:,:)
        addr2 = c_null_ptr
        if (associated(a)) addr2 = c_loc(a)
    end function addr2

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= WINDHIP_OK) then
            write(*,*) 'wind_iterative_hip: ', what, ' failed (', rc, '): ', windhip_error_message()
            error stop
        end if
    end subroutine check

end module wind_iterative_hip
