!>----------------------------------------------------------------------------
!! Drop-in for HICAR's module `wind_iterative` (src/physics/wind_iterative.F90)
!! with its exact public signatures:
!!   init_iter_winds(domain, options)                     :215
!!   calc_iter_winds(domain, alpha_in, div_in, adv_den)   :295
!!   finalize_iter_winds()                                :2142
!! over libwindhip.so (through wind_iterative_hip). It reads from domain_t
!! and options_t exactly the members the reference reads: the metrics in
!! vars_3d(var_indx(kVARS%...)%v)%data_3d, the neighbour terrain in vars_2d,
!! u, v, w in dqdt_3d, density, dx, the memory bounds and nest_indx;
!! options%wind%wind_solver_iterations, options%general%debug and
!! options%physics%windtype.
!!
!! Multi-nest cache (:134-181, :2613-2851): every nest keeps its own solver
!! state (operator, warm-start lambda, probed operator) on the GPU, one handle
!! per nest. init_iter_winds with the active nest returns at once; with another
!! nest it makes that nest the active one, setting it up on its first visit
!! (O(1) afterwards); calc_iter_winds solves for the active nest, as the
!! reference's module state does.
!!
!! The solver choices of this build that options_t does not carry (precision,
!! start, tolerances) are set with set_wind_solver_choice before a nest's
!! first init_iter_winds; the defaults are the reference's algorithm and
!! tolerances.
!!----------------------------------------------------------------------------
module wind_iterative
    use iso_c_binding
    use domain_interface,  only : domain_t
    use icar_constants,    only : kVARS, kITERATIVE_WINDS
    use options_interface, only : options_t
    use windhip_interface, only : windhip_solve_info, WINDHIP_PLAIN_F64, WINDHIP_START_WARM
    use wind_iterative_hip, only : wind_tile_t, wind_solver_options_t, tile_init => init_iter_winds, &
                                   tile_calc => calc_iter_winds, tile_finalize => finalize_iter_winds, &
                                   tile_probe => probe_operator, last_solve_info, &
                                   tile_pattern => probe_lambda_pattern, tile_zero => probe_zero_corrections, &
                                   tile_apply => probe_apply_corrections, tile_record => probe_record, &
                                   tile_finalize_probe => probe_finalize
    implicit none
    private

    public :: init_iter_winds, calc_iter_winds, finalize_iter_winds
    ! the probing API that calibrate_projection_operator (wind.F90:1010-1099) uses
    public :: probe_lambda_pattern, probe_zero_corrections, probe_apply_corrections, probe_record, probe_finalize
    public :: probe_operator, set_wind_solver_choice, wind_tile_of, active_nest, last_solve_info

    integer, parameter :: MAX_NESTS = 4   ! wind_iterative.F90:139
    integer :: active_nest_indx = -1      ! wind_iterative.F90:140

    integer :: choice_precision = WINDHIP_PLAIN_F64
    integer :: choice_start = WINDHIP_START_WARM
    real(c_double) :: choice_tol_rel = 1.0e-5_c_double   ! wind_iterative.F90:60
    real(c_double) :: choice_tol_abs = 1.0e-10_c_double  ! wind_iterative.F90:59

contains

    !> The wind_tile_t view of a domain: pointers into its arrays, nothing copied.
    !! Sizes as init_module_vars takes them (wind_iterative.F90:2279-2298):
    !! the domain extent from domain%grid, the levels from domain%kde. The
    !! library solves one tile on one GPU, so the memory bounds must span the
    !! domain (ims = ids, ime = ide, jms = jds, jme = jde).
    subroutine wind_tile_of(domain, tile)
        type(domain_t), intent(in), target :: domain
        type(wind_tile_t), intent(out) :: tile
        if (domain%ims /= domain%grid%ids .or. domain%ime /= domain%grid%ide .or. &
            domain%jms /= domain%grid%jds .or. domain%jme /= domain%grid%jde) &
            error stop "wind_iterative: the domain is split into tiles; this library solves one tile per GPU"
        tile%nest_indx = domain%nest_indx
        tile%nx = domain%grid%ide - domain%grid%ids + 1
        tile%ny = domain%grid%jde - domain%grid%jds + 1
        tile%nz = domain%kde
        tile%dx = domain%dx
        tile%ims = domain%ims; tile%ime = domain%ime
        tile%kms = domain%kms; tile%kme = domain%kme
        tile%jms = domain%jms; tile%jme = domain%jme
        associate(v => domain%vars_3d, x => domain%var_indx)
            tile%dzdx         => v(x(kVARS%dzdx)%v)%data_3d
            tile%dzdy         => v(x(kVARS%dzdy)%v)%data_3d
            tile%jacobian     => v(x(kVARS%jacobian)%v)%data_3d
            tile%jacobian_u   => v(x(kVARS%jacobian_u)%v)%data_3d
            tile%jacobian_v   => v(x(kVARS%jacobian_v)%v)%data_3d
            tile%jacobian_w   => v(x(kVARS%jacobian_w)%v)%data_3d
            tile%dzdx_u       => v(x(kVARS%dzdx_u)%v)%data_3d
            tile%dzdy_v       => v(x(kVARS%dzdy_v)%v)%data_3d
            tile%advection_dz => v(x(kVARS%advection_dz)%v)%data_3d
            tile%density      => v(x(kVARS%density)%v)%data_3d
            tile%u            => v(x(kVARS%u)%v)%dqdt_3d
            tile%v            => v(x(kVARS%v)%v)%dqdt_3d
            tile%w            => v(x(kVARS%w)%v)%dqdt_3d
        end associate
        tile%terrain => domain%vars_2d(domain%var_indx(kVARS%neighbor_terrain)%v)%data_2d
        ! map factors, as calc_updated_winds (wind_iterative.F90:1796-1797) and
        ! calc_divergence (wind.F90:283-285) read them
        if (allocated(domain%mapfac_mx_u)) tile%mapfac_mx_u => domain%mapfac_mx_u
        if (allocated(domain%mapfac_my_u)) tile%mapfac_my_u => domain%mapfac_my_u
        if (allocated(domain%mapfac_my_v)) tile%mapfac_my_v => domain%mapfac_my_v
        if (allocated(domain%mapfac_mx_v)) tile%mapfac_mx_v => domain%mapfac_mx_v
        if (allocated(domain%mapfac_mxy))  tile%mapfac_mxy  => domain%mapfac_mxy
    end subroutine wind_tile_of

    !> Solver choices of this build for nests set up afterwards (no counterpart
    !! in the reference): precision WINDHIP_PLAIN_F64 (default), WINDHIP_MIXED,
    !! WINDHIP_MIXED_LINE or WINDHIP_MIXED_MG; start WINDHIP_START_WARM
    !! (default) or another WINDHIP_START_*; tolerances (defaults 1e-5, 1e-10).
    subroutine set_wind_solver_choice(precision, start, tol_rel, tol_abs)
        integer, intent(in) :: precision, start
        real(c_double), intent(in), optional :: tol_rel, tol_abs
        choice_precision = precision
        choice_start = start
        if (present(tol_rel)) choice_tol_rel = tol_rel
        if (present(tol_abs)) choice_tol_abs = tol_abs
    end subroutine set_wind_solver_choice

    !> init_iter_winds(domain, options) (wind_iterative.F90:215-291).
    subroutine init_iter_winds(domain, options)
        type(domain_t), intent(in), target :: domain
        type(options_t), intent(in) :: options
        type(wind_tile_t) :: tile
        type(wind_solver_options_t) :: o
        integer :: target_nest

        target_nest = domain%nest_indx
        if (target_nest == active_nest_indx) return
        if (target_nest < 1 .or. target_nest > MAX_NESTS) error stop "init_iter_winds: nest index out of range"
        call wind_tile_of(domain, tile)
        o%wind_solver_iterations = options%wind%wind_solver_iterations
        o%verbose = options%general%debug .or. (options%physics%windtype == kITERATIVE_WINDS)
        o%precision = choice_precision
        o%start = choice_start
        o%tol_rel = choice_tol_rel
        o%tol_abs = choice_tol_abs
        call tile_init(tile, o)   ! returns at once for a nest set up before
        active_nest_indx = target_nest
    end subroutine init_iter_winds

    !> calc_iter_winds(domain, alpha_in, div_in, adv_den) (wind_iterative.F90:295-532):
    !! solve for lambda on the active nest and correct u, v, w (dqdt_3d) in place.
    subroutine calc_iter_winds(domain, alpha_in, div_in, adv_den)
        type(domain_t), intent(inout), target :: domain
        real, dimension(domain%ims:domain%ime, domain%kms:domain%kme, domain%jms:domain%jme), intent(in), &
            target :: alpha_in, div_in
        logical, intent(in) :: adv_den
        type(wind_tile_t) :: tile
        if (active_nest_indx < 1) error stop "calc_iter_winds: init_iter_winds has not been called"
        call wind_tile_of(domain, tile)
        tile%nest_indx = active_nest_indx
        call tile_calc(tile, alpha_in, div_in, adv_den)
    end subroutine calc_iter_winds

    !> calibrate_projection_operator (wind.F90:1010-1099) for the active nest in
    !! one call: the 27 probing rounds run on the GPU and the nest switches to
    !! the probed operator A = 2 D(G(.)).
    subroutine probe_operator(domain, alpha_in, adv_den, max_leak)
        type(domain_t), intent(inout), target :: domain
        real, dimension(domain%ims:domain%ime, domain%kms:domain%kme, domain%jms:domain%jme), intent(in), &
            target :: alpha_in
        logical, intent(in) :: adv_den
        real(c_double), intent(out) :: max_leak
        type(wind_tile_t) :: tile
        if (active_nest_indx < 1) error stop "probe_operator: init_iter_winds has not been called"
        call wind_tile_of(domain, tile)
        tile%nest_indx = active_nest_indx
        call tile_probe(tile, alpha_in, adv_den, max_leak)
    end subroutine probe_operator

    !> probe_lambda_pattern(ca, cb, cc) (wind_iterative.F90:1096-1116): lambda of
    !! the active nest = 1 on the physical cells of colour (ca, cb, cc), else 0.
    subroutine probe_lambda_pattern(ca, cb, cc)
        integer, intent(in) :: ca, cb, cc
        if (active_nest_indx < 1) error stop "probe_lambda_pattern: init_iter_winds has not been called"
        call tile_pattern(active_nest_indx, ca, cb, cc)
    end subroutine probe_lambda_pattern

    !> probe_zero_corrections(domain) (:1118-1153): u, v, w dqdt_3d set to zero.
    subroutine probe_zero_corrections(domain)
        type(domain_t), intent(inout), target :: domain
        type(wind_tile_t) :: tile
        call wind_tile_of(domain, tile)
        call tile_zero(tile)
    end subroutine probe_zero_corrections

    !> probe_apply_corrections(domain, adv_den) (:1155-1161): calc_updated_winds
    !! of the active nest's lambda into u, v, w dqdt_3d.
    subroutine probe_apply_corrections(domain, adv_den)
        type(domain_t), intent(inout), target :: domain
        logical, intent(in) :: adv_den
        type(wind_tile_t) :: tile
        if (active_nest_indx < 1) error stop "probe_apply_corrections: init_iter_winds has not been called"
        call wind_tile_of(domain, tile)
        tile%nest_indx = active_nest_indx
        call tile_apply(tile, adv_den)
    end subroutine probe_apply_corrections

    !> probe_record(domain, Tfield, ca, cb, cc, max_leak) (:1163-1230): 2 Tfield
    !! into the active nest's probed coefficients; max_leak raised to the largest
    !! response off the 15-point stencil.
    subroutine probe_record(domain, Tfield, ca, cb, cc, max_leak)
        type(domain_t), intent(in), target :: domain
        real, intent(in) :: Tfield(domain%ims:domain%ime, domain%kms:domain%kme, domain%jms:domain%jme)
        integer, intent(in) :: ca, cb, cc
        real, intent(inout) :: max_leak
        type(wind_tile_t) :: tile
        if (active_nest_indx < 1) error stop "probe_record: init_iter_winds has not been called"
        call wind_tile_of(domain, tile)
        tile%nest_indx = active_nest_indx
        call tile_record(tile, Tfield, ca, cb, cc, max_leak)
    end subroutine probe_record

    !> probe_finalize(max_leak) (:1232-1248): the active nest solves with the
    !! probed operator from now on; its lambda is reset.
    subroutine probe_finalize(max_leak)
        real, intent(in) :: max_leak
        if (active_nest_indx < 1) error stop "probe_finalize: init_iter_winds has not been called"
        call tile_finalize_probe(active_nest_indx, max_leak)
    end subroutine probe_finalize

    !> finalize_iter_winds() (wind_iterative.F90:2142-2267): every nest's state.
    subroutine finalize_iter_winds()
        call tile_finalize()
        active_nest_indx = -1
    end subroutine finalize_iter_winds

    integer function active_nest()
        active_nest = active_nest_indx
    end function active_nest

end module wind_iterative
