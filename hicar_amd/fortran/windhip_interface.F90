!>----------------------------------------------------------------------------
!! ISO_C_BINDING interface to libwindhip.so (include/windhip.h).
!!
!! Thin, one-to-one: every C entry point has a bind(C) interface here with the
!! same name, and every C struct a bind(C) derived type with the same fields in
!! the same order. Integer return codes as in C (WINDHIP_OK = 0). Field arrays
!! are passed by address (c_loc), Fortran column-major (i,k,j), so HICAR's
!! arrays go through without copies or transposes.
!!----------------------------------------------------------------------------
module windhip_interface
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: WINDHIP_OK = 0, WINDHIP_ERR_ARG = 1, WINDHIP_ERR_HIP = 2, &
                                         WINDHIP_ERR_NAN = 3, WINDHIP_ERR_DIVERGED = 4, WINDHIP_ERR_NOMEM = 5
    integer(c_int), parameter, public :: WINDHIP_F32 = 4, WINDHIP_F64 = 8
    integer(c_int), parameter, public :: WINDHIP_HOST = 0, WINDHIP_DEVICE = 1
    integer(c_int), parameter, public :: WINDHIP_START_WARM = 0, WINDHIP_START_ZERO = 1, &
                                         WINDHIP_START_GIVEN = 2, WINDHIP_START_LEARNED = 3
    integer(c_int), parameter, public :: WINDHIP_PLAIN_F64 = 0, WINDHIP_MIXED = 1, WINDHIP_MIXED_LINE = 2, &
                                         WINDHIP_MIXED_MG = 3
    integer(c_int), parameter, public :: WINDHIP_PROFILE_KINDS = 26

    type, bind(C), public :: windhip_desc
        integer(c_int) :: nx, ny, nz
        real(c_double) :: dx
        integer(c_int) :: ims, ime, kms, kme, jms, jme
        integer(c_int) :: field_dtype, field_mem
    end type windhip_desc

    type, bind(C), public :: windhip_solve_opts
        real(c_double) :: tol_rel, tol_abs
        integer(c_int) :: max_iters, start, precision
        real(c_double) :: inner_tol
        integer(c_int) :: inner_max_iters, max_outer, iters_per_sync, use_graph, retry, profile
    end type windhip_solve_opts

    type, bind(C), public :: windhip_solve_info
        integer(c_int) :: status, iters, nsweeps, outer_iters
        real(c_double) :: res0, res_final, bnorm, true_rel_res, ms
        integer(c_int) :: start_used, reserved
        real(c_double) :: guess_res, plain_res
    end type windhip_solve_info

    public :: windhip_default_opts, windhip_create, windhip_destroy, windhip_last_error, &
              windhip_set_metrics, windhip_solve, windhip_apply_correction, windhip_calc_iter_winds, &
              windhip_divergence, windhip_get_lambda, windhip_set_lambda, windhip_probe_operator, &
              windhip_net_param_count, windhip_net_create, windhip_net_destroy, windhip_net_set_params, &
              windhip_set_net, windhip_set_multigrid, windhip_balance_uvw, windhip_calc_w_real, &
              windhip_idealized_wgrid, windhip_error_message, windhip_probe_lambda_pattern, &
              windhip_probe_record, windhip_probe_finalize, windhip_apply_operator, windhip_profile

    interface
        subroutine windhip_default_opts(opts) bind(C, name="windhip_default_opts")
            import :: windhip_solve_opts
            type(windhip_solve_opts), intent(out) :: opts
        end subroutine

        integer(c_int) function windhip_create(desc, handle) bind(C, name="windhip_create")
            import :: c_int, c_ptr, windhip_desc
            type(windhip_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: handle
        end function

        integer(c_int) function windhip_destroy(handle) bind(C, name="windhip_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        type(c_ptr) function windhip_last_error() bind(C, name="windhip_last_error")
            import :: c_ptr
        end function

        integer(c_int) function windhip_set_metrics(handle, dzdx, dzdy, jacobian, jacobian_u, jacobian_v, &
                jacobian_w, dzdx_u, dzdy_v, advection_dz, terrain, mapfac_mx_u, mapfac_my_u, mapfac_my_v, &
                mapfac_mx_v, mapfac_mxy) bind(C, name="windhip_set_metrics")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, dzdx, dzdy, jacobian, jacobian_u, jacobian_v, jacobian_w, &
                                  dzdx_u, dzdy_v, advection_dz, terrain, mapfac_mx_u, mapfac_my_u, &
                                  mapfac_my_v, mapfac_mx_v, mapfac_mxy
        end function

        integer(c_int) function windhip_solve(handle, alpha, div, opts, info) bind(C, name="windhip_solve")
            import :: c_int, c_ptr, windhip_solve_opts, windhip_solve_info
            type(c_ptr), value :: handle, alpha, div
            type(windhip_solve_opts), intent(in) :: opts
            type(windhip_solve_info), intent(out) :: info
        end function

        integer(c_int) function windhip_apply_correction(handle, u, v, w, density, adv_den) &
                bind(C, name="windhip_apply_correction")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, u, v, w, density
            integer(c_int), value :: adv_den
        end function

        integer(c_int) function windhip_calc_iter_winds(handle, alpha, div, u, v, w, density, adv_den, opts, info) &
                bind(C, name="windhip_calc_iter_winds")
            import :: c_int, c_ptr, windhip_solve_opts, windhip_solve_info
            type(c_ptr), value :: handle, alpha, div, u, v, w, density
            integer(c_int), value :: adv_den
            type(windhip_solve_opts), intent(in) :: opts
            type(windhip_solve_info), intent(out) :: info
        end function

        integer(c_int) function windhip_divergence(handle, u, v, w, density, adv_den, horz_only, use_dqdt, div_out) &
                bind(C, name="windhip_divergence")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, u, v, w, density, div_out
            integer(c_int), value :: adv_den, horz_only, use_dqdt
        end function

        integer(c_int) function windhip_balance_uvw(handle, u, v, w, density, adv_den, use_dqdt) &
                bind(C, name="windhip_balance_uvw")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, u, v, w, density
            integer(c_int), value :: adv_den, use_dqdt
        end function

        integer(c_int) function windhip_calc_w_real(handle, u, v, w_grid, w_real) bind(C, name="windhip_calc_w_real")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, u, v, w_grid, w_real
        end function

        integer(c_int) function windhip_idealized_wgrid(handle, u, v, w_real, w) bind(C, name="windhip_idealized_wgrid")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, u, v, w_real, w
        end function

        integer(c_int) function windhip_get_lambda(handle, lambda_rows) bind(C, name="windhip_get_lambda")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, lambda_rows
        end function

        integer(c_int) function windhip_set_lambda(handle, lambda_rows) bind(C, name="windhip_set_lambda")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, lambda_rows
        end function

        integer(c_int) function windhip_probe_operator(handle, alpha, density, adv_den, max_leak) &
                bind(C, name="windhip_probe_operator")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle, alpha, density
            integer(c_int), value :: adv_den
            real(c_double), intent(out) :: max_leak
        end function

        ! the reference's probing hooks one by one (wind_iterative.F90:1096-1248)
        integer(c_int) function windhip_probe_lambda_pattern(handle, ca, cb, cc) &
                bind(C, name="windhip_probe_lambda_pattern")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: ca, cb, cc
        end function

        integer(c_int) function windhip_probe_record(handle, t, ca, cb, cc, max_leak) &
                bind(C, name="windhip_probe_record")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle, t
            integer(c_int), value :: ca, cb, cc
            real(c_double), intent(inout) :: max_leak
        end function

        integer(c_int) function windhip_probe_finalize(handle) bind(C, name="windhip_probe_finalize")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function

        ! y = A x with the solver's current operator, rows (0:nx+1, 0:nz+1, 0:ny+1) FP64
        integer(c_int) function windhip_apply_operator(handle, precision, x_rows, y_rows) &
                bind(C, name="windhip_apply_operator")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
            integer(c_int), value :: precision
            type(c_ptr), value :: x_rows, y_rows
        end function

        ! learned initial guess: load trained parameters and attach the network
        integer(c_long) function windhip_net_param_count() bind(C, name="windhip_net_param_count")
            import :: c_long
        end function

        integer(c_int) function windhip_net_create(net) bind(C, name="windhip_net_create")
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: net
        end function

        integer(c_int) function windhip_net_destroy(net) bind(C, name="windhip_net_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: net
        end function

        integer(c_int) function windhip_net_set_params(net, params) bind(C, name="windhip_net_set_params")
            import :: c_int, c_ptr
            type(c_ptr), value :: net, params
        end function

        integer(c_int) function windhip_set_net(handle, net) bind(C, name="windhip_set_net")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle, net
        end function

        integer(c_int) function windhip_profile(handle, reset, calls, ms) bind(C, name="windhip_profile")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle
            integer(c_int), value :: reset
            real(c_double), intent(out) :: calls(*), ms(*)
        end function

        integer(c_int) function windhip_set_multigrid(handle, pre, post, coarse_pre, coarse_post, depth, &
                                                      omega) bind(C, name="windhip_set_multigrid")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: handle
            integer(c_int), value :: pre, post, coarse_pre, coarse_post, depth
            real(c_double), value :: omega
        end function
    end interface

contains

    !> The text of windhip_last_error() as a Fortran string.
    function windhip_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n
        interface
            integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
                import :: c_ptr, c_size_t
                type(c_ptr), value :: s
            end function
        end interface
        p = windhip_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        msg = transfer(chars(1:n), msg)
    end function windhip_error_message

end module windhip_interface
