!>----------------------------------------------------------------------------
!! A minimal stand-in for the three HICAR modules the `wind_iterative`
!! adapter uses (icar_constants, domain_interface, options_interface), so the
!! adapter compiles and runs here against the same member names HICAR
!! declares:
!!   kVARS%<name> / kITERATIVE_WINDS   src/constants/icar_constants.F90:440, 578
!!   variable_t data_3d, data_2d, dqdt_3d   src/objects/variable_h.F90:12-17
!!   index_type v                      src/objects/data_structures.F90:23-26
!!   domain_t vars_2d, vars_3d, var_indx, dx, memory/domain bounds, grid,
!!     nest_indx, mapfac_*              src/objects/domain_h.F90:40-110, flow_h.F90:20
!!   options_t general%debug, physics%windtype, wind%wind_solver_iterations
!!                                      src/objects/options_h.F90:27-43, opt_types.F90:16-46
!! Only what the adapter reads is declared; the index values are arbitrary.
!!----------------------------------------------------------------------------
module icar_constants
    implicit none
    type var_constants_type
        integer :: u = 1, v = 2, w = 3, density = 4
        integer :: dzdx = 5, dzdy = 6, jacobian = 7, jacobian_u = 8, jacobian_v = 9, jacobian_w = 10
        integer :: dzdx_u = 11, dzdy_v = 12, advection_dz = 13, neighbor_terrain = 14
    end type var_constants_type
    type(var_constants_type) :: kVARS
    integer, parameter :: kMAX_STORAGE_VARS = 14
    integer, parameter :: kITERATIVE_WINDS = 1
end module icar_constants

module domain_interface
    use icar_constants, only : kMAX_STORAGE_VARS
    implicit none
    type variable_t
        real, allocatable :: data_3d(:,:,:)
        real, allocatable :: data_2d(:,:)
        real, allocatable :: dqdt_3d(:,:,:)
    end type variable_t
    type index_type
        integer :: v = -1
        integer :: id = -1
    end type index_type
    type grid_t
        integer :: ids, ide, jds, jde, kds, kde
    end type grid_t
    type domain_t
        integer :: nest_indx = 0
        type(grid_t) :: grid
        real :: dx
        type(variable_t), allocatable :: vars_2d(:)
        type(variable_t), allocatable :: vars_3d(:)
        type(index_type) :: var_indx(kMAX_STORAGE_VARS)
        integer :: ids, ide, jds, jde, kds, kde
        integer :: ims, ime, jms, jme, kms, kme
        real, allocatable :: mapfac_mx_u(:,:), mapfac_my_u(:,:)
        real, allocatable :: mapfac_mx_v(:,:), mapfac_my_v(:,:)
        real, allocatable :: mapfac_mxy(:,:)
    end type domain_t
end module domain_interface

module options_interface
    implicit none
    type general_options_type
        logical :: debug = .false.
    end type general_options_type
    type physics_type
        integer :: windtype = 0
    end type physics_type
    type wind_type
        integer :: wind_solver_iterations = 1500
    end type wind_type
    type options_t
        type(general_options_type) :: general
        type(physics_type) :: physics
        type(wind_type) :: wind
    end type options_t
end module options_interface
