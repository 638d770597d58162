! atm_c_api.f90 -- ISO_C_BINDING interfaces to libatmos_hip.so (include/atm.h).
!
! A Fortran host binds the GPU library through this module; the drop-in
! modules in this folder (mod_sw_dyn, mod_remap, mod_radiation, mod_physics,
! mod_sync) call it with the reference's public procedure names.
module atm_c_api
   use iso_c_binding, only: c_int, c_int32_t, c_int64_t, c_double, c_ptr, c_null_ptr, c_char
   implicit none
   private

   public :: atm_grid, atm_params
   public :: atm_create, atm_destroy, atm_last_error, atm_set_coordinate, &
             atm_put_field, atm_get_field, atm_init_prognostic_fields, &
             atm_initial_halo_exchange, atm_init_sw_dyn, atm_cgrid_step, &
             atm_cgrid_halo, atm_dgrid_step, atm_dgrid_halo, atm_remap_init, &
             atm_remap, atm_radiation_init, atm_radiation, atm_radiation_halo, &
             atm_halo_exchange, atm_check_stability, atm_sync, atm_physics_init, &
             atm_physics_tendencies, atm_apply_physics, atm_physics_halo, &
             atm_output_init, atm_accumulate_output, atm_output_mean
   public :: ATM_DP, ATM_PT, ATM_UD, ATM_VD, ATM_TS, ATM_GZ, ATM_NET_FLUX, &
             ATM_PLEV, ATM_PKAP, ATM_RADIUS

   integer(c_int), parameter :: ATM_DP = 0, ATM_PT = 1, ATM_UD = 2, ATM_VD = 3, &
                                ATM_TS = 4, ATM_GZ = 5, ATM_NET_FLUX = 6, &
                                ATM_PLEV = 7, ATM_PKAP = 8, ATM_RADIUS = 9

   type, bind(c) :: atm_grid
      integer(c_int32_t) :: nx, ny, nlay
      real(c_double) :: dx, dy
   end type atm_grid

   type, bind(c) :: atm_params
      real(c_double) :: gravity, coriolis_parameter, &
                        divergence_damping_coefficient, &
                        vorticity_damping_coefficient
   end type atm_params

   interface
      integer(c_int) function atm_create(grid, params, precision, ctx) bind(c, name='atm_create')
         import :: c_int, c_ptr, atm_grid, atm_params
         type(atm_grid), intent(in) :: grid
         type(atm_params), intent(in) :: params
         integer(c_int), value :: precision
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function atm_destroy(ctx) bind(c, name='atm_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function atm_last_error(ctx) bind(c, name='atm_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_set_coordinate(ctx, ak, bk) bind(c, name='atm_set_coordinate')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: ak(*), bk(*)
      end function
      integer(c_int) function atm_put_field(ctx, field, host) bind(c, name='atm_put_field')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field
         real(c_double), intent(in) :: host(*)
      end function
      integer(c_int) function atm_get_field(ctx, field, host) bind(c, name='atm_get_field')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: field
         real(c_double), intent(inout) :: host(*)
      end function
      integer(c_int) function atm_init_prognostic_fields(ctx) bind(c, name='atm_init_prognostic_fields')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_initial_halo_exchange(ctx) bind(c, name='atm_initial_halo_exchange')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_init_sw_dyn(ctx) bind(c, name='atm_init_sw_dyn')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_cgrid_step(ctx, dt) bind(c, name='atm_cgrid_step')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
      end function
      integer(c_int) function atm_cgrid_halo(ctx) bind(c, name='atm_cgrid_halo')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_dgrid_step(ctx, dt) bind(c, name='atm_dgrid_step')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
      end function
      integer(c_int) function atm_dgrid_halo(ctx) bind(c, name='atm_dgrid_halo')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_remap_init(ctx) bind(c, name='atm_remap_init')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_remap(ctx) bind(c, name='atm_remap')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_radiation_init(ctx) bind(c, name='atm_radiation_init')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_radiation(ctx, time) bind(c, name='atm_radiation')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: time
      end function
      integer(c_int) function atm_radiation_halo(ctx) bind(c, name='atm_radiation_halo')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_physics_init(ctx) bind(c, name='atm_physics_init')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_physics_tendencies(ctx) bind(c, name='atm_physics_tendencies')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_apply_physics(ctx, dt) bind(c, name='atm_apply_physics')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
      end function
      integer(c_int) function atm_physics_halo(ctx) bind(c, name='atm_physics_halo')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_output_init(ctx) bind(c, name='atm_output_init')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function atm_accumulate_output(ctx, dt) bind(c, name='atm_accumulate_output')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
      end function
      integer(c_int) function atm_output_mean(ctx, host) bind(c, name='atm_output_mean')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: host(*)
      end function
      integer(c_int) function atm_halo_exchange(ctx, field) bind(c, name='atm_halo_exchange')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: field
      end function
      integer(c_int) function atm_check_stability(ctx, stable) bind(c, name='atm_check_stability')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: stable
      end function
      integer(c_int) function atm_sync(ctx) bind(c, name='atm_sync')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
   end interface
end module atm_c_api
