! mod_sw_dyn.f90 -- drop-in for the reference's mod_sw_dyn (src/mod_sw_dyn.f90):
! same module name, same public procedures and arguments (17-19), computed by
! the gfx950 library through atm_c_api.  Link this file instead of the
! reference's; the callers (src/mod_model.f90:12-14, 147-159) stay unchanged.
module mod_sw_dyn

   use mod_kinds, only: rk
   use atm_c_api
   use atm_device

   implicit none

   private
   public :: init_sw_dyn, cgrid_dynamics_step, cgrid_halo_exchange, &
             dgrid_dynamics_step, dgrid_halo_exchange

contains

   subroutine init_sw_dyn()
      ! mod_sw_dyn.f90:958-1023: creates the device context (uploading the
      ! host fields and the coordinate) and derives the step constants.
      call device_start()
      call device_check(atm_init_sw_dyn(device_ctx()), 'init_sw_dyn')
   end subroutine init_sw_dyn

   subroutine cgrid_dynamics_step(dt)
      ! mod_sw_dyn.f90:80-416, called with dt/2 (mod_model.f90:147)
      real(rk), intent(in) :: dt
      if (.not. device_resident()) then
         call device_push(ATM_DP); call device_push(ATM_PT)
         call device_push(ATM_UD); call device_push(ATM_VD)
         call device_push(ATM_NET_FLUX)
      end if
      call device_check(atm_cgrid_step(device_ctx(), dt), 'cgrid_dynamics_step')
   end subroutine cgrid_dynamics_step

   subroutine cgrid_halo_exchange()
      ! mod_sw_dyn.f90:418-422; the C-grid arrays live on the device only
      call device_check(atm_cgrid_halo(device_ctx()), 'cgrid_halo_exchange')
   end subroutine cgrid_halo_exchange

   subroutine dgrid_dynamics_step(dt)
      ! mod_sw_dyn.f90:424-773
      real(rk), intent(in) :: dt
      call device_check(atm_dgrid_step(device_ctx(), dt), 'dgrid_dynamics_step')
      if (.not. device_resident()) then
         call device_pull(ATM_DP); call device_pull(ATM_PT)
         call device_pull(ATM_UD); call device_pull(ATM_VD)
         call device_pull(ATM_PLEV); call device_pull(ATM_PKAP); call device_pull(ATM_GZ)
      end if
   end subroutine dgrid_dynamics_step

   subroutine dgrid_halo_exchange()
      ! mod_sw_dyn.f90:775-779: halo of dp, pt, ud, vd (after the host's
      ! physics tendencies and the remap, mod_model.f90:162-172)
      if (.not. device_resident()) then
         call device_push(ATM_DP); call device_push(ATM_PT)
         call device_push(ATM_UD); call device_push(ATM_VD)
      end if
      call device_check(atm_dgrid_halo(device_ctx()), 'dgrid_halo_exchange')
      if (.not. device_resident()) then
         call device_pull(ATM_DP); call device_pull(ATM_PT)
         call device_pull(ATM_UD); call device_pull(ATM_VD)
      end if
   end subroutine dgrid_halo_exchange

end module mod_sw_dyn
