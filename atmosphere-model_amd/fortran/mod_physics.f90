! mod_physics.f90 -- drop-in for the reference's mod_physics
! (src/mod_physics.f90): same module name and public procedures (13-14),
! computed by the gfx950 library through atm_c_api.
module mod_physics

   use mod_kinds, only: rk
   use atm_c_api
   use atm_device

   implicit none

   private
   public :: allocate_physics_arrays, physics_halo_exchange, &
             calculate_physics_tendencies, apply_physics_tendencies

contains

   subroutine allocate_physics_arrays()
      ! mod_physics.f90:193-214
      call device_check(atm_physics_init(device_ctx()), 'allocate_physics_arrays')
   end subroutine allocate_physics_arrays

   subroutine calculate_physics_tendencies()
      ! mod_physics.f90:27-78: dry convective adjustment and surface heating;
      ! plev, pkap, gz (and playkap) are recomputed on is+1..ie-1
      if (.not. device_resident()) then
         call device_push(ATM_DP); call device_push(ATM_PT)
         call device_push(ATM_UD); call device_push(ATM_VD)
         call device_push(ATM_NET_FLUX)
      end if
      call device_check(atm_physics_tendencies(device_ctx()), 'calculate_physics_tendencies')
      if (.not. device_resident()) then
         call device_pull(ATM_PLEV); call device_pull(ATM_PKAP); call device_pull(ATM_GZ)
      end if
   end subroutine calculate_physics_tendencies

   subroutine apply_physics_tendencies(dt)
      ! mod_physics.f90:162-185 on the interior
      real(rk), intent(in) :: dt
      if (.not. device_resident()) then
         call device_push(ATM_PT); call device_push(ATM_UD)
         call device_push(ATM_VD); call device_push(ATM_TS)
      end if
      call device_check(atm_apply_physics(device_ctx(), dt), 'apply_physics_tendencies')
      if (.not. device_resident()) then
         call device_pull(ATM_PT); call device_pull(ATM_UD)
         call device_pull(ATM_VD); call device_pull(ATM_TS)
      end if
   end subroutine apply_physics_tendencies

   subroutine physics_halo_exchange()
      ! mod_physics.f90:187-191: halo of ts
      if (.not. device_resident()) call device_push(ATM_TS)
      call device_check(atm_physics_halo(device_ctx()), 'physics_halo_exchange')
      if (.not. device_resident()) call device_pull(ATM_TS)
   end subroutine physics_halo_exchange

end module mod_physics
