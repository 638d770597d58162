! mod_remap.f90 -- drop-in for the reference's mod_remap (src/mod_remap.f90):
! same module name and public procedures (13), computed by the gfx950
! library through atm_c_api.
module mod_remap

   use atm_c_api
   use atm_device

   implicit none

   private
   public :: allocate_remapping_arrays, perform_vertical_remapping

contains

   subroutine allocate_remapping_arrays()
      ! mod_remap.f90:358-388
      call device_check(atm_remap_init(device_ctx()), 'allocate_remapping_arrays')
   end subroutine allocate_remapping_arrays

   subroutine perform_vertical_remapping()
      ! mod_remap.f90:27-163: dp, pt, ud, vd on the interior; plev, pkap, gz
      ! recomputed on isd-3..ied+3
      if (.not. device_resident()) then
         call device_push(ATM_DP); call device_push(ATM_PT)
         call device_push(ATM_UD); call device_push(ATM_VD)
      end if
      call device_check(atm_remap(device_ctx()), 'perform_vertical_remapping')
      if (.not. device_resident()) then
         call device_pull(ATM_DP); call device_pull(ATM_PT)
         call device_pull(ATM_UD); call device_pull(ATM_VD)
         call device_pull(ATM_PLEV); call device_pull(ATM_PKAP); call device_pull(ATM_GZ)
      end if
   end subroutine perform_vertical_remapping

end module mod_remap
