! mod_radiation.f90 -- drop-in for the reference's mod_radiation
! (src/mod_radiation.f90): same module name and public procedures (20-21),
! computed by the gfx950 library through atm_c_api.
module mod_radiation

   use mod_kinds, only: rk
   use atm_c_api
   use atm_device

   implicit none

   private
   public :: allocate_radiation_arrays, calculate_radiative_heating, &
             radiation_halo_exchange

contains

   subroutine allocate_radiation_arrays()
      ! mod_radiation.f90:406-446
      call device_check(atm_radiation_init(device_ctx()), 'allocate_radiation_arrays')
   end subroutine allocate_radiation_arrays

   subroutine calculate_radiative_heating(time)
      ! mod_radiation.f90:41-156: net_flux on the interior columns
      real(rk), intent(in) :: time
      if (.not. device_resident()) then
         call device_push(ATM_DP); call device_push(ATM_PT)
         call device_push(ATM_TS); call device_push(ATM_RADIUS)
      end if
      call device_check(atm_radiation(device_ctx(), time), 'calculate_radiative_heating')
      if (.not. device_resident()) then
         call device_pull(ATM_NET_FLUX)
         call device_pull(ATM_PLEV); call device_pull(ATM_PKAP)
      end if
   end subroutine calculate_radiative_heating

   subroutine radiation_halo_exchange()
      ! mod_radiation.f90:400-404
      call device_check(atm_radiation_halo(device_ctx()), 'radiation_halo_exchange')
      if (.not. device_resident()) call device_pull(ATM_NET_FLUX)
   end subroutine radiation_halo_exchange

end module mod_radiation
