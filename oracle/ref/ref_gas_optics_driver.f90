! Test harness (oracle support, never shipped): evaluates the reference's own
! public gas-optics functions and constants, compiled unmodified from
! src/mod_gas_optics.f90 and src/mod_constants.f90, on query points read from
! a little-endian binary file.
!   input  : int32 n, then n records of (real64 kind, g, p, T)
!            kind 1 = shortwave_absorption_coefficient(g, p, T)
!                 2 = shortwave_scattering_coefficient(g)
!                 3 = longwave_absorption_coefficient(g, p, T)
!                 4 = planck_function(g, T)
!   output : n real64 results, then pi, R_dry, cp_dry, cv_dry, kappa
! Queries must be inside the tabulated range (the reference aborts otherwise).
program ref_gas_optics_driver
   use mod_kinds, only: ik, rk
   use mod_gas_optics, only: shortwave_absorption_coefficient, &
                             shortwave_scattering_coefficient, &
                             longwave_absorption_coefficient, planck_function
   use mod_constants, only: pi, R_dry, cp_dry, cv_dry, kappa
   implicit none
   character(len=512) :: fin, fout
   integer :: uin, uout, n, i, g
   real(rk) :: rec(4)
   real(rk), allocatable :: res(:)

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read (uin) n
   allocate (res(n))
   do i = 1, n
      read (uin) rec
      g = nint(rec(2))
      select case (nint(rec(1)))
      case (1)
         res(i) = shortwave_absorption_coefficient(g, rec(3), rec(4))
      case (2)
         res(i) = shortwave_scattering_coefficient(g)
      case (3)
         res(i) = longwave_absorption_coefficient(g, rec(3), rec(4))
      case (4)
         res(i) = planck_function(g, rec(4))
      case default
         res(i) = 0
      end select
   end do
   close (uin)
   open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (uout) res
   write (uout) pi, R_dry, cp_dry, cv_dry, kappa
   close (uout)
end program ref_gas_optics_driver
