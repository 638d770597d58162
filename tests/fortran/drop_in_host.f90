! drop_in_host.f90 -- a Fortran host that drives the gfx950 library through
! the drop-in modules with the reference's names, in the call order of
! run_model (src/mod_model.f90:115-193; output is not part of this test).
!
! usage: drop_in_host <input file> <output file>
!   input (stream, native endian): int32 nx, ny, nlay, nsteps, resident,
!   physics_on; float64 dx, dy, gravity, f, c_div, c_vort; ak(nlay+1),
!   bk(nlay+1); per step float64 time, dt and int32 radiation, physics,
!   remap flags; then the halo'd arrays dp, pt, ud, vd, ts, gz, plev, pkap,
!   net_flux, radius.
!   output: int32 steps completed, int32 stable; dp, pt, ud, vd, net_flux, ts;
!   the time means of the device output accumulation (nx, ny, nlay+1, 6).
program drop_in_host
   use iso_fortran_env, only: int32, real64
   use mod_kinds, only: ik, rk
   use mod_config, only: config => main_config
   use mod_constants, only: set_constants
   use mod_tiles, only: init_tiles
   use mod_fields, only: allocate_fields, check_stability, dp, pt, ud, vd, ts, gz, plev, pkap, &
                         net_flux, radius, coord_Ak, coord_Bk
   use mod_sync, only: halo_exchange, allocate_sync_buffers
   use mod_sw_dyn, only: init_sw_dyn, cgrid_dynamics_step, cgrid_halo_exchange, &
                         dgrid_dynamics_step, dgrid_halo_exchange
   use mod_radiation, only: allocate_radiation_arrays, calculate_radiative_heating, &
                            radiation_halo_exchange
   use mod_remap, only: allocate_remapping_arrays, perform_vertical_remapping
   use mod_physics, only: allocate_physics_arrays, physics_halo_exchange, &
                          calculate_physics_tendencies, apply_physics_tendencies
   use atm_device, only: device_set_resident, device_check_stability, device_pull_all, device_stop, &
                         device_output_init, device_accumulate_output, device_output_mean
   implicit none

   character(len=512) :: fin, fout
   integer(int32) :: hdr(6), nsteps, resident, physics_on, n, done, stable_out
   real(real64) :: prm(6)
   real(real64), allocatable :: time(:), dt(:)
   integer(int32), allocatable :: rad(:), phys(:), remap(:)
   logical :: stable
   integer :: u
   real(real64), allocatable :: means(:, :, :, :)

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
   read (u) hdr
   config%nx = hdr(1)
   config%ny = hdr(2)
   config%nlay = hdr(3)
   nsteps = hdr(4)
   resident = hdr(5)
   physics_on = hdr(6)
   read (u) prm
   config%dx = prm(1)
   config%dy = prm(2)
   call set_constants(prm(3), prm(4), prm(5), prm(6))
   call init_tiles()
   call allocate_fields()
   read (u) coord_Ak, coord_Bk
   allocate (time(nsteps), dt(nsteps), rad(nsteps), phys(nsteps), remap(nsteps))
   read (u) time, dt, rad, phys, remap
   read (u) dp, pt, ud, vd, ts, gz, plev, pkap, net_flux, radius
   close (u)

   ! init_model order (src/mod_model.f90:59-76)
   call allocate_sync_buffers()
   call halo_exchange(dp, pt, gz, ud, vd)
   call halo_exchange(ts)
   call init_sw_dyn()
   call allocate_radiation_arrays()
   call allocate_physics_arrays()
   call allocate_remapping_arrays()
   if (resident /= 0) call device_set_resident(.true.)
   call device_output_init()

   done = 0
   stable = .true.
   do n = 1, nsteps
      if (rad(n) /= 0) then
         call calculate_radiative_heating(time(n))
         call radiation_halo_exchange()
      end if
      if (phys(n) /= 0) call calculate_physics_tendencies()
      call cgrid_dynamics_step(dt(n)/2.)
      call cgrid_halo_exchange()
      call dgrid_dynamics_step(dt(n))
      if (physics_on /= 0) call apply_physics_tendencies(dt(n))
      if (remap(n) /= 0) call perform_vertical_remapping()
      call dgrid_halo_exchange()
      if (physics_on /= 0) call physics_halo_exchange()
      if (resident /= 0) then
         stable = device_check_stability()
      else
         stable = check_stability()
      end if
      if (.not. stable) exit
      call device_accumulate_output(dt(n))
      done = n
   end do
   allocate (means(config%nx, config%ny, config%nlay + 1, 6))
   call device_output_mean(means)
   if (resident /= 0) call device_pull_all()
   call device_stop()

   stable_out = merge(1, 0, stable)
   open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
   write (u) done, stable_out
   write (u) dp, pt, ud, vd, net_flux, ts
   write (u) means
   close (u)
   print '(a,i0,a)', 'drop_in_host: ', done, ' steps'
end program drop_in_host
