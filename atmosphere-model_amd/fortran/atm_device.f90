! atm_device.f90 -- the device context behind the drop-in modules
! (mod_sw_dyn, mod_remap, mod_radiation, mod_physics, mod_sync in this
! folder), the coherence of the host's mod_fields arrays with the device
! copy, and the device form of mod_output's time-mean accumulation.
!
! The reference keeps its state in the module-global allocatables of
! mod_fields (src/mod_fields.f90:15-37) and its callers read and write them
! between the hot-path calls (src/mod_model.f90:115-193).  Two coherence
! modes:
!
!   mirrored (default)  every drop-in call uploads the fields it reads from
!                       the host arrays and downloads the fields it writes,
!                       so an unmodified host (physics tendencies, stability
!                       check, output on the host) sees the reference's
!                       arrays after every call;
!   resident            the device copy is authoritative; nothing moves
!                       unless the host calls device_push / device_pull
!                       (or the *_all forms) around its own reads and writes.
!
! Errors: a non-zero status from the library is reported through the
! reference's logger as fatal, then abort_now, as the reference does for
! its own failures (src/mod_util.f90:13-19).
module atm_device

   use iso_c_binding, only: c_ptr, c_null_ptr, c_associated, c_int, c_char, c_f_pointer, c_null_char
   use mod_kinds, only: ik, rk
   use mod_config, only: config => main_config
   use mod_constants, only: gravity, coriolis_parameter, &
                            divergence_damping_coeff, vorticity_damping_coeff
   use mod_fields, only: dp, pt, ud, vd, ts, gz, plev, pkap, net_flux, &
                         radius, coord_Ak, coord_Bk
   use mod_log, only: logger => main_logger
   use mod_util, only: abort_now
   use atm_c_api

   implicit none

   private
   public :: device_ctx, device_start, device_check, device_resident, &
             device_set_resident, device_push, device_pull, device_push_all, &
             device_pull_all, device_check_stability, device_stop, &
             device_output_init, device_accumulate_output, device_output_mean

   type(c_ptr), save :: ctx = c_null_ptr
   logical, save :: resident = .false.

contains

   function device_ctx() result(c)
      type(c_ptr) :: c
      if (.not. c_associated(ctx)) call device_start()
      c = ctx
   end function device_ctx

   logical function device_resident()
      device_resident = resident
   end function device_resident

   subroutine device_set_resident(on)
      ! Switching to mirrored mode first downloads the device state.
      logical, intent(in) :: on
      if (resident .and. .not. on .and. c_associated(ctx)) call device_pull_all()
      resident = on
   end subroutine device_set_resident

   subroutine device_start()
      ! Creates the context from main_config and mod_constants and uploads
      ! the coordinate and every field the hot path reads.
      type(atm_grid) :: grid
      type(atm_params) :: params
      integer(c_int) :: rc

      if (c_associated(ctx)) return
      grid%nx = config%nx
      grid%ny = config%ny
      grid%nlay = config%nlay
      grid%dx = config%dx
      grid%dy = config%dy
      params%gravity = gravity
      params%coriolis_parameter = coriolis_parameter
      params%divergence_damping_coefficient = divergence_damping_coeff
      params%vorticity_damping_coefficient = vorticity_damping_coeff
      rc = atm_create(grid, params, 0_c_int, ctx)
      if (rc /= 0) then
         call logger%fatal('atm_device', 'could not create the device context (status ' &
                           //trim(itoa(int(rc)))//'): no usable GPU or out of memory')
         call abort_now()
      end if
      call device_check(atm_set_coordinate(ctx, coord_Ak, coord_Bk), 'atm_set_coordinate')
      call device_push_all()
   end subroutine device_start

   subroutine device_stop()
      if (.not. c_associated(ctx)) return
      if (resident) call device_pull_all()
      call device_check(atm_destroy(ctx), 'atm_destroy')
      ctx = c_null_ptr
   end subroutine device_stop

   subroutine device_check(rc, source)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: source
      if (rc == 0) return
      call logger%fatal(source, last_error())
      call abort_now()
   end subroutine device_check

   function last_error() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: n
      p = atm_last_error(ctx)
      msg = 'unknown library error'
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [4096])
      n = 0
      do while (n < 4096)
         if (s(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate (character(len=n) :: msg)
      msg = transfer(s(1:n), msg)
   end function last_error

   pure function itoa(i) result(str)
      integer, intent(in) :: i
      character(len=16) :: str
      write (str, '(i0)') i
   end function itoa

   subroutine device_push(field)
      ! Host array -> device for one field (ATM_* field id of atm_c_api).
      integer(c_int), intent(in) :: field
      integer(c_int) :: rc
      select case (field)
      case (ATM_DP); rc = atm_put_field(ctx, field, dp)
      case (ATM_PT); rc = atm_put_field(ctx, field, pt)
      case (ATM_UD); rc = atm_put_field(ctx, field, ud)
      case (ATM_VD); rc = atm_put_field(ctx, field, vd)
      case (ATM_TS); rc = atm_put_field(ctx, field, ts)
      case (ATM_GZ); rc = atm_put_field(ctx, field, gz)
      case (ATM_NET_FLUX); rc = atm_put_field(ctx, field, net_flux)
      case (ATM_PLEV); rc = atm_put_field(ctx, field, plev)
      case (ATM_PKAP); rc = atm_put_field(ctx, field, pkap)
      case (ATM_RADIUS); rc = atm_put_field(ctx, field, radius)
      case default; rc = 1
      end select
      call device_check(rc, 'device_push')
   end subroutine device_push

   subroutine device_pull(field)
      ! Device -> host array for one field.
      integer(c_int), intent(in) :: field
      integer(c_int) :: rc
      select case (field)
      case (ATM_DP); rc = atm_get_field(ctx, field, dp)
      case (ATM_PT); rc = atm_get_field(ctx, field, pt)
      case (ATM_UD); rc = atm_get_field(ctx, field, ud)
      case (ATM_VD); rc = atm_get_field(ctx, field, vd)
      case (ATM_TS); rc = atm_get_field(ctx, field, ts)
      case (ATM_GZ); rc = atm_get_field(ctx, field, gz)
      case (ATM_NET_FLUX); rc = atm_get_field(ctx, field, net_flux)
      case (ATM_PLEV); rc = atm_get_field(ctx, field, plev)
      case (ATM_PKAP); rc = atm_get_field(ctx, field, pkap)
      case (ATM_RADIUS); rc = atm_get_field(ctx, field, radius)
      case default; rc = 1
      end select
      call device_check(rc, 'device_pull')
   end subroutine device_pull

   subroutine device_push_all()
      integer(c_int) :: f
      do f = ATM_DP, ATM_RADIUS
         call device_push(f)
      end do
   end subroutine device_push_all

   subroutine device_pull_all()
      integer(c_int) :: f
      do f = ATM_DP, ATM_RADIUS
         call device_pull(f)
      end do
   end subroutine device_pull_all

   logical function device_check_stability()
      ! check_stability (src/mod_fields.f90:69-106) on the device copy: the
      ! prognostic interiors are all finite.  For resident mode.
      integer(c_int) :: stable
      call device_check(atm_check_stability(device_ctx(), stable), 'check_stability')
      device_check_stability = stable /= 0
   end function device_check_stability

   subroutine device_output_init()
      ! init_output's accumulators (src/mod_output.f90:37-55) on the device
      call device_check(atm_output_init(device_ctx()), 'init_output')
   end subroutine device_output_init

   subroutine device_accumulate_output(dt)
      ! accumulate_output (src/mod_output.f90:94-130) on the device copy,
      ! after dgrid_halo_exchange
      real(rk), intent(in) :: dt
      call device_check(atm_accumulate_output(device_ctx(), dt), 'accumulate_output')
   end subroutine device_accumulate_output

   subroutine device_output_mean(output_field)
      ! the time means write_output writes (src/mod_output.f90:144-154):
      ! output_field(isd:ied, jsd:jed, nlay+1, 6) of dp, pt, u, v, ts, nflx;
      ! the device accumulators restart from zero
      real(rk), intent(inout) :: output_field(:, :, :, :)
      call device_check(atm_output_mean(device_ctx(), output_field), 'write_output')
   end subroutine device_output_mean

end module atm_device
