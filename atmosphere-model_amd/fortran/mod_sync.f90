! mod_sync.f90 -- drop-in for the reference's mod_sync (src/mod_sync.f90) on
! one image: the halo_exchange generic (14-17, 41-107) with its four
! specifics and allocate_sync_buffers.  On one image every neighbour is the
! image itself, so the exchange is the periodic fill
! halo(i, j) = q(wrap(i), wrap(j)) of the 7-wide halo, edges and corners.
!
! These specifics act on the host arrays the caller passes (initial fields
! from mod_input.f90:249-250, ts from mod_physics.f90:189).  The halo fills
! of the device-resident state are part of the drop-in mod_sw_dyn /
! mod_radiation procedures and run on the GPU.
module mod_sync

   use mod_kinds, only: ik, rk
   use mod_config, only: config => main_config
   use mod_tiles, only: is, ie, js, je, isd, ied, jsd, jed

   implicit none

   private
   public :: halo_exchange, allocate_sync_buffers

   interface halo_exchange
      module procedure :: halo_exchange_1arg_2d, halo_exchange_1arg_levels, &
         halo_exchange_4arg_layers, halo_exchange_5arg_layers
   end interface

contains

   subroutine allocate_sync_buffers()
      ! no buffers: the fill copies within the array
   end subroutine allocate_sync_buffers

   subroutine fill_2d(q)
      real(rk), intent(inout) :: q(is:ie, js:je)
      integer(ik) :: i, j, nx, ny
      nx = ied - isd + 1
      ny = jed - jsd + 1
      do j = jsd, jed
         do i = is, isd - 1
            q(i, j) = q(i + nx, j)
         end do
         do i = ied + 1, ie
            q(i, j) = q(i - nx, j)
         end do
      end do
      do j = js, jsd - 1
         q(:, j) = q(:, j + ny)
      end do
      do j = jed + 1, je
         q(:, j) = q(:, j - ny)
      end do
   end subroutine fill_2d

   subroutine fill_3d(q, nk)
      integer(ik), intent(in) :: nk
      real(rk), intent(inout) :: q(is:ie, js:je, nk)
      integer(ik) :: k
      do k = 1, nk
         call fill_2d(q(:, :, k))
      end do
   end subroutine fill_3d

   subroutine halo_exchange_1arg_2d(a)
      real(rk), dimension(is:ie, js:je), intent(inout) :: a
      call fill_2d(a)
   end subroutine halo_exchange_1arg_2d

   subroutine halo_exchange_1arg_levels(a)
      real(rk), dimension(is:ie, js:je, 1:config%nlay + 1), intent(inout) :: a
      call fill_3d(a, config%nlay + 1)
   end subroutine halo_exchange_1arg_levels

   subroutine halo_exchange_4arg_layers(a, b, c, d)
      real(rk), dimension(is:ie, js:je, 1:config%nlay), intent(inout) :: a, b, c, d
      call fill_3d(a, config%nlay)
      call fill_3d(b, config%nlay)
      call fill_3d(c, config%nlay)
      call fill_3d(d, config%nlay)
   end subroutine halo_exchange_4arg_layers

   subroutine halo_exchange_5arg_layers(a, b, c, d, e)
      real(rk), dimension(is:ie, js:je, 1:config%nlay), intent(inout) :: a, b, c, d, e
      call fill_3d(a, config%nlay)
      call fill_3d(b, config%nlay)
      call fill_3d(c, config%nlay)
      call fill_3d(d, config%nlay)
      call fill_3d(e, config%nlay)
   end subroutine halo_exchange_5arg_layers

end module mod_sync
