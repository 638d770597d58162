! nc_host.f90 -- test program for the module `netcdf` of
! atmosphere-model_amd/fortran/netcdf.f90 (CPU only).
!
! It drives the nf90_* calls in the sequences the reference's mod_netcdf
! makes (src/mod_netcdf.f90):
!   read  <initial.nc> <nx> <ny> <nlay> <out.bin>
!       open_netcdf's inquiry loops (55-118), then read_3d_variable /
!       read_2d_variable (550-624, start/count/map built from the dimension
!       names, last record) for dp, pt, u, v, gzs, ts, and the 1-D reads of
!       ak, bk, xc, t (324-330, 460-477); writes them to out.bin as
!       real64 in that order;
!   write <out.nc> <nx> <ny> <nlay>
!       create_netcdf (160-176), create_time_axis (189-206), advance_time
!       (208-219), create_real_axis / create_integer_axis (221-300),
!       create_variable (376-441) and write_3d_variable / write_2d_variable
!       (479-535) with the values v(i,j,k) = i + 100 j + 10000 k + 1e6 r for
!       records r = 1, 2.
program nc_host
   use netcdf
   implicit none
   character(len=256) :: mode, path, out, arg
   integer :: nx, ny, nlay

   call get_command_argument(1, mode)
   call get_command_argument(2, path)
   call get_command_argument(3, arg)
   read (arg, *) nx
   call get_command_argument(4, arg)
   read (arg, *) ny
   call get_command_argument(5, arg)
   read (arg, *) nlay
   if (trim(mode) == 'read') then
      call get_command_argument(6, out)
      call read_file()
   else
      call write_file()
   end if

contains

   subroutine check(status, where)
      integer, intent(in) :: status
      character(*), intent(in) :: where
      if (status /= nf90_noerr) then
         write (*, '(a)') 'error in '//where//': '//trim(nf90_strerror(status))
         error stop 1
      end if
   end subroutine check

   subroutine read_file()
      integer :: ncid, ndims, nvars, tdim, tlen, i, j, n, nd, varid, u
      character(len=10) :: name, tname
      character(len=10), allocatable :: dnames(:)
      integer, allocatable :: dimids(:), start(:), count(:), dmap(:), lens(:)
      real(8), allocatable :: v3(:, :, :), v2(:, :), v1(:)
      character(len=3), parameter :: names3(4) = ['dp ', 'pt ', 'u  ', 'v  ']
      character(len=3), parameter :: names2(2) = ['gzs', 'ts ']

      call check(nf90_open(trim(path), nf90_nowrite, ncid), 'open')
      call check(nf90_inquire(ncid, nDimensions=ndims, nVariables=nvars, unlimitedDimId=tdim), 'inquire')
      if (tdim == -1) error stop 2
      call check(nf90_inquire_dimension(ncid, tdim, name=tname, len=tlen), 'inquire_dimension')
      allocate (lens(ndims))
      do i = 1, ndims
         call check(nf90_inquire_dimension(ncid, i, name=name, len=lens(i)), 'inquire_dimension')
      end do
      open (newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
      allocate (v3(nx, ny, nlay))
      do n = 1, 6
         if (n <= 4) then
            name = names3(n)
         else
            name = names2(n - 4)
         end if
         call check(nf90_inq_varid(ncid, trim(name), varid), 'inq_varid')
         call check(nf90_inquire_variable(ncid, varid, ndims=nd), 'inquire_variable')
         allocate (dimids(nd), dnames(nd), start(nd), count(nd), dmap(nd))
         call check(nf90_inquire_variable(ncid, varid, dimids=dimids), 'inquire_variable dimids')
         do j = 1, nd
            call check(nf90_inquire_dimension(ncid, dimids(j), name=dnames(j)), 'inquire_dimension')
            if (dnames(j) == tname) then
               start(j) = tlen
               count(j) = 1
               dmap(j) = 0
            else
               start(j) = 1
               count(j) = lens(dimids(j))
               if (any(dnames(j) == ['x ', 'xc', 'xf'])) then
                  dmap(j) = 1
               else if (any(dnames(j) == ['y ', 'yc', 'yf'])) then
                  dmap(j) = nx
               else if (dnames(j) == 'lay') then
                  dmap(j) = nx*ny
               end if
            end if
         end do
         if (n <= 4) then
            call check(nf90_get_var(ncid, varid, v3, start=start, count=count, map=dmap), 'get_var 3d')
            write (u) v3
         else
            allocate (v2(nx, ny))
            call check(nf90_get_var(ncid, varid, v2, start=start, count=count, map=dmap), 'get_var 2d')
            write (u) v2
            deallocate (v2)
         end if
         deallocate (dimids, dnames, start, count, dmap)
      end do
      allocate (v1(nlay + 1))
      call check(nf90_inq_varid(ncid, 'ak', varid), 'inq_varid ak')
      call check(nf90_get_var(ncid, varid, v1), 'get_var ak')
      write (u) v1
      call check(nf90_inq_varid(ncid, 'bk', varid), 'inq_varid bk')
      call check(nf90_get_var(ncid, varid, v1), 'get_var bk')
      write (u) v1
      deallocate (v1)
      allocate (v1(nx))
      call check(nf90_inq_varid(ncid, 'xc', varid), 'inq_varid xc')
      call check(nf90_get_var(ncid, varid, v1), 'get_var xc')
      write (u) v1
      deallocate (v1)
      allocate (v1(tlen))
      call check(nf90_inq_varid(ncid, trim(tname), varid), 'inq_varid t')
      call check(nf90_get_var(ncid, varid, v1), 'get_var t')
      write (u) v1
      close (u)
      call check(nf90_close(ncid), 'close')
   end subroutine read_file

   subroutine write_file()
      integer :: ncid, tdim, tvar, dimid, varid, i, j, k, r
      integer :: dx, dy, dlay, dlev
      real(8), allocatable :: f3(:, :, :), f2(:, :)

      call check(nf90_create(path=trim(path), cmode=nf90_clobber, ncid=ncid), 'create')
      call check(nf90_enddef(ncid), 'enddef')
      ! time axis
      call check(nf90_redef(ncid), 'redef t')
      call check(nf90_def_dim(ncid, 't', nf90_unlimited, tdim), 'def_dim t')
      call check(nf90_def_var(ncid, 't', nf90_float, tdim, tvar), 'def_var t')
      call check(nf90_enddef(ncid), 'enddef t')
      call check(nf90_put_var(ncid, tvar, 1800.5d0, start=[1]), 'put t 1')
      ! axes, each in its own define block
      call axis_r(ncid, 'x', [(-.5d0*nx + (i + .5d0), i=0, nx - 1)], dx)
      call axis_r(ncid, 'y', [(-.5d0*ny + (i + .5d0), i=0, ny - 1)], dy)
      call axis_i(ncid, 'lay', [(i, i=1, nlay)], dlay)
      call axis_i(ncid, 'lev', [(i, i=1, nlay + 1)], dlev)
      ! variables: the record dimension goes last in Fortran order
      call check(nf90_redef(ncid), 'redef var')
      call check(nf90_def_var(ncid, 'dp', nf90_float, [dx, dy, dlay, tdim], varid), 'def_var dp')
      call check(nf90_enddef(ncid), 'enddef var')
      call check(nf90_redef(ncid), 'redef var')
      call check(nf90_def_var(ncid, 'ts', nf90_float, [dx, dy, tdim], varid), 'def_var ts')
      call check(nf90_enddef(ncid), 'enddef var')
      call check(nf90_redef(ncid), 'redef var')
      call check(nf90_def_var(ncid, 'gzs', nf90_double, [dx, dy], varid), 'def_var gzs')
      call check(nf90_enddef(ncid), 'enddef var')

      allocate (f3(nx, ny, nlay), f2(nx, ny))
      do r = 1, 2
         if (r == 2) call check(nf90_put_var(ncid, tvar, 3600.25d0, start=[2]), 'put t 2')
         do k = 1, nlay
            do j = 1, ny
               do i = 1, nx
                  f3(i, j, k) = i + 100*j + 10000*k + 1000000*r
               end do
            end do
         end do
         f2 = f3(:, :, 1)
         call check(nf90_inq_varid(ncid, 'dp', varid), 'inq_varid dp')
         call check(nf90_put_var(ncid, varid, f3, start=[1, 1, 1, r], count=[nx, ny, nlay, 1]), 'put dp')
         call check(nf90_inq_varid(ncid, 'ts', varid), 'inq_varid ts')
         call check(nf90_put_var(ncid, varid, f2, start=[1, 1, r], count=[nx, ny, 1]), 'put ts')
      end do
      call check(nf90_inq_varid(ncid, 'gzs', varid), 'inq_varid gzs')
      call check(nf90_put_var(ncid, varid, f2 + 0.125d0, start=[1, 1], count=[nx, ny]), 'put gzs')
      call check(nf90_close(ncid), 'close')
   end subroutine write_file

   subroutine axis_r(ncid_, name, points, dimid)
      integer, intent(in) :: ncid_
      character(*), intent(in) :: name
      real(8), intent(in) :: points(:)
      integer, intent(out) :: dimid
      integer :: varid
      call check(nf90_redef(ncid_), 'redef axis')
      call check(nf90_def_dim(ncid_, name, size(points), dimid), 'def_dim axis')
      call check(nf90_def_var(ncid_, name, nf90_float, dimid, varid), 'def_var axis')
      call check(nf90_enddef(ncid_), 'enddef axis')
      call check(nf90_put_var(ncid_, varid, points), 'put axis')
   end subroutine axis_r

   subroutine axis_i(ncid_, name, points, dimid)
      integer, intent(in) :: ncid_
      character(*), intent(in) :: name
      integer, intent(in) :: points(:)
      integer, intent(out) :: dimid
      integer :: varid
      call check(nf90_redef(ncid_), 'redef axis')
      call check(nf90_def_dim(ncid_, name, size(points), dimid), 'def_dim axis')
      call check(nf90_def_var(ncid_, name, nf90_int, dimid, varid), 'def_var axis')
      call check(nf90_enddef(ncid_), 'enddef axis')
      call check(nf90_put_var(ncid_, varid, points), 'put axis')
   end subroutine axis_i

end program nc_host
