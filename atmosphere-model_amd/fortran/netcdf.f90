! netcdf.f90 -- module `netcdf`: the part of the netCDF-Fortran 90 interface
! that the reference's mod_netcdf imports (src/mod_netcdf.f90:7-11), over
! the classic-format library libatmos_io.so (include/atm_nc.h).  With it the
! reference's mod_netcdf, mod_input and mod_fields compile and run unchanged
! on a host without netCDF; mod_output compiles unchanged too where the
! compiler takes its coarrays (gfortran -fcoarray=single).
!
! Conventions are netCDF-Fortran's: ids are 1-based, dimension order is
! Fortran order (the first dimension varies fastest; the record dimension
! is the last), start/count/map follow that order, unlimitedDimId is -1 when
! there is no record dimension, and nf90_global (0) names the file's own
! attributes.  Values are transferred in double (real64); integer arrays go
! through it exactly.  Strides other than 1 are not supported.
module netcdf

   use iso_c_binding, only: c_int, c_char, c_double, c_size_t, c_ptrdiff_t, c_ptr, c_null_char, &
                            c_f_pointer, c_associated

   implicit none

   private
   public :: nf90_open, nf90_create, nf90_enddef, nf90_redef, nf90_sync, nf90_close, &
             nf90_inquire, nf90_inquire_dimension, nf90_inquire_variable, nf90_inq_dimid, &
             nf90_inq_varid, nf90_def_dim, nf90_def_var, nf90_put_var, nf90_get_var, &
             nf90_put_att, nf90_get_att, nf90_inquire_attribute, nf90_strerror

   integer, parameter, public :: nf90_noerr = 0
   integer, parameter, public :: nf90_nowrite = 0, nf90_write = 1
   integer, parameter, public :: nf90_clobber = 0, nf90_noclobber = 4
   integer, parameter, public :: nf90_64bit_offset = 512, nf90_64bit_data = 32
   integer, parameter, public :: nf90_unlimited = 0, nf90_global = 0
   integer, parameter, public :: nf90_byte = 1, nf90_char = 2, nf90_short = 3, nf90_int = 4, &
                                 nf90_float = 5, nf90_double = 6
   integer, parameter, public :: nf90_max_name = 256
   integer, parameter :: einval = -36, ebadtype = -45, echar = -56

   interface
      integer(c_int) function c_open(path, mode, ncid) bind(c, name='atm_nc_open')
         import :: c_int, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: mode
         integer(c_int), intent(out) :: ncid
      end function
      integer(c_int) function c_create(path, cmode, ncid) bind(c, name='atm_nc_create')
         import :: c_int, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: cmode
         integer(c_int), intent(out) :: ncid
      end function
      integer(c_int) function c_enddef(ncid) bind(c, name='atm_nc_enddef')
         import :: c_int
         integer(c_int), value :: ncid
      end function
      integer(c_int) function c_redef(ncid) bind(c, name='atm_nc_redef')
         import :: c_int
         integer(c_int), value :: ncid
      end function
      integer(c_int) function c_sync(ncid) bind(c, name='atm_nc_sync')
         import :: c_int
         integer(c_int), value :: ncid
      end function
      integer(c_int) function c_close(ncid) bind(c, name='atm_nc_close')
         import :: c_int
         integer(c_int), value :: ncid
      end function
      integer(c_int) function c_inq(ncid, ndims, nvars, ngatts, unlimdimid) bind(c, name='atm_nc_inq')
         import :: c_int
         integer(c_int), value :: ncid
         integer(c_int), intent(out) :: ndims, nvars, ngatts, unlimdimid
      end function
      integer(c_int) function c_inq_format(ncid, fmt) bind(c, name='atm_nc_inq_format')
         import :: c_int
         integer(c_int), value :: ncid
         integer(c_int), intent(out) :: fmt
      end function
      integer(c_int) function c_inq_dim(ncid, dimid, name, len) bind(c, name='atm_nc_inq_dim')
         import :: c_int, c_char, c_size_t
         integer(c_int), value :: ncid, dimid
         character(kind=c_char), intent(out) :: name(*)
         integer(c_size_t), intent(out) :: len
      end function
      integer(c_int) function c_inq_dimid(ncid, name, dimid) bind(c, name='atm_nc_inq_dimid')
         import :: c_int, c_char
         integer(c_int), value :: ncid
         character(kind=c_char), intent(in) :: name(*)
         integer(c_int), intent(out) :: dimid
      end function
      integer(c_int) function c_inq_var(ncid, varid, name, xtype, ndims, dimids, natts) &
         bind(c, name='atm_nc_inq_var')
         import :: c_int, c_char
         integer(c_int), value :: ncid, varid
         character(kind=c_char), intent(out) :: name(*)
         integer(c_int), intent(out) :: xtype, ndims, dimids(*), natts
      end function
      integer(c_int) function c_inq_varid(ncid, name, varid) bind(c, name='atm_nc_inq_varid')
         import :: c_int, c_char
         integer(c_int), value :: ncid
         character(kind=c_char), intent(in) :: name(*)
         integer(c_int), intent(out) :: varid
      end function
      integer(c_int) function c_def_dim(ncid, name, len, dimid) bind(c, name='atm_nc_def_dim')
         import :: c_int, c_char, c_size_t
         integer(c_int), value :: ncid
         character(kind=c_char), intent(in) :: name(*)
         integer(c_size_t), value :: len
         integer(c_int), intent(out) :: dimid
      end function
      integer(c_int) function c_def_var(ncid, name, xtype, ndims, dimids, varid) bind(c, name='atm_nc_def_var')
         import :: c_int, c_char
         integer(c_int), value :: ncid, xtype, ndims
         character(kind=c_char), intent(in) :: name(*)
         integer(c_int), intent(in) :: dimids(*)
         integer(c_int), intent(out) :: varid
      end function
      integer(c_int) function c_inq_att(ncid, varid, name, xtype, len) bind(c, name='atm_nc_inq_att')
         import :: c_int, c_char, c_size_t
         integer(c_int), value :: ncid, varid
         character(kind=c_char), intent(in) :: name(*)
         integer(c_int), intent(out) :: xtype
         integer(c_size_t), intent(out) :: len
      end function
      integer(c_int) function c_get_att_double(ncid, varid, name, values) bind(c, name='atm_nc_get_att_double')
         import :: c_int, c_char, c_double
         integer(c_int), value :: ncid, varid
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(out) :: values(*)
      end function
      integer(c_int) function c_get_att_text(ncid, varid, name, text) bind(c, name='atm_nc_get_att_text')
         import :: c_int, c_char
         integer(c_int), value :: ncid, varid
         character(kind=c_char), intent(in) :: name(*)
         character(kind=c_char), intent(out) :: text(*)
      end function
      integer(c_int) function c_put_att_double(ncid, varid, name, xtype, len, values) &
         bind(c, name='atm_nc_put_att_double')
         import :: c_int, c_char, c_double, c_size_t
         integer(c_int), value :: ncid, varid, xtype
         character(kind=c_char), intent(in) :: name(*)
         integer(c_size_t), value :: len
         real(c_double), intent(in) :: values(*)
      end function
      integer(c_int) function c_put_att_text(ncid, varid, name, len, text) bind(c, name='atm_nc_put_att_text')
         import :: c_int, c_char, c_size_t
         integer(c_int), value :: ncid, varid
         character(kind=c_char), intent(in) :: name(*)
         integer(c_size_t), value :: len
         character(kind=c_char), intent(in) :: text(*)
      end function
      integer(c_int) function c_get_varm(ncid, varid, start, count, imap, values) &
         bind(c, name='atm_nc_get_varm_double')
         import :: c_int, c_size_t, c_ptrdiff_t, c_double
         integer(c_int), value :: ncid, varid
         integer(c_size_t), intent(in) :: start(*), count(*)
         integer(c_ptrdiff_t), intent(in) :: imap(*)
         real(c_double), intent(out) :: values(*)
      end function
      integer(c_int) function c_put_varm(ncid, varid, start, count, imap, values) &
         bind(c, name='atm_nc_put_varm_double')
         import :: c_int, c_size_t, c_ptrdiff_t, c_double
         integer(c_int), value :: ncid, varid
         integer(c_size_t), intent(in) :: start(*), count(*)
         integer(c_ptrdiff_t), intent(in) :: imap(*)
         real(c_double), intent(in) :: values(*)
      end function
      type(c_ptr) function c_strerror(status) bind(c, name='atm_nc_strerror')
         import :: c_ptr, c_int
         integer(c_int), value :: status
      end function
   end interface

   interface nf90_def_var
      module procedure def_var_nodims, def_var_onedim, def_var_dims
   end interface

   interface nf90_put_var
      module procedure put_r8_0d, put_r8_1d, put_r8_2d, put_r8_3d, put_r8_4d, &
         put_r4_0d, put_r4_1d, put_r4_2d, put_r4_3d, put_i4_0d, put_i4_1d, put_i4_2d
   end interface

   interface nf90_get_var
      module procedure get_r8_0d, get_r8_1d, get_r8_2d, get_r8_3d, get_r8_4d, &
         get_r4_0d, get_r4_1d, get_r4_2d, get_r4_3d, get_i4_0d, get_i4_1d, get_i4_2d
   end interface

   interface nf90_put_att
      module procedure put_att_text, put_att_r8, put_att_r8_1d, put_att_i4
   end interface

   interface nf90_get_att
      module procedure get_att_text, get_att_r8, get_att_r8_1d, get_att_i4
   end interface

contains

   ! ---- helpers -------------------------------------------------------------

   pure function cstr(s) result(out)
      character(*), intent(in) :: s
      character(kind=c_char, len=len_trim(s) + 1) :: out
      out = trim(s)//c_null_char
   end function cstr

   subroutine fstr(buf, out)
      ! C string of buf into the blank-padded out (truncated to len(out))
      character(kind=c_char), intent(in) :: buf(:)
      character(*), intent(out) :: out
      integer :: i
      out = ''
      do i = 1, min(size(buf), len(out))
         if (buf(i) == c_null_char) exit
         out(i:i) = buf(i)
      end do
   end subroutine fstr

   integer function var_ndims(ncid, varid, nd)
      integer, intent(in) :: ncid, varid
      integer, intent(out) :: nd
      character(kind=c_char) :: name(nf90_max_name + 1)
      integer(c_int) :: xt, ndc, na, dimids(1024)
      var_ndims = c_inq_var(ncid - 1, varid - 1, name, xt, ndc, dimids, na)
      nd = ndc
   end function var_ndims

   integer function transfer(ncid, varid, buf, shp, start, count, map, write)
      ! start/count/map in Fortran order, 1-based start; buf holds the
      ! values in array element order
      integer, intent(in) :: ncid, varid
      real(c_double), intent(inout) :: buf(:)
      integer, intent(in) :: shp(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      logical, intent(in) :: write
      integer :: nd, d, r
      integer, allocatable :: fs(:), fc(:), fm(:)
      integer(c_size_t), allocatable :: cs(:), cc(:)
      integer(c_ptrdiff_t), allocatable :: cm(:)

      transfer = var_ndims(ncid, varid, nd)
      if (transfer /= nf90_noerr) return
      allocate (fs(max(nd, 1)), fc(max(nd, 1)), fm(max(nd, 1)))
      fs = 1
      fc = 1
      r = min(size(shp), nd)
      if (r > 0) fc(1:r) = shp(1:r)
      if (present(start)) then
         if (size(start) < nd) then
            transfer = einval
            return
         end if
         fs(1:nd) = start(1:nd)
      end if
      if (present(count)) then
         if (size(count) < nd) then
            transfer = einval
            return
         end if
         fc(1:nd) = count(1:nd)
      end if
      if (present(map)) then
         if (size(map) < nd) then
            transfer = einval
            return
         end if
         fm(1:nd) = map(1:nd)
      else
         fm(1) = 1
         do d = 2, nd
            fm(d) = fm(d - 1)*fc(d - 1)
         end do
      end if
      if (any(fs(1:nd) < 1)) then
         transfer = -40  ! index exceeds dimension bound
         return
      end if
      allocate (cs(max(nd, 1)), cc(max(nd, 1)), cm(max(nd, 1)))
      do d = 1, nd  ! C order is the reverse
         cs(nd + 1 - d) = int(fs(d) - 1, c_size_t)
         cc(nd + 1 - d) = int(fc(d), c_size_t)
         cm(nd + 1 - d) = int(fm(d), c_ptrdiff_t)
      end do
      if (write) then
         transfer = c_put_varm(ncid - 1, varid - 1, cs, cc, cm, buf)
      else
         transfer = c_get_varm(ncid - 1, varid - 1, cs, cc, cm, buf)
      end if
   end function transfer

   ! ---- files ---------------------------------------------------------------

   integer function nf90_open(path, mode, ncid)
      character(*), intent(in) :: path
      integer, intent(in) :: mode
      integer, intent(out) :: ncid
      integer(c_int) :: id
      nf90_open = c_open(cstr(path), mode, id)
      ncid = id + 1
   end function nf90_open

   integer function nf90_create(path, cmode, ncid)
      character(*), intent(in) :: path
      integer, intent(in) :: cmode
      integer, intent(out) :: ncid
      integer(c_int) :: id
      nf90_create = c_create(cstr(path), cmode, id)
      ncid = id + 1
   end function nf90_create

   integer function nf90_enddef(ncid)
      integer, intent(in) :: ncid
      nf90_enddef = c_enddef(ncid - 1)
   end function nf90_enddef

   integer function nf90_redef(ncid)
      integer, intent(in) :: ncid
      nf90_redef = c_redef(ncid - 1)
   end function nf90_redef

   integer function nf90_sync(ncid)
      integer, intent(in) :: ncid
      nf90_sync = c_sync(ncid - 1)
   end function nf90_sync

   integer function nf90_close(ncid)
      integer, intent(in) :: ncid
      nf90_close = c_close(ncid - 1)
   end function nf90_close

   function nf90_strerror(ncerr) result(msg)
      integer, intent(in) :: ncerr
      character(len=128) :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      msg = ''
      p = c_strerror(ncerr)
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [128])
      call fstr(s, msg)
   end function nf90_strerror

   ! ---- inquiry -------------------------------------------------------------

   integer function nf90_inquire(ncid, nDimensions, nVariables, nAttributes, unlimitedDimId, formatNum)
      integer, intent(in) :: ncid
      integer, intent(out), optional :: nDimensions, nVariables, nAttributes, unlimitedDimId, formatNum
      integer(c_int) :: nd, nv, na, ul, fm
      nf90_inquire = c_inq(ncid - 1, nd, nv, na, ul)
      if (nf90_inquire /= nf90_noerr) return
      if (present(nDimensions)) nDimensions = nd
      if (present(nVariables)) nVariables = nv
      if (present(nAttributes)) nAttributes = na
      if (present(unlimitedDimId)) then
         unlimitedDimId = -1
         if (ul >= 0) unlimitedDimId = ul + 1
      end if
      if (present(formatNum)) then
         nf90_inquire = c_inq_format(ncid - 1, fm)
         formatNum = fm
      end if
   end function nf90_inquire

   integer function nf90_inquire_dimension(ncid, dimid, name, len)
      integer, intent(in) :: ncid, dimid
      character(*), intent(out), optional :: name
      integer, intent(out), optional :: len
      character(kind=c_char) :: buf(nf90_max_name + 1)
      integer(c_size_t) :: n
      nf90_inquire_dimension = c_inq_dim(ncid - 1, dimid - 1, buf, n)
      if (nf90_inquire_dimension /= nf90_noerr) return
      if (present(name)) call fstr(buf, name)
      if (present(len)) len = int(n)
   end function nf90_inquire_dimension

   integer function nf90_inquire_variable(ncid, varid, name, xtype, ndims, dimids, nAtts)
      integer, intent(in) :: ncid, varid
      character(*), intent(out), optional :: name
      integer, intent(out), optional :: xtype, ndims, nAtts
      integer, intent(out), optional :: dimids(:)
      character(kind=c_char) :: buf(nf90_max_name + 1)
      integer(c_int) :: xt, nd, na, ids(1024)
      integer :: d
      nf90_inquire_variable = c_inq_var(ncid - 1, varid - 1, buf, xt, nd, ids, na)
      if (nf90_inquire_variable /= nf90_noerr) return
      if (present(name)) call fstr(buf, name)
      if (present(xtype)) xtype = xt
      if (present(ndims)) ndims = nd
      if (present(nAtts)) nAtts = na
      if (present(dimids)) then
         if (size(dimids) < nd) then
            nf90_inquire_variable = einval
            return
         end if
         do d = 1, nd
            dimids(d) = ids(nd + 1 - d) + 1
         end do
      end if
   end function nf90_inquire_variable

   integer function nf90_inq_dimid(ncid, name, dimid)
      integer, intent(in) :: ncid
      character(*), intent(in) :: name
      integer, intent(out) :: dimid
      integer(c_int) :: id
      nf90_inq_dimid = c_inq_dimid(ncid - 1, cstr(name), id)
      dimid = id + 1
   end function nf90_inq_dimid

   integer function nf90_inq_varid(ncid, name, varid)
      integer, intent(in) :: ncid
      character(*), intent(in) :: name
      integer, intent(out) :: varid
      integer(c_int) :: id
      nf90_inq_varid = c_inq_varid(ncid - 1, cstr(name), id)
      varid = id + 1
   end function nf90_inq_varid

   ! ---- definitions ---------------------------------------------------------

   integer function nf90_def_dim(ncid, name, len, dimid)
      integer, intent(in) :: ncid, len
      character(*), intent(in) :: name
      integer, intent(out) :: dimid
      integer(c_int) :: id
      if (len < 0) then
         nf90_def_dim = einval
         return
      end if
      nf90_def_dim = c_def_dim(ncid - 1, cstr(name), int(len, c_size_t), id)
      dimid = id + 1
   end function nf90_def_dim

   integer function def_var_dims(ncid, name, xtype, dimids, varid)
      integer, intent(in) :: ncid, xtype
      character(*), intent(in) :: name
      integer, intent(in) :: dimids(:)
      integer, intent(out) :: varid
      integer(c_int) :: ids(max(size(dimids), 1)), id
      integer :: d, nd
      nd = size(dimids)
      do d = 1, nd
         ids(nd + 1 - d) = dimids(d) - 1
      end do
      def_var_dims = c_def_var(ncid - 1, cstr(name), xtype, nd, ids, id)
      varid = id + 1
   end function def_var_dims

   integer function def_var_onedim(ncid, name, xtype, dimids, varid)
      integer, intent(in) :: ncid, xtype, dimids
      character(*), intent(in) :: name
      integer, intent(out) :: varid
      def_var_onedim = def_var_dims(ncid, name, xtype, [dimids], varid)
   end function def_var_onedim

   integer function def_var_nodims(ncid, name, xtype, varid)
      integer, intent(in) :: ncid, xtype
      character(*), intent(in) :: name
      integer, intent(out) :: varid
      integer :: none(0)
      def_var_nodims = def_var_dims(ncid, name, xtype, none, varid)
   end function def_var_nodims

   ! ---- attributes ----------------------------------------------------------

   integer function nf90_inquire_attribute(ncid, varid, name, xtype, len)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      integer, intent(out), optional :: xtype, len
      integer(c_int) :: xt
      integer(c_size_t) :: n
      nf90_inquire_attribute = c_inq_att(ncid - 1, varid - 1, cstr(name), xt, n)
      if (present(xtype)) xtype = xt
      if (present(len)) len = int(n)
   end function nf90_inquire_attribute

   integer function put_att_text(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name, values
      put_att_text = c_put_att_text(ncid - 1, varid - 1, cstr(name), int(len(values), c_size_t), values)
   end function put_att_text

   integer function put_att_r8_1d(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      real(c_double), intent(in) :: values(:)
      put_att_r8_1d = c_put_att_double(ncid - 1, varid - 1, cstr(name), nf90_double, &
                                       int(size(values), c_size_t), values)
   end function put_att_r8_1d

   integer function put_att_r8(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      real(c_double), intent(in) :: values
      put_att_r8 = put_att_r8_1d(ncid, varid, name, [values])
   end function put_att_r8

   integer function put_att_i4(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      integer, intent(in) :: values
      put_att_i4 = c_put_att_double(ncid - 1, varid - 1, cstr(name), nf90_int, 1_c_size_t, [real(values, c_double)])
   end function put_att_i4

   integer function get_att_text(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      character(*), intent(out) :: values
      integer(c_int) :: xt
      integer(c_size_t) :: n
      character(kind=c_char), allocatable :: buf(:)
      integer :: i
      values = ''
      get_att_text = c_inq_att(ncid - 1, varid - 1, cstr(name), xt, n)
      if (get_att_text /= nf90_noerr) return
      if (xt /= nf90_char) then
         get_att_text = echar
         return
      end if
      allocate (buf(max(n, 1_c_size_t)))
      get_att_text = c_get_att_text(ncid - 1, varid - 1, cstr(name), buf)
      do i = 1, min(int(n), len(values))
         values(i:i) = buf(i)
      end do
   end function get_att_text

   integer function get_att_r8_1d(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      real(c_double), intent(out) :: values(:)
      integer(c_int) :: xt
      integer(c_size_t) :: n
      real(c_double), allocatable :: buf(:)
      get_att_r8_1d = c_inq_att(ncid - 1, varid - 1, cstr(name), xt, n)
      if (get_att_r8_1d /= nf90_noerr) return
      if (n > size(values)) then
         get_att_r8_1d = einval
         return
      end if
      allocate (buf(max(n, 1_c_size_t)))
      get_att_r8_1d = c_get_att_double(ncid - 1, varid - 1, cstr(name), buf)
      values(1:n) = buf(1:n)
   end function get_att_r8_1d

   integer function get_att_r8(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      real(c_double), intent(out) :: values
      real(c_double) :: v(1)
      get_att_r8 = get_att_r8_1d(ncid, varid, name, v)
      values = v(1)
   end function get_att_r8

   integer function get_att_i4(ncid, varid, name, values)
      integer, intent(in) :: ncid, varid
      character(*), intent(in) :: name
      integer, intent(out) :: values
      real(c_double) :: v(1)
      get_att_i4 = get_att_r8_1d(ncid, varid, name, v)
      values = int(v(1))
   end function get_att_i4

   ! ---- values: put ---------------------------------------------------------

   integer function put_r8_0d(ncid, varid, values, start)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(in) :: values
      integer, intent(in), optional :: start(:)
      real(c_double) :: b(1)
      integer :: none(0)
      b(1) = values
      put_r8_0d = transfer(ncid, varid, b, none, start=start, write=.true.)
   end function put_r8_0d

   integer function put_r8_1d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(in) :: values(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double) :: b(size(values))
      b = values
      put_r8_1d = transfer(ncid, varid, b, shape(values), start, count, map, .true.)
   end function put_r8_1d

   integer function put_r8_2d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(in) :: values(:, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:)
      b = reshape(values, [size(values)])
      put_r8_2d = transfer(ncid, varid, b, shape(values), start, count, map, .true.)
   end function put_r8_2d

   integer function put_r8_3d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(in) :: values(:, :, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:)
      b = reshape(values, [size(values)])
      put_r8_3d = transfer(ncid, varid, b, shape(values), start, count, map, .true.)
   end function put_r8_3d

   integer function put_r8_4d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(in) :: values(:, :, :, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:)
      b = reshape(values, [size(values)])
      put_r8_4d = transfer(ncid, varid, b, shape(values), start, count, map, .true.)
   end function put_r8_4d

   integer function put_r4_0d(ncid, varid, values, start)
      integer, intent(in) :: ncid, varid
      real, intent(in) :: values
      integer, intent(in), optional :: start(:)
      put_r4_0d = put_r8_0d(ncid, varid, real(values, c_double), start)
   end function put_r4_0d

   integer function put_r4_1d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real, intent(in) :: values(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      put_r4_1d = put_r8_1d(ncid, varid, real(values, c_double), start, count, map)
   end function put_r4_1d

   integer function put_r4_2d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real, intent(in) :: values(:, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      put_r4_2d = put_r8_2d(ncid, varid, real(values, c_double), start, count, map)
   end function put_r4_2d

   integer function put_r4_3d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real, intent(in) :: values(:, :, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      put_r4_3d = put_r8_3d(ncid, varid, real(values, c_double), start, count, map)
   end function put_r4_3d

   integer function put_i4_0d(ncid, varid, values, start)
      integer, intent(in) :: ncid, varid
      integer, intent(in) :: values
      integer, intent(in), optional :: start(:)
      put_i4_0d = put_r8_0d(ncid, varid, real(values, c_double), start)
   end function put_i4_0d

   integer function put_i4_1d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      integer, intent(in) :: values(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      put_i4_1d = put_r8_1d(ncid, varid, real(values, c_double), start, count, map)
   end function put_i4_1d

   integer function put_i4_2d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      integer, intent(in) :: values(:, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      put_i4_2d = put_r8_2d(ncid, varid, real(values, c_double), start, count, map)
   end function put_i4_2d

   ! ---- values: get ---------------------------------------------------------

   integer function get_r8_0d(ncid, varid, values, start)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(out) :: values
      integer, intent(in), optional :: start(:)
      real(c_double) :: b(1)
      integer :: none(0)
      get_r8_0d = transfer(ncid, varid, b, none, start=start, write=.false.)
      values = b(1)
   end function get_r8_0d

   integer function get_r8_1d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(inout) :: values(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double) :: b(size(values))
      b = values
      get_r8_1d = transfer(ncid, varid, b, shape(values), start, count, map, .false.)
      values = b
   end function get_r8_1d

   integer function get_r8_2d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(inout) :: values(:, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:)
      b = reshape(values, [size(values)])
      get_r8_2d = transfer(ncid, varid, b, shape(values), start, count, map, .false.)
      values = reshape(b, shape(values))
   end function get_r8_2d

   integer function get_r8_3d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(inout) :: values(:, :, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:)
      b = reshape(values, [size(values)])
      get_r8_3d = transfer(ncid, varid, b, shape(values), start, count, map, .false.)
      values = reshape(b, shape(values))
   end function get_r8_3d

   integer function get_r8_4d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real(c_double), intent(inout) :: values(:, :, :, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:)
      b = reshape(values, [size(values)])
      get_r8_4d = transfer(ncid, varid, b, shape(values), start, count, map, .false.)
      values = reshape(b, shape(values))
   end function get_r8_4d

   integer function get_r4_0d(ncid, varid, values, start)
      integer, intent(in) :: ncid, varid
      real, intent(out) :: values
      integer, intent(in), optional :: start(:)
      real(c_double) :: b
      get_r4_0d = get_r8_0d(ncid, varid, b, start)
      values = real(b)
   end function get_r4_0d

   integer function get_r4_1d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real, intent(inout) :: values(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double) :: b(size(values))
      b = values
      get_r4_1d = get_r8_1d(ncid, varid, b, start, count, map)
      values = real(b)
   end function get_r4_1d

   integer function get_r4_2d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real, intent(inout) :: values(:, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:, :)
      b = values
      get_r4_2d = get_r8_2d(ncid, varid, b, start, count, map)
      values = real(b)
   end function get_r4_2d

   integer function get_r4_3d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      real, intent(inout) :: values(:, :, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:, :, :)
      b = values
      get_r4_3d = get_r8_3d(ncid, varid, b, start, count, map)
      values = real(b)
   end function get_r4_3d

   integer function get_i4_0d(ncid, varid, values, start)
      integer, intent(in) :: ncid, varid
      integer, intent(out) :: values
      integer, intent(in), optional :: start(:)
      real(c_double) :: b
      get_i4_0d = get_r8_0d(ncid, varid, b, start)
      values = int(b)
   end function get_i4_0d

   integer function get_i4_1d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      integer, intent(inout) :: values(:)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double) :: b(size(values))
      b = values
      get_i4_1d = get_r8_1d(ncid, varid, b, start, count, map)
      values = int(b)
   end function get_i4_1d

   integer function get_i4_2d(ncid, varid, values, start, count, map)
      integer, intent(in) :: ncid, varid
      integer, intent(inout) :: values(:, :)
      integer, intent(in), optional :: start(:), count(:), map(:)
      real(c_double), allocatable :: b(:, :)
      b = values
      get_i4_2d = get_r8_2d(ncid, varid, b, start, count, map)
      values = int(b)
   end function get_i4_2d

end module netcdf
