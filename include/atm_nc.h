/*
 * atm_nc.h -- C ABI of the model's netCDF classic-format I/O (host only).
 *
 * The reference reads its initial state and writes its time means and
 * restart through the netCDF-Fortran library (src/mod_netcdf.f90:1-632,
 * called from src/mod_input.f90:21-252 and src/mod_output.f90:37-261).
 * This library reads and writes the netCDF classic formats directly:
 * CDF-1 (classic), CDF-2 (64-bit offset) and CDF-5 (64-bit data), so a
 * host without netCDF can load initial files and write output.  The
 * Fortran module `netcdf` of atmosphere-model_amd/fortran/netcdf.f90 binds
 * these functions under the nf90_* names src/mod_netcdf.f90:7-11 imports.
 * netCDF-4 (HDF5) files are not read: open reports ATM_NC_ENOTNC.
 *
 * Conventions follow the netCDF C interface: ids are 0-based, dimension
 * order is C order (the last dimension varies fastest on disk), a record
 * (unlimited) variable has the record dimension first, and the mode flags,
 * type codes and status values carry netCDF's numbers, so a caller can map
 * them one to one.  Transfers convert between the file's external type and
 * double; a value outside the external type's range is written converted
 * and reported as ATM_NC_ERANGE.
 *
 * Files are not thread-safe: one file id is used from one thread at a time.
 * Separate ids may be used from separate threads.
 */
#ifndef ATM_NC_H
#define ATM_NC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (netCDF's values) */
enum {
    ATM_NC_NOERR = 0,
    ATM_NC_EBADID = -33,        /* not a valid id */
    ATM_NC_EEXIST = -35,        /* file exists and NOCLOBBER was given */
    ATM_NC_EINVAL = -36,        /* invalid argument */
    ATM_NC_EPERM = -37,         /* write to a read-only file */
    ATM_NC_ENOTINDEFINE = -38,  /* operation needs define mode */
    ATM_NC_EINDEFINE = -39,     /* operation not allowed in define mode */
    ATM_NC_EINVALCOORDS = -40,  /* start index out of range */
    ATM_NC_ENAMEINUSE = -42,    /* name already used */
    ATM_NC_ENOTATT = -43,       /* attribute not found */
    ATM_NC_EBADTYPE = -45,      /* unknown or unsupported type */
    ATM_NC_EBADDIM = -46,       /* invalid dimension id or name */
    ATM_NC_EUNLIMPOS = -47,     /* record dimension not first */
    ATM_NC_ENOTVAR = -49,       /* variable not found */
    ATM_NC_ENOTNC = -51,        /* not a netCDF classic-format file */
    ATM_NC_EUNLIMIT = -54,      /* second record dimension */
    ATM_NC_ECHAR = -56,         /* numeric transfer on a char variable */
    ATM_NC_EEDGE = -57,         /* start + count out of range */
    ATM_NC_EBADNAME = -59,      /* invalid name */
    ATM_NC_ERANGE = -60,        /* value out of the external type's range */
    ATM_NC_ENOMEM = -61,        /* out of memory */
    ATM_NC_EVARSIZE = -62,      /* variable or offset too large for the format */
    ATM_NC_ETRUNC = -64,        /* file shorter than its header says */
    ATM_NC_EIO = -68            /* read or write of the file failed */
};

/* external types (netCDF's codes) */
enum {
    ATM_NC_BYTE = 1, ATM_NC_CHAR = 2, ATM_NC_SHORT = 3, ATM_NC_INT = 4,
    ATM_NC_FLOAT = 5, ATM_NC_DOUBLE = 6,
    ATM_NC_UBYTE = 7, ATM_NC_USHORT = 8, ATM_NC_UINT = 9, ATM_NC_INT64 = 10, ATM_NC_UINT64 = 11
};

/* open and create modes (netCDF's flags) */
enum {
    ATM_NC_NOWRITE = 0x0000,
    ATM_NC_WRITE = 0x0001,
    ATM_NC_CLOBBER = 0x0000,
    ATM_NC_NOCLOBBER = 0x0004,
    ATM_NC_64BIT_DATA = 0x0020,    /* CDF-5 */
    ATM_NC_64BIT_OFFSET = 0x0200   /* CDF-2; neither flag: CDF-1 */
};

enum { ATM_NC_UNLIMITED = 0, ATM_NC_GLOBAL = -1, ATM_NC_MAX_NAME = 256 };

/* File lifetime.  open: nf90_open (mod_netcdf.f90:66); create (define
   mode): nf90_create (mod_netcdf.f90:165); enddef / redef: nf90_enddef /
   nf90_redef (mod_netcdf.f90:168, 200-214 ...); close: nf90_close
   (mod_netcdf.f90:181).  sync writes the header's record count. */
int atm_nc_open(const char *path, int mode, int *ncid);
int atm_nc_create(const char *path, int cmode, int *ncid);
int atm_nc_enddef(int ncid);
int atm_nc_redef(int ncid);
int atm_nc_sync(int ncid);
int atm_nc_close(int ncid);

/* Inquiry: nf90_inquire, nf90_inquire_dimension, nf90_inquire_variable,
   nf90_inq_dimid, nf90_inq_varid (mod_netcdf.f90:69-118, 123-145).
   Name buffers hold at least ATM_NC_MAX_NAME + 1 bytes; unlimdimid is -1
   when the file has no record dimension.  The length of the record
   dimension is the number of records.  format: 1, 2 or 5 (CDF-n). */
int atm_nc_inq(int ncid, int *ndims, int *nvars, int *ngatts, int *unlimdimid);
int atm_nc_inq_format(int ncid, int *format);
int atm_nc_inq_dim(int ncid, int dimid, char *name, size_t *len);
int atm_nc_inq_dimid(int ncid, const char *name, int *dimid);
int atm_nc_inq_var(int ncid, int varid, char *name, int *xtype, int *ndims, int *dimids, int *natts);
int atm_nc_inq_varid(int ncid, const char *name, int *varid);

/* Definitions (define mode): nf90_def_dim, nf90_def_var
   (mod_netcdf.f90:197-202, 227-232, 301-305, 407-408).  len 0 defines the
   record dimension. */
int atm_nc_def_dim(int ncid, const char *name, size_t len, int *dimid);
int atm_nc_def_var(int ncid, const char *name, int xtype, int ndims, const int *dimids, int *varid);

/* Attributes (varid ATM_NC_GLOBAL for the file's).  The reference writes
   none; files from other writers carry them (units, _FillValue ...).  A
   variable's _FillValue of its own type replaces the default fill. */
int atm_nc_inq_att(int ncid, int varid, const char *name, int *xtype, size_t *len);
int atm_nc_inq_attname(int ncid, int varid, int attnum, char *name);
int atm_nc_get_att_double(int ncid, int varid, const char *name, double *values);
int atm_nc_get_att_text(int ncid, int varid, const char *name, char *text);
int atm_nc_put_att_double(int ncid, int varid, const char *name, int xtype, size_t len, const double *values);
int atm_nc_put_att_text(int ncid, int varid, const char *name, size_t len, const char *text);

/* Data (data mode): nf90_get_var / nf90_put_var with start and count
   (mod_netcdf.f90:251-252, 324-330, 460-505, 520-631).  start and count
   have one entry per dimension in C order; vara moves a contiguous block
   (last dimension fastest), varm places element (i_0 .. i_n-1) of the
   block at values[sum i_d * imap[d]].  Writing past the last record adds
   records; records in between are filled. */
int atm_nc_get_vara_double(int ncid, int varid, const size_t *start, const size_t *count, double *values);
int atm_nc_get_varm_double(int ncid, int varid, const size_t *start, const size_t *count,
                           const ptrdiff_t *imap, double *values);
int atm_nc_put_vara_double(int ncid, int varid, const size_t *start, const size_t *count, const double *values);
int atm_nc_put_varm_double(int ncid, int varid, const size_t *start, const size_t *count,
                           const ptrdiff_t *imap, const double *values);

/* Message for a status code (nf90_strerror, mod_netcdf.f90:627). */
const char *atm_nc_strerror(int status);

#ifdef __cplusplus
}
#endif

#endif /* ATM_NC_H */
