/*
 * nc3 -- NetCDF-3 classic files (CDF-1 read/write, CDF-2 read) without the
 * NetCDF library, for the Cliffs driver on machines that have none.
 *
 * The C API below uses C conventions (0-based ids and indices, C dimension
 * order: slowest first).  nc3.c also exports the Fortran-77 entry points the
 * Cliffs driver calls (nf_open, nf_get_var_double, ...), with Fortran
 * conventions: 1-based ids and indices, dimension order reversed, blank-
 * padded strings.  Error codes and type codes are the NetCDF ones.
 *
 * Supported: dimensions (one unlimited), variables of type byte, char,
 * short, int, float, double, text and numeric attributes, whole-variable
 * and hyperslab reads with conversion to float or double, whole-variable
 * and hyperslab writes from float or double, record variables, default fill
 * values.  Not supported: strided/mapped access, renaming, redefining an
 * existing file.
 */
#ifndef NC3_H
#define NC3_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NC3_NOERR 0
#define NC3_EBADID (-33)
#define NC3_ENFILE (-34)
#define NC3_EEXIST (-35)
#define NC3_EINVAL (-36)
#define NC3_EPERM (-37)
#define NC3_ENOTINDEFINE (-38)
#define NC3_EINDEFINE (-39)
#define NC3_EINVALCOORDS (-40)
#define NC3_EMAXDIMS (-41)
#define NC3_ENAMEINUSE (-42)
#define NC3_ENOTATT (-43)
#define NC3_EMAXATTS (-44)
#define NC3_EBADTYPE (-45)
#define NC3_EBADDIM (-46)
#define NC3_EUNLIMPOS (-47)
#define NC3_EMAXVARS (-48)
#define NC3_ENOTVAR (-49)
#define NC3_EGLOBAL (-50)
#define NC3_ENOTNC (-51)
#define NC3_EMAXNAME (-53)
#define NC3_EUNLIMIT (-54)
#define NC3_ECHAR (-56)
#define NC3_EEDGE (-57)
#define NC3_ERANGE (-60)
#define NC3_ENOMEM (-61)
#define NC3_EIO (-68)

#define NC3_BYTE 1
#define NC3_CHAR 2
#define NC3_SHORT 3
#define NC3_INT 4
#define NC3_FLOAT 5
#define NC3_DOUBLE 6

#define NC3_NOWRITE 0
#define NC3_WRITE 1
#define NC3_CLOBBER 0
#define NC3_NOCLOBBER 4
#define NC3_64BIT_OFFSET 512
#define NC3_UNLIMITED 0L
#define NC3_GLOBAL (-1)

int nc3_create(const char *path, int cmode, int *ncid);
int nc3_open(const char *path, int mode, int *ncid);
int nc3_def_dim(int ncid, const char *name, size_t len, int *dimid);
int nc3_def_var(int ncid, const char *name, int type, int ndims, const int *dimids, int *varid);
int nc3_put_att(int ncid, int varid, const char *name, int type, size_t n, const void *values);
int nc3_copy_att(int ncid_in, int varid_in, const char *name, int ncid_out, int varid_out);
int nc3_enddef(int ncid);
int nc3_sync(int ncid);
int nc3_close(int ncid);

int nc3_inq(int ncid, int *ndims, int *nvars, int *ngatts, int *unlimdimid);
int nc3_inq_dim(int ncid, int dimid, char *name, size_t *len);
int nc3_inq_var(int ncid, int varid, char *name, int *type, int *ndims, int *dimids, int *natts);
int nc3_inq_attname(int ncid, int varid, int attnum, char *name);
int nc3_inq_att(int ncid, int varid, const char *name, int *type, size_t *n);
int nc3_get_att(int ncid, int varid, const char *name, void *values); /* file type, host order */

/* memtype NC3_FLOAT or NC3_DOUBLE; count = NULL: the whole variable */
int nc3_get_vara(int ncid, int varid, const size_t *start, const size_t *count, void *out, int memtype);
int nc3_put_vara(int ncid, int varid, const size_t *start, const size_t *count, const void *in,
                 int memtype);

const char *nc3_strerror(int status);

#ifdef __cplusplus
}
#endif

#endif /* NC3_H */
