! Fortran declarations of the NetCDF-3 entry points exported by nc3.c
! (the Fortran-77 NetCDF interface: 1-based ids, fastest dimension first).
! The constants carry the published NetCDF values.
module netcdf3_mod
  implicit none
  public

  integer, parameter :: NF_NOERR = 0
  integer, parameter :: NF_NOWRITE = 0, NF_WRITE = 1
  integer, parameter :: NF_CLOBBER = 0, NF_NOCLOBBER = 4, NF_64BIT_OFFSET = 512
  integer, parameter :: NF_UNLIMITED = 0, NF_GLOBAL = 0
  integer, parameter :: NF_BYTE = 1, NF_CHAR = 2, NF_SHORT = 3, NF_INT = 4
  integer, parameter :: NF_FLOAT = 5, NF_REAL = 5, NF_DOUBLE = 6

  integer, external :: nf_create, nf_open, nf_close, nf_sync, nf_enddef
  integer, external :: nf_inq, nf_inq_dim, nf_inq_dimlen, nf_inq_varname, nf_inq_vartype
  integer, external :: nf_inq_var, nf_inq_attname, nf_def_dim, nf_def_var
  integer, external :: nf_put_att_text, nf_copy_att
  integer, external :: nf_get_var_double, nf_get_var_real, nf_get_vara_double, nf_get_vara_real
  integer, external :: nf_put_var_double, nf_put_var_real, nf_put_vara_double, nf_put_vara_real
  character(len=80), external :: nf_strerror
end module netcdf3_mod
