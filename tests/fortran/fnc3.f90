! Test program for the NetCDF-3 shim through its Fortran-77 entry points.
!
!   fnc3 <grid.nc> <out prefix>
!
! Reads a grid file as MasterGrid does (dimension lengths, variable types,
! coordinates and depth as real*8 and real*4) and prints what it read; then
! writes <prefix>_fixed.nc (fixed-size variables, text attributes) and
! <prefix>_rec.nc (an unlimited dimension, two record variables written
! record by record and a hyperslab), reads a record back and prints it.
! tests/test_netcdf3.py checks everything against scipy's reader.
program fnc3
  use netcdf3_mod
  implicit none
  character(len=400) :: gname, prefix
  integer :: ncid, nx, ny, tx, ty, tb, ndims, nvar, ngatts, unlim, k, i, j
  integer :: dx, dy, dp, dt, vx, vy, ve, vg, vt, vs, st
  double precision, allocatable :: x(:), y(:), dep(:, :), g(:, :)
  real, allocatable :: dep4(:, :), e(:, :), sea(:, :, :)
  character(len=64) :: dname
  integer :: dlen

  call get_command_argument(1, gname)
  call get_command_argument(2, prefix)

  call check(nf_open(trim(gname), NF_NOWRITE, ncid))
  call check(nf_inq(ncid, ndims, nvar, ngatts, unlim))
  print '(a,4i6)', 'inq', ndims, nvar, ngatts, unlim
  call check(nf_inq_dimlen(ncid, 1, nx))
  call check(nf_inq_dimlen(ncid, 2, ny))
  call check(nf_inq_dim(ncid, 1, dname, dlen))
  print '(a,1x,a,1x,i6)', 'dim1', trim(dname), dlen
  call check(nf_inq_vartype(ncid, 1, tx))
  call check(nf_inq_vartype(ncid, 2, ty))
  call check(nf_inq_vartype(ncid, 3, tb))
  call check(nf_inq_varname(ncid, 3, dname))
  print '(a,2i6,3i3,1x,a)', 'grid', nx, ny, tx, ty, tb, trim(dname)
  allocate(x(nx), y(ny), dep(nx, ny), dep4(nx, ny))
  call check(nf_get_var_double(ncid, 1, x))
  call check(nf_get_var_double(ncid, 2, y))
  call check(nf_get_var_double(ncid, 3, dep))
  call check(nf_get_var_real(ncid, 3, dep4))
  call check(nf_close(ncid))
  print '(a,3es25.16)', 'x', x(1), x(nx), sum(x)
  print '(a,3es25.16)', 'dep', dep(1, 1), dep(nx, ny), sum(dep)
  print '(a,es25.16)', 'dep4', sum(dble(dep4))
  print '(a,i6)', 'badopen', merge(1, 0, nf_open(trim(prefix)//'_missing.nc', NF_NOWRITE, k) /= NF_NOERR)
  print '(a,1x,a)', 'strerror', trim(nf_strerror(-49))

  ! fixed-size variables: the maxwave schema (opennc4max.f)
  allocate(e(nx, ny))
  do j = 1, ny
    do i = 1, nx
      e(i, j) = real(i) + 1000.0 * real(j) + 0.25
    end do
  end do
  call check(nf_create(trim(prefix)//'_fixed.nc', NF_CLOBBER, ncid))
  call check(nf_def_dim(ncid, 'xxx', nx, dx))
  call check(nf_def_dim(ncid, 'yyy', ny, dy))
  call check(nf_def_var(ncid, 'xxx', NF_DOUBLE, 1, [dx], vx))
  call check(nf_def_var(ncid, 'yyy', NF_DOUBLE, 1, [dy], vy))
  call check(nf_def_var(ncid, 'MaxE', NF_REAL, 2, [dx, dy], ve))
  call check(nf_put_att_text(ncid, ve, 'name', 23, 'Maximal Water Elevation'))
  call check(nf_put_att_text(ncid, ve, 'units', 1, 'm'))
  call check(nf_put_att_text(ncid, NF_GLOBAL, 'model', 6, 'Cliffs'))
  call check(nf_enddef(ncid))
  call check(nf_put_var_double(ncid, vx, x))
  call check(nf_put_var_double(ncid, vy, y))
  call check(nf_put_var_real(ncid, ve, e))
  call check(nf_close(ncid))

  ! record variables: the gauge and sea-state schemas (opennc4gages.f, opennc4sea.f)
  allocate(g(3, 1), sea(nx, ny, 1))
  call check(nf_create(trim(prefix)//'_rec.nc', NF_CLOBBER, ncid))
  call check(nf_def_dim(ncid, 'point', 3, dp))
  call check(nf_def_dim(ncid, 'xxx', nx, dx))
  call check(nf_def_dim(ncid, 'yyy', ny, dy))
  call check(nf_def_dim(ncid, 'time', NF_UNLIMITED, dt))
  call check(nf_def_var(ncid, 'time', NF_DOUBLE, 1, [dt], vt))
  call check(nf_def_var(ncid, 'gage', NF_DOUBLE, 2, [dp, dt], vg))
  call check(nf_def_var(ncid, 'ha', NF_REAL, 3, [dx, dy, dt], vs))
  call check(nf_enddef(ncid))
  do k = 1, 3
    g(:, 1) = [1.5d0 * k, -2.0d0 * k, 1.0d-3 * k]
    call check(nf_put_vara_double(ncid, vt, [k], [1], [10.0d0 * k]))
    call check(nf_put_vara_double(ncid, vg, [1, k], [3, 1], g))
    if (k /= 2) then          ! record 2 of 'ha' stays at the fill value
      sea(:, :, 1) = e * real(k)
      call check(nf_put_vara_real(ncid, vs, [1, 1, k], [nx, ny, 1], sea))
    end if
    call check(nf_sync(ncid))
  end do
  ! a hyperslab inside record 3: columns 2..4 of the first row
  call check(nf_put_vara_real(ncid, vs, [2, 1, 3], [3, 1, 1], [-1.0, -2.0, -3.0]))
  call check(nf_inq_dimlen(ncid, 4, k))
  print '(a,i6)', 'records', k
  call check(nf_get_vara_real(ncid, vs, [1, 1, 3], [nx, ny, 1], sea))
  print '(a,4es16.7)', 'rec3', sea(1, 1, 1), sea(2, 1, 1), sea(4, 1, 1), sea(5, 1, 1)
  call check(nf_close(ncid))
  st = nf_close(ncid)
  print '(a,i6)', 'closeagain', st

contains

  subroutine check(status)
    integer, intent(in) :: status
    if (status /= NF_NOERR) then
      print '(a,i6,1x,a)', 'error', status, trim(nf_strerror(status))
      stop 1
    end if
  end subroutine check

end program fnc3
