! The Cliffs driver with the time steps on the GPU.
!
!   cliffs_gpu_driver <output dir/prefix> <source dir/> <boundary prefix|0>
!                     <area prefix|0> <param dir/param file> [notes]
!
! Same command line, parameter file, input files and output files as the
! reference's driver (cliffs_main.f:56-112 command line, allabout.f parameter
! order, MasterGrid.f / NestedGrid.f grids, get_boundary_feed.f and
! get_initial_state.f inputs, opennc4*.f / output_*.f outputs).  The state
! lives on the GPU (libcliffs_gpu through cliffs_gpu_mod); the host keeps the
! inputs, interpolates the boundary feed in time and writes the outputs at
! their cadences.  NetCDF files go through the nc3 shim (netcdf3_mod), so no
! NetCDF library is needed.
!
! Compile without floating-point contraction: the host arithmetic that feeds
! the solver (initial-field interpolation, edge blending) must round as the
! reference's real*8 code does.

module drv_consts
  implicit none
  integer, parameter :: dp = kind(1.d0)
  ! the reference's constants (global_modules.f:17-18): single-precision
  ! literals stored in real*8
  real(dp), parameter :: GRAV = real(9.8, dp), PI = real(3.1415926536, dp), EARTHR = 6371009
  real(4), parameter :: NAN4 = -1.0e32
  integer, parameter :: LOGU = 9, PARU = 15
end module drv_consts

module drv_nc
  use netcdf3_mod
  use drv_consts
  implicit none
contains
  ! any NetCDF failure ends the run, with the library's message in the log
  subroutine nc(status)
    integer, intent(in) :: status
    if (status /= NF_NOERR) then
      write(LOGU, *) trim(nf_strerror(status))
      flush(LOGU)
      stop 2
    end if
  end subroutine nc
end module drv_nc

module drv_data
  use drv_consts
  implicit none
  ! run parameters (allabout.f order)
  logical :: cartesian, freeze
  integer :: nests, itopo, steps_total, quake, seaout, lonsub, latsub, bndout, maxout
  integer :: ngauges = 0, gout = 1
  integer :: pline = 0   ! lines of the parameter file read so far
  integer, allocatable :: ig(:), jg(:)
  real(dp) :: cuke, ground, crough, dwall, dt
  ! master grid
  character(len=200) :: gridname
  integer :: nx, ny
  real(dp), allocatable :: xcrd(:), ycrd(:), dep(:, :), s1(:, :), s2(:), zeta(:)
  ! nested grids
  character(len=200) :: nestname(99)
  integer :: nxs(99), nys(99)
  real(dp), allocatable :: xn(:, :), yn(:, :)
  ! boundary feed (time, W/E (ny,3,nt), S/N (nx,3,nt))
  integer :: ntin = 0
  real(dp), allocatable :: tin(:), west(:, :, :), east(:, :, :), south(:, :, :), north(:, :, :)
  ! initial fields
  integer :: nxsrc, nysrc
  real(dp), allocatable :: xsrc(:), ysrc(:), usrc(:, :), vsrc(:, :), qsrc(:, :)
  logical :: xrun, yrun
end module drv_data

module drv_input
  use drv_nc
  use drv_data
  implicit none
contains

  ! bisection of a sorted lattice (NestedGrid.f:129-148): xx(ip) < xp <=
  ! xx(ip+1) with the weight cp of node ip
  subroutine pinpoint(xx, n, xp, ip, cp)
    integer, intent(in) :: n
    real(dp), intent(in) :: xx(n), xp
    integer, intent(out) :: ip
    real(dp), intent(out) :: cp
    integer :: hi, mid
    ip = 1
    hi = n
    do while (ip + 2 <= hi)
      mid = (ip + hi) / 2
      if (xp > xx(mid)) then
        ip = mid
      else
        hi = mid
      end if
    end do
    if (n > 1) then
      cp = (xx(ip + 1) - xp) / (xx(ip + 1) - xx(ip))
    else
      cp = 0.5_dp
    end if
  end subroutine pinpoint

  ! a coordinate or field variable of any real type, as real*8
  subroutine get_real8(ncid, varid, out)
    integer, intent(in) :: ncid, varid
    real(dp), intent(out) :: out(*)
    call nc(nf_get_var_double(ncid, varid, out))
  end subroutine get_real8

  ! One line of the parameter file: its leading value (list-directed, the
  ! rest of the line is comment, as allabout.f reads it).  A missing or
  ! malformed value sets err = -1 and names the line, where the reference's
  ! reads would stop the program.
  subroutine param_int(v, err)
    integer, intent(out) :: v
    integer, intent(inout) :: err
    integer :: ios
    v = 0
    if (err /= 0) return
    pline = pline + 1
    read(PARU, *, iostat=ios) v
    if (ios /= 0) call param_bad(err)
  end subroutine param_int

  subroutine param_real(v, err)
    real(dp), intent(out) :: v
    integer, intent(inout) :: err
    integer :: ios
    v = 0.0_dp
    if (err /= 0) return
    pline = pline + 1
    read(PARU, *, iostat=ios) v
    if (ios /= 0) call param_bad(err)
  end subroutine param_real

  subroutine param_name(v, err)
    character(len=*), intent(out) :: v
    integer, intent(inout) :: err
    integer :: ios
    v = ' '
    if (err /= 0) return
    pline = pline + 1
    read(PARU, '(a)', iostat=ios) v
    if (ios /= 0 .or. len_trim(v) == 0) call param_bad(err)
  end subroutine param_name

  subroutine param_bad(err)
    integer, intent(inout) :: err
    err = -1
    write(LOGU, '(1x,a,i0)') 'parameter file: missing or malformed value on line ', pline
  end subroutine param_bad

  ! parameter file and the grids it names (allabout.f:6-97)
  subroutine read_params(indir, pfile, err)
    character(len=*), intent(in) :: indir, pfile
    integer, intent(out) :: err
    integer :: num, i, ios
    err = 0
    pline = 0
    write(LOGU, '(1x,2a)') 'Read Computational parameters: ', trim(indir)//trim(pfile)
    open(unit=PARU, file=trim(indir)//trim(pfile), status='old', form='formatted', iostat=ios)
    if (ios /= 0) then
      err = -1
      write(LOGU, '(1x,2a)') 'parameter file not found: ', trim(indir)//trim(pfile)
      return
    end if
    call param_int(num, err)
    if (err /= 0) return
    cartesian = num == 1
    if (cartesian) then
      write(LOGU, *) 'CARTESIAN COORDINATE SYSTEM'
    else
      write(LOGU, *) 'SPHERICAL (LON,LAT) COORDINATE SYSTEM'
    end if
    call read_master(indir, err)
    if (err /= 0) return
    call param_int(nests, err)
    if (err /= 0) return
    write(LOGU, *) 'Number of grids enclosed in Master Grid: ', nests
    if (nests < 0 .or. nests > 99) then
      err = -1
      write(LOGU, *) 'number of nested grids must be 0..99'
      return
    end if
    if (nests > 0 .and. (nx < 3 .or. ny < 3)) then
      err = -1
      write(LOGU, *) 'nesting in 1D is not permitted'
      return
    end if
    if (nests > 0) call read_nests(indir, err)
    if (err /= 0) return
    call param_real(cuke, err)
    call param_real(ground, err)
    call param_real(crough, err)
    call param_int(itopo, err)
    call param_real(dwall, err)
    call param_real(dt, err)
    call param_int(steps_total, err)
    call param_int(quake, err)
    call param_int(num, err)
    freeze = num == 0
    call param_int(seaout, err)
    call param_int(lonsub, err)
    call param_int(latsub, err)
    call param_int(bndout, err)
    call param_int(maxout, err)
    call param_int(ngauges, err)
    if (err == 0 .and. ngauges >= 1) then
      call param_int(gout, err)
      allocate(ig(ngauges), jg(ngauges))
      do i = 1, ngauges
        if (err /= 0) exit
        pline = pline + 1
        read(PARU, *, iostat=ios) ig(i), jg(i)
        if (ios /= 0) call param_bad(err)
      end do
    end if
    close(PARU)
    if (err /= 0) return
    write(LOGU, *) 'cuke, ground, crough:', cuke, ground, crough
    write(LOGU, *) 'dt, steps:', dt, steps_total, ' topo', itopo, ' quake', quake, ' freeze ', freeze
    write(LOGU, *) 'outputs every (steps): sea', seaout, ' feed', bndout, ' max', maxout, ' gauges', gout
    flush(LOGU)
  end subroutine read_params

  ! master grid: lattice, depth and metrics (MasterGrid.f:7-127)
  subroutine read_master(indir, err)
    character(len=*), intent(in) :: indir
    integer, intent(out) :: err
    character(len=200) :: fname
    integer :: ncid, i, j, ndims, nvar, ngatts, unlim, k1, k2
    real(dp) :: beta, alpha
    err = 0
    call param_name(fname, err)
    if (err /= 0) return
    k1 = index(fname, '/', .true.) + 1
    k2 = index(fname, '.nc') - 1
    gridname = fname(k1:k2)
    write(LOGU, '(1x,2a)') 'Master bathy/topo file:  ', trim(indir)//trim(fname)
    call nc(nf_open(trim(indir)//trim(fname), NF_NOWRITE, ncid))
    call nc(nf_inq_dimlen(ncid, 1, nx))
    call nc(nf_inq_dimlen(ncid, 2, ny))
    write(LOGU, *) 'grid size ', nx, ny
    allocate(xcrd(nx), ycrd(ny), dep(nx, ny))
    call get_real8(ncid, 1, xcrd)
    call get_real8(ncid, 2, ycrd)
    call get_real8(ncid, 3, dep)
    beta = PI * EARTHR / 180.
    if (nx > 2) then
      allocate(s1(nx - 1, ny))
      do j = 1, ny
        if (cartesian) then
          s1(:, j) = xcrd(2:nx) - xcrd(1:nx - 1)
        else
          alpha = cos(PI / 180. * ycrd(j))
          s1(:, j) = alpha * (beta * (xcrd(2:nx) - xcrd(1:nx - 1)))
        end if
      end do
    end if
    if (ny > 2) then
      allocate(s2(ny - 1), zeta(ny))
      if (cartesian) then
        s2 = ycrd(2:ny) - ycrd(1:ny - 1)
        zeta = 0
      else
        s2 = beta * (ycrd(2:ny) - ycrd(1:ny - 1))
        alpha = 0.125 / EARTHR
        do i = 1, ny
          zeta(i) = alpha * tan(PI / 180. * ycrd(i))
        end do
      end if
    end if
    ! an extra variable (river width factor) replaces zeta (MasterGrid.f:118-121)
    call nc(nf_inq(ncid, ndims, nvar, ngatts, unlim))
    if (nvar > 3) then
      if (.not. allocated(zeta)) allocate(zeta(ny))
      call get_real8(ncid, nvar, zeta)
    end if
    call nc(nf_close(ncid))
  end subroutine read_master

  ! nested grid lattices (NestedGrid.f:9-127); the maps are built by the
  ! library (cliffs_nest_maps)
  subroutine read_nests(indir, err)
    character(len=*), intent(in) :: indir
    integer, intent(out) :: err
    character(len=200) :: fname
    integer :: ncid(99), ns, mx, my
    err = 0
    do ns = 1, nests
      call param_name(fname, err)
      if (err /= 0) return
      nestname(ns) = fname(1:index(fname, '.nc') - 1)
      call nc(nf_open(trim(indir)//trim(fname), NF_NOWRITE, ncid(ns)))
      call nc(nf_inq_dimlen(ncid(ns), 1, nxs(ns)))
      call nc(nf_inq_dimlen(ncid(ns), 2, nys(ns)))
      write(LOGU, *) 'nested Grid', ns, trim(indir)//trim(fname), nxs(ns), nys(ns)
    end do
    mx = maxval(nxs(1:nests))
    my = maxval(nys(1:nests))
    allocate(xn(mx, nests), yn(my, nests))
    do ns = 1, nests
      call get_real8(ncid(ns), 1, xn(:, ns))
      call get_real8(ncid(ns), 2, yn(:, ns))
      call nc(nf_close(ncid(ns)))
      if (xn(1, ns) < xcrd(1) .or. xn(nxs(ns), ns) > xcrd(nx) .or. &
          yn(1, ns) < ycrd(1) .or. yn(nys(ns), ns) > ycrd(ny)) then
        err = -1
        write(LOGU, *) 'Grid ', ns, ' not within Master grid limits'
        return
      end if
    end do
  end subroutine read_nests

  ! boundary feed files <src><prefix>_<grid>_{west,east,south,north}.nc and
  ! the first record whose edge |eta| reaches cuke (get_boundary_feed.f:8-170)
  subroutine read_boundary(srcdir, bname, first, err)
    character(len=*), intent(in) :: srcdir, bname
    integer, intent(out) :: first, err
    real(dp) :: w(4), wave
    integer :: k
    err = 0
    first = 0
    write(LOGU, '(1x,2a)') 'Boundary Input: ', trim(srcdir)//trim(bname)
    if (xrun) then
      call read_side('west', ny, west, .true., err)
      if (err == 0) call read_side('east', ny, east, .false., err)
    end if
    if (err == 0 .and. yrun) then
      call read_side('south', nx, south, .not. xrun, err)
      if (err == 0) call read_side('north', nx, north, .false., err)
    end if
    if (err /= 0) return
    w = 0
    do k = 1, ntin
      if (xrun) then
        w(1) = max(maxval(west(:, 3, k)), -minval(west(:, 3, k)))
        w(2) = max(maxval(east(:, 3, k)), -minval(east(:, 3, k)))
      end if
      if (yrun) then
        w(3) = max(maxval(south(:, 3, k)), -minval(south(:, 3, k)))
        w(4) = max(maxval(north(:, 3, k)), -minval(north(:, 3, k)))
      end if
      wave = maxval(w)
      if (wave >= cuke) exit
    end do
    if (k >= ntin) then
      err = -1
      write(LOGU, *) 'bndr signal below threshold'
    else
      first = k
      write(LOGU, *) 'feed exceeds threshold at record', k, ',', tin(k), 's, with ', wave, ' m'
    end if
  contains
    subroutine read_side(side, npts, vals, with_time, err)
      character(len=*), intent(in) :: side
      integer, intent(in) :: npts
      real(dp), allocatable, intent(inout) :: vals(:, :, :)
      logical, intent(in) :: with_time
      integer, intent(out) :: err
      integer :: ncid, np, nt
      err = 0
      if (nf_open(trim(srcdir)//trim(bname)//'_'//side//'.nc', NF_NOWRITE, ncid) /= NF_NOERR) then
        err = -1
        write(LOGU, '(1x,a)') trim(srcdir)//trim(bname)//'_'//side//'.nc not found'
        return
      end if
      call nc(nf_inq_dimlen(ncid, 1, np))
      call nc(nf_inq_dimlen(ncid, 3, nt))
      if (with_time) then
        ntin = nt
        if (ntin < 2) then
          err = -1
          write(LOGU, *) 'unsufficient time-series length: ', ntin, ' samples'
          return
        end if
        allocate(tin(ntin))
        call get_real8(ncid, 2, tin)
      end if
      if (np /= npts .or. nt /= ntin) then
        err = -1
        write(LOGU, *) 'size mismatch: expected ', npts, ntin, ' provided ', np, nt
        return
      end if
      allocate(vals(npts, 3, ntin))
      call get_real8(ncid, 1, vals)
      call nc(nf_close(ncid))
    end subroutine read_side
  end subroutine read_boundary

  ! initial fields <src><prefix>_{h,u,v}.nc at the start time
  ! (get_initial_state.f:9-165); values below -1e10 are 0
  subroutine read_initial(srcdir, aprfx, irec, hin, uin, vin, err)
    character(len=*), intent(in) :: srcdir, aprfx
    integer, intent(inout) :: irec
    logical, intent(out) :: hin, uin, vin
    integer, intent(out) :: err
    integer :: ids(3), ncid, ndims, nvar, ngatts, unlim, i, dlen, tid, xid, yid, nt, isrc, st(3), ct(3)
    character(len=200) :: dname
    real(dp), allocatable :: tsrc(:)
    real(4), allocatable :: sea(:, :, :)
    real(dp) :: wgt
    character(len=1), parameter :: suf(3) = ['u', 'v', 'h']
    logical :: have(3)
    err = 0
    do i = 1, 3
      have(i) = nf_open(trim(srcdir)//trim(aprfx)//'_'//suf(i)//'.nc', NF_NOWRITE, ids(i)) == NF_NOERR
      write(LOGU, '(1x,4a,1x,l1)') 'Initial ', suf(i), ': ', trim(srcdir)//trim(aprfx)//'_'//suf(i)//'.nc', have(i)
    end do
    uin = have(1)
    vin = have(2)
    hin = have(3)
    if (hin) then
      ncid = ids(3)
    else if (uin) then
      ncid = ids(1)
    else if (vin) then
      ncid = ids(2)
    else
      if (ntin < 1) err = -1
      return
    end if
    call nc(nf_inq(ncid, ndims, nvar, ngatts, unlim))
    do i = 1, ndims
      call nc(nf_inq_dim(ncid, i, dname, dlen))
      if (dname(1:3) == 'tim') then
        nt = dlen
      else if (dname(1:3) == 'xxx' .or. dname(1:3) == 'lon') then
        nxsrc = dlen
      else if (dname(1:3) == 'yyy' .or. dname(1:3) == 'lat') then
        nysrc = dlen
      end if
    end do
    do i = 1, nvar
      call nc(nf_inq_varname(ncid, i, dname))
      if (dname(1:3) == 'tim') then
        tid = i
      else if (dname(1:3) == 'xxx' .or. dname(1:3) == 'lon') then
        xid = i
      else if (dname(1:3) == 'yyy' .or. dname(1:3) == 'lat') then
        yid = i
      end if
    end do
    allocate(xsrc(nxsrc), ysrc(nysrc), tsrc(nt))
    call get_real8(ncid, xid, xsrc)
    call get_real8(ncid, yid, ysrc)
    call get_real8(ncid, tid, tsrc)
    if (ntin < 2) then
      if (.not. allocated(tin)) allocate(tin(1))
      irec = 1
      isrc = 1
      tin(1) = tsrc(1)
    else
      call pinpoint(tsrc, nt, tin(irec), isrc, wgt)
    end if
    write(LOGU, *) 'Initial conditions read at ', tsrc(isrc), ' applied at ', tin(irec)
    allocate(sea(nxsrc, nysrc, 1), usrc(nxsrc, nysrc), vsrc(nxsrc, nysrc), qsrc(nxsrc, nysrc))
    usrc = 0
    vsrc = 0
    qsrc = 0
    st = [1, 1, isrc]
    ct = [nxsrc, nysrc, 1]
    do i = 1, 3
      if (.not. have(i)) cycle
      call nc(nf_inq(ids(i), ndims, nvar, ngatts, unlim))
      ! the field is the file's last variable, read as real*4
      call nc(nf_get_vara_real(ids(i), nvar, st, ct, sea))
      call nc(nf_close(ids(i)))
      where (sea(:, :, 1) < -1.0e10) sea(:, :, 1) = 0
      select case (i)
      case (1)
        usrc = sea(:, :, 1)
      case (2)
        vsrc = sea(:, :, 1)
      case (3)
        qsrc = sea(:, :, 1)
      end select
    end do
  end subroutine read_initial

  ! add the source field qq(nxq,nyq) on lattice (xq,yq), interpolated
  ! bilinearly (linearly in 1-D), to pp(kx,ky) on lattice (xp,yp)
  ! (apply_initial_conditions.f:7-70: same weights and operation order)
  subroutine add_field(nxq, nyq, xq, yq, qq, kx, ky, xp, yp, pp)
    integer, intent(in) :: nxq, nyq, kx, ky
    real(dp), intent(in) :: xq(nxq), yq(nyq), qq(nxq, nyq), xp(kx), yp(ky)
    real(dp), intent(inout) :: pp(kx, ky)
    real(dp) :: x1, x2, y1, y2, xcc, ycc, q
    integer :: i, j, ii, jj
    real(dp), parameter :: tiny_ = real(0.1e-20, dp)  ! an FP32 literal (apply_initial_conditions.f:12)
    x1 = xq(1) - tiny_
    x2 = xq(nxq) + tiny_
    y1 = yq(1) - tiny_
    y2 = yq(nyq) + tiny_
    if (nxq > 2 .and. nyq > 2) then
      do i = 1, kx
        if (xp(i) < x1 .or. xp(i) > x2) cycle
        call pinpoint(xq, nxq, xp(i), ii, xcc)
        do j = 1, ky
          if (yp(j) < y1 .or. yp(j) > y2) cycle
          call pinpoint(yq, nyq, yp(j), jj, ycc)
          q = (qq(ii, jj) * xcc + qq(ii + 1, jj) * (1 - xcc)) * ycc + &
              (qq(ii, jj + 1) * xcc + qq(ii + 1, jj + 1) * (1 - xcc)) * (1 - ycc)
          pp(i, j) = pp(i, j) + q
        end do
      end do
    else if (nyq == 1 .and. nxq > 2) then
      do i = 1, kx
        if (xp(i) < x1 .or. xp(i) > x2) cycle
        call pinpoint(xq, nxq, xp(i), ii, xcc)
        q = qq(ii, 1) * xcc + qq(ii + 1, 1) * (1 - xcc)
        pp(i, :) = pp(i, :) + q
      end do
    else if (nxq == 1 .and. nyq > 2) then
      do j = 1, ky
        if (yp(j) < y1 .or. yp(j) > y2) cycle
        call pinpoint(yq, nyq, yp(j), jj, ycc)
        q = qq(1, jj) * ycc + qq(1, jj + 1) * (1 - ycc)
        pp(:, j) = pp(:, j) + q
      end do
    end if
  end subroutine add_field

end module drv_input

module drv_output
  use drv_nc
  use drv_data
  implicit none

  type nc_series
    integer :: ncid = -1, tid = 0, vid(3) = 0, n = 0
  end type nc_series

contains

  subroutine axis_names(xname, yname)
    character(len=*), intent(out) :: xname, yname
    if (cartesian) then
      xname = 'xxx'
      yname = 'yyy'
    else
      xname = 'lon'
      yname = 'lat'
    end if
  end subroutine axis_names

  ! <case>_maxwave.nc: MaxE, MaxV over the grid (opennc4max.f)
  subroutine open_max(fname, f)
    character(len=*), intent(in) :: fname
    type(nc_series), intent(out) :: f
    character(len=3) :: xname, yname
    integer :: dx, dy, xid, yid
    call axis_names(xname, yname)
    call nc(nf_create(trim(fname), NF_CLOBBER, f%ncid))
    call nc(nf_def_dim(f%ncid, xname, nx, dx))
    call nc(nf_def_dim(f%ncid, yname, ny, dy))
    call nc(nf_def_var(f%ncid, xname, NF_DOUBLE, 1, [dx], xid))
    call nc(nf_def_var(f%ncid, yname, NF_DOUBLE, 1, [dy], yid))
    call nc(nf_def_var(f%ncid, 'MaxE', NF_REAL, 2, [dx, dy], f%vid(1)))
    call nc(nf_def_var(f%ncid, 'MaxV', NF_REAL, 2, [dx, dy], f%vid(2)))
    call text(f%ncid, f%vid(1), 'name', 'Maximal Water Elevation')
    call text(f%ncid, f%vid(1), 'units', 'm')
    call text(f%ncid, f%vid(2), 'name', 'Maximal Current')
    call text(f%ncid, f%vid(2), 'units', 'm/s')
    call axis_atts(f%ncid, xid, yid, 'x-axis', 'y-axis', 'longitude', 'latitude', 'degrees East', &
                   'degrees North')
    call nc(nf_enddef(f%ncid))
    call nc(nf_put_var_double(f%ncid, xid, xcrd))
    call nc(nf_put_var_double(f%ncid, yid, ycrd))
    call nc(nf_sync(f%ncid))
    write(LOGU, '(1x,2a)') 'Maxwave output to: ', trim(fname)
  end subroutine open_max

  ! MaxE = hmax^2/g - dep and MaxV = sqrt(velmax) where hmax > celmin, else
  ! the real*4 NAN (output_maxwave.f:14-24)
  subroutine write_max(f, hmax, velmax, celmin)
    type(nc_series), intent(in) :: f
    real(dp), intent(in) :: hmax(nx, ny), velmax(nx, ny), celmin
    real(4), allocatable :: e(:, :), v(:, :)
    allocate(e(nx, ny), v(nx, ny))
    where (hmax > celmin)
      e = hmax**2 / GRAV - dep
      v = sqrt(velmax)
    elsewhere
      e = NAN4
      v = NAN4
    end where
    call nc(nf_put_var_real(f%ncid, f%vid(1), e))
    call nc(nf_put_var_real(f%ncid, f%vid(2), v))
    call nc(nf_sync(f%ncid))
  end subroutine write_max

  ! <case>_sea_{h,u,v}.nc: subsampled snapshots (opennc4sea.f)
  subroutine open_sea(fname, who, gx, gy, f)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: who
    real(dp), intent(in) :: gx(:), gy(:)
    type(nc_series), intent(out) :: f
    character(len=3) :: xname, yname
    character(len=2), parameter :: vname(3) = ['ha', 'ua', 'va']
    integer :: dx, dy, dt_, xid, yid
    call axis_names(xname, yname)
    call nc(nf_create(trim(fname), NF_CLOBBER, f%ncid))
    call nc(nf_def_dim(f%ncid, xname, size(gx), dx))
    call nc(nf_def_dim(f%ncid, yname, size(gy), dy))
    call nc(nf_def_dim(f%ncid, 'time', NF_UNLIMITED, dt_))
    call nc(nf_def_var(f%ncid, xname, NF_DOUBLE, 1, [dx], xid))
    call nc(nf_def_var(f%ncid, yname, NF_DOUBLE, 1, [dy], yid))
    call nc(nf_def_var(f%ncid, 'time', NF_DOUBLE, 1, [dt_], f%tid))
    call nc(nf_def_var(f%ncid, vname(who), NF_REAL, 3, [dx, dy, dt_], f%vid(1)))
    select case (who)
    case (1)
      call text(f%ncid, f%vid(1), 'name', 'Surface Displacement')
      call text(f%ncid, f%vid(1), 'units', 'meters')
    case (2)
      call text(f%ncid, f%vid(1), 'name', 'U (zonal) Velocity')
      call text(f%ncid, f%vid(1), 'units', 'meters/second')
    case (3)
      call text(f%ncid, f%vid(1), 'name', 'V (meridional) Velocity')
      call text(f%ncid, f%vid(1), 'units', 'meters/second')
    end select
    call axis_atts(f%ncid, xid, yid, 'subset_x', 'subset_y', 'subset_longitude', 'subset_latitude', &
                   'degrees_east', 'degrees_north')
    call text(f%ncid, f%tid, 'name', 'time')
    call text(f%ncid, f%tid, 'units', 'seconds')
    call nc(nf_enddef(f%ncid))
    call nc(nf_put_var_double(f%ncid, xid, gx))
    call nc(nf_put_var_double(f%ncid, yid, gy))
    write(LOGU, '(1x,2a)') 'Screenshot output to: ', trim(fname)
  end subroutine open_sea

  ! one snapshot record of a field (output_sea_state.f:14-58)
  subroutine write_sea(f, t, vals)
    type(nc_series), intent(inout) :: f
    real(dp), intent(in) :: t
    real(4), intent(in) :: vals(:, :)
    f%n = f%n + 1
    call nc(nf_put_vara_double(f%ncid, f%tid, [f%n], [1], [t]))
    call nc(nf_put_vara_real(f%ncid, f%vid(1), [1, 1, f%n], [size(vals, 1), size(vals, 2), 1], vals))
    call nc(nf_sync(f%ncid))
  end subroutine write_sea

  ! <case>_gages.nc (opennc4gages.f)
  subroutine open_gauges(fname, f)
    character(len=*), intent(in) :: fname
    type(nc_series), intent(out) :: f
    integer :: dp_, dt_, xid, yid, k
    real(dp) :: gx(ngauges), gy(ngauges)
    call nc(nf_create(trim(fname), NF_CLOBBER, f%ncid))
    call nc(nf_def_dim(f%ncid, 'point', ngauges, dp_))
    call nc(nf_def_dim(f%ncid, 'time', NF_UNLIMITED, dt_))
    if (cartesian) then
      call nc(nf_def_var(f%ncid, 'xxx', NF_DOUBLE, 1, [dp_], xid))
      call nc(nf_def_var(f%ncid, 'yyy', NF_DOUBLE, 1, [dp_], yid))
    else
      call nc(nf_def_var(f%ncid, 'lon', NF_DOUBLE, 1, [dp_], xid))
      call nc(nf_def_var(f%ncid, 'lat', NF_DOUBLE, 1, [dp_], yid))
    end if
    call nc(nf_def_var(f%ncid, 'time', NF_DOUBLE, 1, [dt_], f%tid))
    call nc(nf_def_var(f%ncid, 'gage', NF_DOUBLE, 2, [dp_, dt_], f%vid(1)))
    f%vid(2:3) = -1
    if (xrun) call nc(nf_def_var(f%ncid, 'u', NF_DOUBLE, 2, [dp_, dt_], f%vid(2)))
    if (yrun) call nc(nf_def_var(f%ncid, 'v', NF_DOUBLE, 2, [dp_, dt_], f%vid(3)))
    call text(f%ncid, f%vid(1), 'name', 'wave elevation at the gage')
    call text(f%ncid, f%vid(1), 'units', 'meters')
    if (xrun) then
      call text(f%ncid, f%vid(2), 'name', 'x (zonal) velocity')
      call text(f%ncid, f%vid(2), 'units', 'm/s')
    end if
    if (yrun) then
      call text(f%ncid, f%vid(3), 'name', 'y (meridional) velocity')
      call text(f%ncid, f%vid(3), 'units', 'm/s')
    end if
    call axis_atts(f%ncid, xid, yid, 'x-axis', 'y-axis', 'longitude', 'latitude', 'degrees_east', &
                   'degrees_north')
    call text(f%ncid, f%tid, 'name', 'time')
    call text(f%ncid, f%tid, 'units', 'seconds')
    call text(f%ncid, NF_GLOBAL, 'model', 'Cliffs')
    call nc(nf_enddef(f%ncid))
    do k = 1, ngauges
      gx(k) = xcrd(ig(k))
      gy(k) = ycrd(jg(k))
    end do
    call nc(nf_put_var_double(f%ncid, xid, gx))
    call nc(nf_put_var_double(f%ncid, yid, gy))
    write(LOGU, '(1x,2a)') 'Time-series output to  ', trim(fname)
  end subroutine open_gauges

  subroutine write_gauges(f, t, eta, u, v)
    type(nc_series), intent(inout) :: f
    real(dp), intent(in) :: t, eta(:), u(:), v(:)
    f%n = f%n + 1
    call nc(nf_put_vara_double(f%ncid, f%tid, [f%n], [1], [t]))
    call nc(nf_put_vara_double(f%ncid, f%vid(1), [1, f%n], [ngauges, 1], eta))
    if (f%vid(2) > 0) call nc(nf_put_vara_double(f%ncid, f%vid(2), [1, f%n], [ngauges, 1], u))
    if (f%vid(3) > 0) call nc(nf_put_vara_double(f%ncid, f%vid(3), [1, f%n], [ngauges, 1], v))
    call nc(nf_sync(f%ncid))
  end subroutine write_gauges

  ! <case>_<nest>_{west,east,south,north}.nc: the boundary feed of a nested
  ! grid (opennc4feed.f), one record per feed (output_feed.f)
  subroutine open_feed(bname, mxn, myn, f)
    character(len=*), intent(in) :: bname
    integer, intent(in) :: mxn, myn
    type(nc_series), intent(out) :: f(4)
    character(len=5), parameter :: side(4) = ['west ', 'east ', 'south', 'north']
    integer :: k, dp_, du, dt_, npnt
    do k = 1, 4
      npnt = myn
      if (k > 2) npnt = mxn
      call nc(nf_create(trim(bname)//'_'//trim(side(k))//'.nc', NF_CLOBBER, f(k)%ncid))
      call nc(nf_def_dim(f(k)%ncid, 'pnt', npnt, dp_))
      call nc(nf_def_dim(f(k)%ncid, 'uvq', 3, du))
      call nc(nf_def_dim(f(k)%ncid, 'tim', NF_UNLIMITED, dt_))
      call nc(nf_def_var(f(k)%ncid, 'vals', NF_DOUBLE, 3, [dp_, du, dt_], f(k)%vid(1)))
      call nc(nf_def_var(f(k)%ncid, 'time', NF_DOUBLE, 1, [dt_], f(k)%tid))
      call nc(nf_enddef(f(k)%ncid))
    end do
  end subroutine open_feed

  subroutine write_feed(f, t, uvq)
    type(nc_series), intent(inout) :: f
    real(dp), intent(in) :: t, uvq(:, :)
    f%n = f%n + 1
    call nc(nf_put_vara_double(f%ncid, f%tid, [f%n], [1], [t]))
    call nc(nf_put_vara_double(f%ncid, f%vid(1), [1, 1, f%n], [size(uvq, 1), 3, 1], uvq))
    call nc(nf_sync(f%ncid))
  end subroutine write_feed

  subroutine text(ncid, varid, name, val)
    integer, intent(in) :: ncid, varid
    character(len=*), intent(in) :: name, val
    call nc(nf_put_att_text(ncid, varid, name, len(val), val))
  end subroutine text

  subroutine axis_atts(ncid, xid, yid, cx, cy, sx, sy, ux, uy)
    integer, intent(in) :: ncid, xid, yid
    character(len=*), intent(in) :: cx, cy, sx, sy, ux, uy
    if (cartesian) then
      call text(ncid, xid, 'name', cx)
      call text(ncid, xid, 'units', 'meters')
      call text(ncid, yid, 'name', cy)
      call text(ncid, yid, 'units', 'meters')
    else
      call text(ncid, xid, 'name', sx)
      call text(ncid, xid, 'units', ux)
      call text(ncid, yid, 'name', sy)
      call text(ncid, yid, 'units', uy)
    end if
  end subroutine axis_atts

  subroutine close_nc(f)
    type(nc_series), intent(inout) :: f
    if (f%ncid >= 0) call nc(nf_close(f%ncid))
    f%ncid = -1
  end subroutine close_nc

end module drv_output

program cliffs_gpu_driver
  use iso_c_binding
  use cliffs_gpu_mod
  use drv_consts
  use drv_nc
  use drv_data
  use drv_input
  use drv_output
  implicit none

  character(len=200) :: arg, casetitle, outdir, srcdir, bprfx, aprfx, indir, pfile, casename
  character(len=1000) :: notes
  integer :: nargs, k1, i, j, irec, istep, err, ns, mx, nxsub, nysub, nshot
  logical :: hin, uin, vin, bndr, ended
  logical :: warned = .false.   ! non-finite state reported
  ! CLIFFS_PRECISION=fp64 in the environment: the FP64 line mode
  ! (include/cliffs_fp64.h), the reference's FP64 results bit for bit
  logical :: fp64
  character(len=16) :: prec
  integer :: ii1, ii2, jj1, jj2
  real(dp) :: t, cc, celmin, t0, t1
  real(dp), allocatable, target :: cel(:, :), xvel(:, :), yvel(:, :), hmax(:, :), velmax(:, :)
  real(dp), allocatable, target :: xe1(:, :), xe2(:, :), ye1(:, :), ye2(:, :)
  real(dp), allocatable :: geta(:), gu(:), gv(:), fw(:, :), fe(:, :), fs(:, :), fn(:, :)
  real(4), allocatable :: snap(:, :, :)
  type(c_ptr) :: h, nest(99)
  type(c_ptr) :: pe(4)
  type(nc_series) :: fmax, fsea(3), fgau, ffeed(4, 99)
  integer(c_int) :: rc, ia, fed
  integer, allocatable :: we_ny(:), sn_nx(:)
  real(dp), allocatable :: we_cy(:), sn_cx(:)
  integer :: w_nx, e_nx, s_ny, n_ny
  real(dp) :: w_cx, e_cx, s_cy, n_cy
  integer :: iarv(99)

  nargs = command_argument_count()
  if (nargs < 5) then
    write(*, *) 'Help: cliffs_gpu_driver <output dir/prefix> <source dir/> <boundary prefix, 0 for no input>', &
                ' <area prefix, 0 for no input> <param dir/param file> [<short notes>]'
    stop 1
  end if
  call get_command_argument(1, arg)
  k1 = index(arg, '/', .true.)
  casetitle = arg(k1 + 1:)
  outdir = arg(1:k1)
  call get_command_argument(2, srcdir)
  call get_command_argument(3, bprfx)
  call get_command_argument(4, aprfx)
  call get_command_argument(5, arg)
  k1 = index(arg, '/', .true.)
  pfile = arg(k1 + 1:)
  indir = arg(1:k1)
  notes = ' '
  do i = 6, nargs
    call get_command_argument(i, arg)
    notes = trim(notes)//' '//trim(arg)
  end do

  casename = trim(outdir)//trim(casetitle)
  open(unit=LOGU, file=trim(casename)//'_log.txt', status='unknown', form='formatted')
  call cpu_time(t0)
  write(LOGU, *) trim(adjustl(notes))
  write(LOGU, *) 'Case Title: ', trim(casetitle)
  write(LOGU, *) 'Input Data: ', trim(bprfx), ', ', trim(aprfx)
  call get_environment_variable('CLIFFS_PRECISION', prec)
  fp64 = trim(prec) == 'fp64'
  if (fp64) then
    write(LOGU, *) 'Time steps on the GPU (libcliffs_gpu), FP64 line mode'
  else
    write(LOGU, *) 'Time steps on the GPU (libcliffs_gpu)'
  end if

  call read_params(indir, pfile, err)
  if (err /= 0) call finish(1)

  ! vertical wall at depth dwall when there is no topography (cliffs_main.f:130-137)
  if (itopo == 0) where (dep <= dwall) dep = -999

  ! 1-D and 2-D grids (cliffs_main.f:140-164)
  xrun = nx > 2
  yrun = ny > 2
  if (.not. xrun) lonsub = 1
  if (.not. yrun) latsub = 1
  mx = 1
  if (xrun) mx = max(mx, ny)
  if (yrun) mx = max(mx, nx)
  ii1 = 1
  ii2 = 1
  jj1 = 1
  jj2 = 1
  if (xrun) then
    ii1 = 2
    ii2 = nx - 1
  end if
  if (yrun) then
    jj1 = 2
    jj2 = ny - 1
  end if

  ! inputs (cliffs_main.f:167-187)
  if (trim(bprfx) == '0' .and. trim(aprfx) == '0') then
    write(LOGU, *) 'no input of any kind'
    call finish(1)
  end if
  hin = .false.
  uin = .false.
  vin = .false.
  bndr = .false.
  irec = 1
  if (trim(bprfx) /= '0') then
    call read_boundary(srcdir, trim(bprfx)//'_'//trim(gridname), irec, err)
    if (err /= 0) call finish(1)
    bndr = .true.
  end if
  if (trim(aprfx) /= '0') then
    call read_initial(srcdir, aprfx, irec, hin, uin, vin, err)
    if (err /= 0) call finish(1)
  end if
  if (ntin < 2) freeze = .false.
  t = tin(irec)

  ! state: depth, initial fields, then celerity (cliffs_main.f:190-227)
  allocate(cel(nx, ny), xvel(nx, ny), yvel(nx, ny), hmax(nx, ny), velmax(nx, ny))
  cel = dep
  xvel = 0
  yvel = 0
  if (uin) call add_field(nxsrc, nysrc, xsrc, ysrc, usrc, nx, ny, xcrd, ycrd, xvel)
  if (vin) call add_field(nxsrc, nysrc, xsrc, ysrc, vsrc, nx, ny, xcrd, ycrd, yvel)
  if (hin) then
    if (quake == 1) then          ! the source moves the bottom
      qsrc = -qsrc
      call add_field(nxsrc, nysrc, xsrc, ysrc, qsrc, nx, ny, xcrd, ycrd, dep)
    else                          ! the source moves the surface
      call add_field(nxsrc, nysrc, xsrc, ysrc, qsrc, nx, ny, xcrd, ycrd, cel)
    end if
  end if
  where (cel <= ground)
    cel = 0
  elsewhere
    cel = sqrt(GRAV * cel)
  end where
  celmin = sqrt(GRAV * ground)

  ! the grid on the device
  if (xrun .and. yrun) then
    rc = cliffs_gpu_open(h, nx, ny, dep, dt, ground, crough, 0, s1, s2, zeta, fp64=fp64)
  else if (xrun) then
    rc = cliffs_gpu_open(h, nx, ny, dep, dt, ground, crough, 0, s1=s1, fp64=fp64)
  else
    rc = cliffs_gpu_open(h, nx, ny, dep, dt, ground, crough, 0, s2=s2, zeta=zeta, fp64=fp64)
  end if
  call gpu(rc, 'cliffs_create')
  if (fp64) then
    call gpu(cliffs64_upload_state(h, cel, xvel, yvel), 'cliffs64_upload_state')
  else
    call gpu(cliffs_upload_state(h, cel, xvel, yvel), 'cliffs_upload_state')
  end if

  ! nested grids: parent maps on the device and feed files
  do ns = 1, nests
    allocate(we_ny(nys(ns)), we_cy(nys(ns)), sn_nx(nxs(ns)), sn_cx(nxs(ns)))
    call gpu(cliffs_nest_maps(nx, xcrd, ny, ycrd, nxs(ns), xn(:, ns), nys(ns), yn(:, ns), we_ny, we_cy, &
             sn_nx, sn_cx, w_nx, w_cx, e_nx, e_cx, s_ny, s_cy, n_ny, n_cy), 'cliffs_nest_maps')
    if (fp64) then
      call gpu(cliffs64_nest_create(h, nxs(ns), nys(ns), we_ny, we_cy, sn_nx, sn_cx, w_nx, w_cx, e_nx, e_cx, &
               s_ny, s_cy, n_ny, n_cy, nest(ns)), 'cliffs64_nest_create')
    else
      call gpu(cliffs_nest_create(h, nxs(ns), nys(ns), we_ny, we_cy, sn_nx, sn_cx, w_nx, w_cx, e_nx, e_cx, &
               s_ny, s_cy, n_ny, n_cy, nest(ns)), 'cliffs_nest_create')
    end if
    deallocate(we_ny, we_cy, sn_nx, sn_cx)
    call open_feed(trim(casename)//'_'//trim(nestname(ns)), nxs(ns), nys(ns), ffeed(:, ns))
  end do
  iarv = 0

  ! outputs and their first records (cliffs_main.f:229-285)
  call open_max(trim(casename)//'_maxwave.nc', fmax)
  nshot = 0
  istep = 0                         ! the outputs at the start time
  if (seaout <= steps_total) then
    nxsub = (nx - 1) / lonsub + 1
    nysub = (ny - 1) / latsub + 1
    call open_sea(trim(casename)//'_sea_h.nc', 1, xcrd(1:nx:lonsub), ycrd(1:ny:latsub), fsea(1))
    if (xrun) call open_sea(trim(casename)//'_sea_u.nc', 2, xcrd(1:nx:lonsub), ycrd(1:ny:latsub), fsea(2))
    if (yrun) call open_sea(trim(casename)//'_sea_v.nc', 3, xcrd(1:nx:lonsub), ycrd(1:ny:latsub), fsea(3))
    allocate(snap(nxsub, nysub, 3))
    call snapshot()
  end if
  if (ngauges >= 1) then
    allocate(geta(ngauges), gu(ngauges), gv(ngauges))
    call open_gauges(trim(casename)//'_gages.nc', fgau)
    call gauges()
  end if
  allocate(xe1(mx, 3), xe2(mx, 3), ye1(mx, 3), ye2(mx, 3))
  xe1 = 0
  xe2 = 0
  ye1 = 0
  ye2 = 0
  call feed_nests()                 ! at the start time (cliffs_main.f:292-300)

  ! time loop (cliffs_main.f:303-378)
  ended = .false.
  do istep = 1, steps_total
    t = t + dt
    do while (bndr .and. t >= tin(irec))   ! tin(irec-1) <= t < tin(irec)
      if (irec == ntin) then
        bndr = .false.
        xe1 = 0
        xe2 = 0
        ye1 = 0
        ye2 = 0
        write(LOGU, *) 'bndr feed ended at step', istep
        ended = freeze                    ! stop when the forcing stops
      else
        irec = irec + 1
      end if
    end do
    if (ended) exit
    if (bndr) then
      cc = (tin(irec) - t) / (tin(irec) - tin(irec - 1))
      if (xrun) then
        xe1(1:ny, :) = west(:, :, irec - 1) * cc + west(:, :, irec) * (1 - cc)
        xe2(1:ny, :) = east(:, :, irec - 1) * cc + east(:, :, irec) * (1 - cc)
      end if
      if (yrun) then
        ye1(1:nx, :) = south(:, :, irec - 1) * cc + south(:, :, irec) * (1 - cc)
        ye2(1:nx, :) = north(:, :, irec - 1) * cc + north(:, :, irec) * (1 - cc)
      end if
      pe = [c_loc(xe1), c_loc(xe2), c_loc(ye1), c_loc(ye2)]
    else
      pe = c_null_ptr
    end if
    if (fp64) then   ! x-sweep, y-sweep, maxima (cliffs_main.f:325-355)
      if (xrun) call gpu(cliffs64_sweep(h, 1, int(jj1, c_int), int(jj2, c_int), pe(1), pe(2), int(mx, c_int)), &
                         'cliffs64_sweep')
      if (yrun) call gpu(cliffs64_sweep(h, 2, int(ii1, c_int), int(ii2, c_int), pe(3), pe(4), int(mx, c_int)), &
                         'cliffs64_sweep')
      call gpu(cliffs64_maxwave(h), 'cliffs64_maxwave')
    else
      call gpu(cliffs_step(h, pe(1), pe(2), pe(3), pe(4), int(mx, c_int)), 'cliffs_step')
    end if
    if (nests > 0 .and. mod(istep, bndout) == 0) call feed_nests()
    if (mod(istep, seaout) == 0) call snapshot()
    if (ngauges >= 1 .and. mod(istep, gout) == 0) call gauges()
    if (mod(istep, maxout) == 0) call maxwave()
  end do

  call maxwave()
  do ns = 1, nests
    do i = 1, 4
      call close_nc(ffeed(i, ns))
    end do
    if (fp64) then
      call gpu(cliffs64_nest_destroy(nest(ns)), 'cliffs64_nest_destroy')
    else
      call gpu(cliffs_nest_destroy(nest(ns)), 'cliffs_nest_destroy')
    end if
  end do
  call close_nc(fgau)
  do i = 1, 3
    call close_nc(fsea(i))
  end do
  call close_nc(fmax)
  if (fp64) then
    call gpu(cliffs64_destroy(h), 'cliffs64_destroy')
  else
    call gpu(cliffs_destroy(h), 'cliffs_destroy')
  end if
  call cpu_time(t1)
  write(LOGU, *) 'Run finished: ', istep - 1, ' steps, ', t1 - t0, ' s of host CPU time'
  call finish(0)

contains

  subroutine gpu(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= CLIFFS_OK) then
      write(LOGU, *) what, ' failed: ', trim(cliffs_error_message())
      write(*, *) what, ' failed: ', trim(cliffs_error_message())
      call finish(3)
    end if
  end subroutine gpu

  subroutine finish(code)
    integer, intent(in) :: code
    flush(LOGU)
    close(LOGU)
    if (code /= 0) stop code
    stop
  end subroutine finish

  ! sea-state snapshot (output_sea_state.f:14-58)
  subroutine snapshot()
    integer :: ii, jj
    real(dp) :: c
    if (fp64) then
      call gpu(cliffs64_download_state(h, cel, xvel, yvel), 'cliffs64_download_state')
    else
      call gpu(cliffs_download_state(h, cel, xvel, yvel), 'cliffs_download_state')
    end if
    do j = 1, nysub
      jj = latsub * (j - 1) + 1
      do i = 1, nxsub
        ii = lonsub * (i - 1) + 1
        c = cel(ii, jj)
        if (c > celmin) then
          snap(i, j, 1) = c**2 / GRAV - dep(ii, jj)
          snap(i, j, 2) = xvel(ii, jj)
          snap(i, j, 3) = yvel(ii, jj)
        else
          snap(i, j, :) = NAN4
        end if
      end do
    end do
    nshot = nshot + 1
    call write_sea(fsea(1), t, snap(:, :, 1))
    if (xrun) call write_sea(fsea(2), t, snap(:, :, 2))
    if (yrun) call write_sea(fsea(3), t, snap(:, :, 3))
    write(LOGU, *) 'screenshot', nshot, 'at', t, 'sec'
    call health()
  end subroutine snapshot

  ! largest Courant number and non-finite cells of the state (cliffs_health);
  ! the reference checks neither, so a blow-up is reported, not acted on
  subroutine health()
    real(c_double) :: courant
    integer(c_long_long) :: nbad
    if (fp64) then
      call gpu(cliffs64_health(h, courant, nbad), 'cliffs64_health')
    else
      call gpu(cliffs_health(h, courant, nbad), 'cliffs_health')
    end if
    write(LOGU, '(1x,a,f8.4)') 'max Courant number ', courant
    if (nbad > 0 .and. .not. warned) then
      write(LOGU, '(1x,a,i0,a,i0)') 'WARNING: state not finite in ', nbad, ' cells at step ', istep
      warned = .true.
    end if
  end subroutine health

  subroutine gauges()
    if (fp64) then
      call gpu(cliffs64_gauges(h, int(ngauges, c_int), ig, jg, geta, gu, gv), 'cliffs64_gauges')
    else
      call gpu(cliffs_gauges(h, int(ngauges, c_int), ig, jg, geta, gu, gv), 'cliffs_gauges')
    end if
    call write_gauges(fgau, t, geta, gu, gv)
  end subroutine gauges

  subroutine maxwave()
    if (fp64) then
      call gpu(cliffs64_download_max(h, hmax, velmax), 'cliffs64_download_max')
    else
      call gpu(cliffs_download_max(h, hmax, velmax), 'cliffs_download_max')
    end if
    call write_max(fmax, hmax, velmax, celmin)
  end subroutine maxwave

  ! feed_children_grid for every nest: arrival test, then one record per
  ! call on each of the four edge files
  subroutine feed_nests()
    do ns = 1, nests
      allocate(fw(nys(ns), 3), fe(nys(ns), 3), fs(nxs(ns), 3), fn(nxs(ns), 3))
      ia = iarv(ns)
      if (fp64) then
        call gpu(cliffs64_nest_feed(nest(ns), cuke, ground, ia, fed, fw, fe, fs, fn), 'cliffs64_nest_feed')
      else
        call gpu(cliffs_nest_feed(nest(ns), cuke, ground, ia, fed, fw, fe, fs, fn), 'cliffs_nest_feed')
      end if
      if (ia /= 0 .and. iarv(ns) == 0) write(LOGU, *) 'reached child grid ', ns, ' at ', t, 'sec'
      iarv(ns) = ia
      if (fed /= 0) then
        call write_feed(ffeed(1, ns), t, fw)
        call write_feed(ffeed(2, ns), t, fe)
        call write_feed(ffeed(3, ns), t, fs)
        call write_feed(ffeed(4, ns), t, fn)
      end if
      deallocate(fw, fe, fs, fn)
    end do
  end subroutine feed_nests

end program cliffs_gpu_driver
