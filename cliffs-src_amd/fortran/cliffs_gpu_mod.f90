! ISO_C_BINDING interface of libcliffs_gpu (include/cliffs_gpu.h) for the
! Fortran driver.  Self-contained: it depends on no driver module, every
! array is passed explicitly in the reference's (nX,nY) column-major layout.
!
! The drop-in subroutines with the reference call surface
! (cliffs(iu,irw,n), cliffs.f:17) are in cliffs_dropin.f90, which is
! compiled together with the driver's own MASTER/PARAMETERS modules.
module cliffs_gpu_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CLIFFS_OK = 0
  integer(c_int), parameter, public :: CLIFFS_ERR_ARG = -1
  integer(c_int), parameter, public :: CLIFFS_ERR_HIP = -2
  integer(c_int), parameter, public :: CLIFFS_ERR_NOMEM = -3
  integer(c_int), parameter, public :: CLIFFS_ERR_METRIC = -4
  integer(c_int), parameter, public :: CLIFFS_ERR_NODEVICE = -5
  integer(c_int), parameter, public :: CLIFFS_ERR_STATE = -6

  ! struct cliffs_grid_desc
  type, bind(C), public :: cliffs_grid_desc
    integer(c_int) :: nx, ny
    type(c_ptr) :: dep, s1, s2, zeta
    real(c_double) :: dt, ground, crough
    integer(c_int) :: device
  end type cliffs_grid_desc

  public :: cliffs_create, cliffs_destroy, cliffs_upload_state, cliffs_download_state
  public :: cliffs_download_max, cliffs_reset_max, cliffs_sweep, cliffs_maxwave
  public :: cliffs_step, cliffs_run, cliffs_run_timed, cliffs_gauges, cliffs_sync
  public :: cliffs_stream, cliffs_info, cliffs_last_error, cliffs_health
  public :: cliffs_nest_maps, cliffs_nest_create, cliffs_nest_destroy, cliffs_nest_feed
  public :: cliffs_nest_couple, cliffs_nest_record, cliffs_nest_end_feed, cliffs_nest_child_step
  public :: cliffs_nest_child_start, cliffs_run_gauges_noalloc, cliffs_set_tile_width, cliffs_ensemble_shape
  public :: cliffs_nest_run
  public :: cliffs_run_gauges, cliffs_set_fused, cliffs_set_tile_rows
  public :: cliffs_ensemble_create, cliffs_ensemble_destroy, cliffs_ensemble_run
  public :: cliffs_ensemble_run_timed, cliffs_ensemble_tune, cliffs_ensemble_sync
  public :: cliffs_ensemble_run_gauges
  public :: cliffs_gpu_open, cliffs_error_message
  ! FP64 line mode (include/cliffs_fp64.h)
  public :: cliffs64_create, cliffs64_destroy, cliffs64_upload_state, cliffs64_download_state
  public :: cliffs64_download_max, cliffs64_sweep, cliffs64_maxwave, cliffs64_run
  public :: cliffs64_gauges, cliffs64_health, cliffs64_nest_create, cliffs64_nest_feed, cliffs64_nest_destroy

  interface
    integer(c_int) function cliffs_create(desc, h) bind(C, name="cliffs_create")
      import :: c_int, c_ptr, cliffs_grid_desc
      type(cliffs_grid_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function cliffs_destroy(h) bind(C, name="cliffs_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cliffs_upload_state(h, cel, xvel, yvel) bind(C, name="cliffs_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: cel(*), xvel(*), yvel(*)
    end function
    integer(c_int) function cliffs_download_state(h, cel, xvel, yvel) bind(C, name="cliffs_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: cel(*), xvel(*), yvel(*)
    end function
    integer(c_int) function cliffs_download_max(h, hmax, velmax) bind(C, name="cliffs_download_max")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: hmax(*), velmax(*)
    end function
    integer(c_int) function cliffs_reset_max(h) bind(C, name="cliffs_reset_max")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! edge1/edge2 are passed as c_ptr so that absent edges can be c_null_ptr
    integer(c_int) function cliffs_sweep(h, iu, first, last, edge1, edge2, mx) bind(C, name="cliffs_sweep")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, edge1, edge2
      integer(c_int), value :: iu, first, last, mx
    end function
    integer(c_int) function cliffs_maxwave(h) bind(C, name="cliffs_maxwave")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cliffs_step(h, xe1, xe2, ye1, ye2, mx) bind(C, name="cliffs_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, xe1, xe2, ye1, ye2
      integer(c_int), value :: mx
    end function
    integer(c_int) function cliffs_run(h, nsteps) bind(C, name="cliffs_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function cliffs_run_timed(h, nsteps, ms_total, ms_x, ms_y) bind(C, name="cliffs_run_timed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: ms_total, ms_x, ms_y
    end function
    integer(c_int) function cliffs_gauges(h, n, ig, jg, eta, u, v) bind(C, name="cliffs_gauges")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      integer(c_int), intent(in) :: ig(*), jg(*)
      real(c_double), intent(out) :: eta(*), u(*), v(*)
    end function
    integer(c_int) function cliffs_health(h, courant, nonfinite) bind(C, name="cliffs_health")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h
      real(c_double), intent(out) :: courant
      integer(c_long_long), intent(out) :: nonfinite
    end function
    integer(c_int) function cliffs_sync(h) bind(C, name="cliffs_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cliffs_stream(h, stream) bind(C, name="cliffs_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), intent(out) :: stream
    end function
    integer(c_int) function cliffs_info(h, grav, celmin, ii1, ii2, jj1, jj2) bind(C, name="cliffs_info")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: grav, celmin
      integer(c_int), intent(out) :: ii1, ii2, jj1, jj2
    end function
    type(c_ptr) function cliffs_last_error() bind(C, name="cliffs_last_error")
      import :: c_ptr
    end function
    integer(c_int) function cliffs_nest_maps(nxp, xp, nyp, yp, nxc, xc, nyc, yc, we_ny, we_cy, &
        sn_nx, sn_cx, w_nx, w_cx, e_nx, e_cx, s_ny, s_cy, n_ny, n_cy) bind(C, name="cliffs_nest_maps")
      import :: c_int, c_double
      integer(c_int), value :: nxp, nyp, nxc, nyc
      real(c_double), intent(in) :: xp(*), yp(*), xc(*), yc(*)
      integer(c_int), intent(out) :: we_ny(*), sn_nx(*), w_nx, e_nx, s_ny, n_ny
      real(c_double), intent(out) :: we_cy(*), sn_cx(*), w_cx, e_cx, s_cy, n_cy
    end function
    integer(c_int) function cliffs_nest_create(parent, nxc, nyc, we_ny, we_cy, sn_nx, sn_cx, &
        w_nx, w_cx, e_nx, e_cx, s_ny, s_cy, n_ny, n_cy, nest) bind(C, name="cliffs_nest_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: parent
      integer(c_int), value :: nxc, nyc, w_nx, e_nx, s_ny, n_ny
      integer(c_int), intent(in) :: we_ny(*), sn_nx(*)
      real(c_double), intent(in) :: we_cy(*), sn_cx(*)
      real(c_double), value :: w_cx, e_cx, s_cy, n_cy
      type(c_ptr), intent(out) :: nest
    end function
    integer(c_int) function cliffs_nest_destroy(nest) bind(C, name="cliffs_nest_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: nest
    end function
    integer(c_int) function cliffs_nest_feed(nest, cuke, hmin, iarv, fed, uvq_w, uvq_e, uvq_s, &
        uvq_n) bind(C, name="cliffs_nest_feed")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: nest
      real(c_double), value :: cuke, hmin
      integer(c_int), intent(inout) :: iarv
      integer(c_int), intent(out) :: fed
      real(c_double), intent(out) :: uvq_w(*), uvq_e(*), uvq_s(*), uvq_n(*)
    end function
    integer(c_int) function cliffs_run_gauges(h, every, nsamples, n, ig, jg, eta) &
        bind(C, name="cliffs_run_gauges")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: every, nsamples, n
      integer(c_int), intent(in) :: ig(*), jg(*)
      real(c_float), intent(out) :: eta(*)
    end function
    ! no device allocation: a library-wide device scratch, n*(nsamples+2) <= 2**20
    integer(c_int) function cliffs_run_gauges_noalloc(h, every, nsamples, n, ig, jg, eta) &
        bind(C, name="cliffs_run_gauges_noalloc")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: every, nsamples, n
      integer(c_int), intent(in) :: ig(*), jg(*)
      real(c_float), intent(out) :: eta(*)
    end function
    integer(c_int) function cliffs_set_fused(h, on) bind(C, name="cliffs_set_fused")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function cliffs_set_tile_rows(h, rows) bind(C, name="cliffs_set_tile_rows")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: rows
    end function
    integer(c_int) function cliffs_set_tile_width(h, width) bind(C, name="cliffs_set_tile_width")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: width
    end function
    integer(c_int) function cliffs_ensemble_shape(e, width, rows) bind(C, name="cliffs_ensemble_shape")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: width, rows
    end function
    integer(c_int) function cliffs_nest_couple(nest, child, capacity) bind(C, name="cliffs_nest_couple")
      import :: c_int, c_ptr
      type(c_ptr), value :: nest, child
      integer(c_int), value :: capacity
    end function
    integer(c_int) function cliffs_nest_record(nest, cuke, hmin, t, iarv, nrec) &
        bind(C, name="cliffs_nest_record")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: nest
      real(c_double), value :: cuke, hmin, t
      integer(c_int), intent(inout) :: iarv
      integer(c_int), intent(out) :: nrec
    end function
    integer(c_int) function cliffs_nest_end_feed(nest) bind(C, name="cliffs_nest_end_feed")
      import :: c_int, c_ptr
      type(c_ptr), value :: nest
    end function
    integer(c_int) function cliffs_nest_child_start(nest, cuke, t0) bind(C, name="cliffs_nest_child_start")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: nest
      real(c_double), value :: cuke
      real(c_double), intent(out) :: t0
    end function
    integer(c_int) function cliffs_nest_child_step(nest, t) bind(C, name="cliffs_nest_child_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: nest
      real(c_double), value :: t
    end function
    integer(c_int) function cliffs_nest_run(nest, nrec, every, t0, dt, per, dt_child, cuke, hmin, &
                                            iarv) bind(C, name="cliffs_nest_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: nest
      integer(c_int), value :: nrec, every, per
      real(c_double), value :: t0, dt, dt_child, cuke, hmin
      integer(c_int), intent(inout) :: iarv
    end function
    ! members: an array of n grid handles
    integer(c_int) function cliffs_ensemble_create(members, n, e) bind(C, name="cliffs_ensemble_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: members(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function cliffs_ensemble_destroy(e) bind(C, name="cliffs_ensemble_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function cliffs_ensemble_run(e, nsteps) bind(C, name="cliffs_ensemble_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function cliffs_ensemble_run_timed(e, nsteps, ms_total) &
        bind(C, name="cliffs_ensemble_run_timed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: ms_total
    end function
    integer(c_int) function cliffs_ensemble_tune(e, trial_steps, rows) bind(C, name="cliffs_ensemble_tune")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: trial_steps
      integer(c_int), intent(out) :: rows
    end function
    integer(c_int) function cliffs_ensemble_sync(e) bind(C, name="cliffs_ensemble_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function cliffs_ensemble_run_gauges(e, every, nsamples, n, ig, jg, eta) &
        bind(C, name="cliffs_ensemble_run_gauges")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: e
      integer(c_int), value :: every, nsamples, n
      integer(c_int), intent(in) :: ig(*), jg(*)
      real(c_float), intent(out) :: eta(*)
    end function
    ! FP64 line mode (include/cliffs_fp64.h): the reference's state and line
    ! procedure in double precision on the device
    integer(c_int) function cliffs64_create(desc, h) bind(C, name="cliffs64_create")
      import :: c_int, c_ptr, cliffs_grid_desc
      type(cliffs_grid_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function cliffs64_destroy(h) bind(C, name="cliffs64_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cliffs64_upload_state(h, cel, xvel, yvel) bind(C, name="cliffs64_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: cel(*), xvel(*), yvel(*)
    end function
    integer(c_int) function cliffs64_download_state(h, cel, xvel, yvel) bind(C, name="cliffs64_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: cel(*), xvel(*), yvel(*)
    end function
    integer(c_int) function cliffs64_download_max(h, hmax, velmax) bind(C, name="cliffs64_download_max")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: hmax(*), velmax(*)
    end function
    integer(c_int) function cliffs64_sweep(h, iu, first, last, edge1, edge2, mx) bind(C, name="cliffs64_sweep")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, edge1, edge2
      integer(c_int), value :: iu, first, last, mx
    end function
    integer(c_int) function cliffs64_maxwave(h) bind(C, name="cliffs64_maxwave")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cliffs64_run(h, nsteps) bind(C, name="cliffs64_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function cliffs64_gauges(h, n, I, J, eta, u, v) bind(C, name="cliffs64_gauges")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      integer(c_int), intent(in) :: I(*), J(*)
      real(c_double), intent(out) :: eta(*), u(*), v(*)
    end function
    integer(c_int) function cliffs64_health(h, courant, nonfinite) bind(C, name="cliffs64_health")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h
      real(c_double), intent(out) :: courant
      integer(c_long_long), intent(out) :: nonfinite
    end function
    integer(c_int) function cliffs64_nest_create(parent, nxc, nyc, we_ny, we_cy, sn_nx, sn_cx, &
        w_nx, w_cx, e_nx, e_cx, s_ny, s_cy, n_ny, n_cy, nest) bind(C, name="cliffs64_nest_create")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: parent
      integer(c_int), value :: nxc, nyc, w_nx, e_nx, s_ny, n_ny
      integer(c_int), intent(in) :: we_ny(*), sn_nx(*)
      real(c_double), intent(in) :: we_cy(*), sn_cx(*)
      real(c_double), value :: w_cx, e_cx, s_cy, n_cy
      type(c_ptr), intent(out) :: nest
    end function
    integer(c_int) function cliffs64_nest_feed(nest, cuke, hmin, iarv, fed, uvq_w, uvq_e, uvq_s, uvq_n) &
        bind(C, name="cliffs64_nest_feed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: nest
      real(c_double), value :: cuke, hmin
      integer(c_int), intent(inout) :: iarv
      integer(c_int), intent(out) :: fed
      real(c_double), intent(out) :: uvq_w(*), uvq_e(*), uvq_s(*), uvq_n(*)
    end function
    integer(c_int) function cliffs64_nest_destroy(nest) bind(C, name="cliffs64_nest_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: nest
    end function
  end interface

contains

  ! Create a device grid from the arrays MasterGrid fills (MasterGrid.f:85-121)
  ! and the PARAMETERS dt, ground, crough.  s1/s2/zeta may be absent for 1-D
  ! grids (nx <= 2 or ny <= 2).  fp64 = .true.: a grid in the FP64 line mode
  ! (cliffs64_create) instead of the FP32 one.
  integer(c_int) function cliffs_gpu_open(h, nx, ny, dep, dt, ground, crough, device, s1, s2, zeta, fp64)
    type(c_ptr), intent(out) :: h
    integer, intent(in) :: nx, ny, device
    real(c_double), intent(in), target :: dep(nx, ny)
    real(c_double), intent(in) :: dt, ground, crough
    real(c_double), intent(in), target, contiguous, optional :: s1(:, :), s2(:), zeta(:)
    logical, intent(in), optional :: fp64
    type(cliffs_grid_desc) :: d
    d%nx = nx
    d%ny = ny
    d%dep = c_loc(dep)
    d%s1 = c_null_ptr
    d%s2 = c_null_ptr
    d%zeta = c_null_ptr
    if (present(s1)) d%s1 = c_loc(s1)
    if (present(s2)) d%s2 = c_loc(s2)
    if (present(zeta)) d%zeta = c_loc(zeta)
    d%dt = dt
    d%ground = ground
    d%crough = crough
    d%device = device
    if (present(fp64)) then
      if (fp64) then
        cliffs_gpu_open = cliffs64_create(d, h)
        return
      end if
    end if
    cliffs_gpu_open = cliffs_create(d, h)
  end function cliffs_gpu_open

  ! cliffs_last_error() as a Fortran string
  function cliffs_error_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(cliffs_last_error(), p, [256])
    do i = 1, 256
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function cliffs_error_message

end module cliffs_gpu_mod
