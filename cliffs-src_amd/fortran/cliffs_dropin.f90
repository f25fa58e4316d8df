! Drop-in replacements for the reference hot path, compiled together with
! the driver's own MASTER / PARAMETERS modules (global_modules.f:16-55).
!
!   cliffs_gpu_attach(device)     after the H->c conversion (cliffs_main.f:227):
!                                 device copy of dep/s1/s2/zeta/dt/ground/crough
!                                 and of the state cel/xvel/yvel; hmax/velmax = 0
!   cliffs_gpu_sweep(iu,first,last)
!                                 one whole OpenMP loop of cliffs calls
!                                 (cliffs_main.f:325-329 or 340-344), with the
!                                 driver's edge1/edge2 (cliffs_main.f:318-339)
!   cliffs_gpu_maxwave()          the max-wave loop (cliffs_main.f:347-355)
!   cliffs_gpu_to_host()          state and maxima back into cel/xvel/yvel/
!                                 hmax/velmax, before output_* or
!                                 feed_children_grid read them
!   cliffs_gpu_detach()
!   subroutine cliffs(iu,irw,n)   the reference call surface (cliffs.f:17);
!                                 one line on the device state.  One handle
!                                 belongs to one host thread: call it outside
!                                 OpenMP regions.
module cliffs_dropin
  use iso_c_binding
  use cliffs_gpu_mod
  implicit none
  private
  type(c_ptr), save, public :: cliffs_gpu_handle = c_null_ptr
  type(c_ptr), save, public :: cliffs_gpu_nests(99) = c_null_ptr   ! up to 99 nests (global_modules.f:38)
  public :: cliffs_gpu_attach, cliffs_gpu_sweep, cliffs_gpu_maxwave, cliffs_gpu_attach_nests
  public :: cliffs_gpu_to_host, cliffs_gpu_detach, cliffs_gpu_check

contains

  subroutine cliffs_gpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= CLIFFS_OK) then
      write(*, *) 'cliffs_gpu: ', where, ' failed: ', trim(cliffs_error_message())
      stop 1
    end if
  end subroutine cliffs_gpu_check

  subroutine cliffs_gpu_attach(device)
    use MASTER, only: nXn, nYn, dep, cel, xvel, yvel, s1, s2, zeta
    use PARAMETERS, only: dt, ground, crough
    integer, intent(in) :: device
    integer(c_int) :: rc
    ! the module arrays are not TARGETs: their addresses are taken on the
    ! TARGET dummies of cliffs_gpu_open, inside the call
    if (allocated(s1) .and. allocated(s2)) then
      rc = cliffs_gpu_open(cliffs_gpu_handle, nXn, nYn, dep, dt, ground, crough, device, &
                           s1=s1, s2=s2, zeta=zeta)
    else if (allocated(s1)) then
      rc = cliffs_gpu_open(cliffs_gpu_handle, nXn, nYn, dep, dt, ground, crough, device, s1=s1)
    else
      rc = cliffs_gpu_open(cliffs_gpu_handle, nXn, nYn, dep, dt, ground, crough, device, &
                           s2=s2, zeta=zeta)
    end if
    call cliffs_gpu_check(rc, 'cliffs_create')
    call cliffs_gpu_check(cliffs_upload_state(cliffs_gpu_handle, cel, xvel, yvel), &
                          'cliffs_upload_state')
  end subroutine cliffs_gpu_attach

  subroutine cliffs_gpu_sweep(iu, first, last)
    use MASTER, only: edge1, edge2
    integer, intent(in) :: iu, first, last
    if (allocated(edge1)) then
      call sweep_with_edges(iu, first, last, edge1, edge2)
    else
      call cliffs_gpu_check(cliffs_sweep(cliffs_gpu_handle, iu, first, last, c_null_ptr, &
                                         c_null_ptr, 0), 'cliffs_sweep')
    end if
  end subroutine cliffs_gpu_sweep

  subroutine sweep_with_edges(iu, first, last, e1, e2)
    integer, intent(in) :: iu, first, last
    real(c_double), intent(in), target, contiguous :: e1(:, :), e2(:, :)
    call cliffs_gpu_check(cliffs_sweep(cliffs_gpu_handle, iu, first, last, c_loc(e1), &
                                       c_loc(e2), int(size(e1, 1), c_int)), 'cliffs_sweep')
  end subroutine sweep_with_edges

  subroutine cliffs_gpu_maxwave()
    call cliffs_gpu_check(cliffs_maxwave(cliffs_gpu_handle), 'cliffs_maxwave')
  end subroutine cliffs_gpu_maxwave

  subroutine cliffs_gpu_to_host()
    use MASTER, only: cel, xvel, yvel, hmax, velmax
    call cliffs_gpu_check(cliffs_download_state(cliffs_gpu_handle, cel, xvel, yvel), &
                          'cliffs_download_state')
    call cliffs_gpu_check(cliffs_download_max(cliffs_gpu_handle, hmax, velmax), &
                          'cliffs_download_max')
  end subroutine cliffs_gpu_to_host

  ! Device nests from the maps NestedGrid left in module NESTED
  ! (NestedGrid.f:98-119); call after cliffs_gpu_attach.
  subroutine cliffs_gpu_attach_nests()
    use NESTED, only: nests, nXnst, nYnst, WE_ny, WE_cy, SN_nx, SN_cx, W_nx, W_cx, E_nx, E_cx, &
                      S_ny, S_cy, N_ny, N_cy
    integer :: ns
    do ns = 1, nests
      call nest_create(ns, nXnst(ns), nYnst(ns), WE_ny(:, ns), WE_cy(:, ns), SN_nx(:, ns), &
                       SN_cx(:, ns), W_nx(ns), W_cx(ns), E_nx(ns), E_cx(ns), S_ny(ns), &
                       S_cy(ns), N_ny(ns), N_cy(ns))
    end do
  end subroutine cliffs_gpu_attach_nests

  subroutine nest_create(ns, nxc, nyc, we_ny, we_cy, sn_nx, sn_cx, w_nx, w_cx, e_nx, e_cx, &
                         s_ny, s_cy, n_ny, n_cy)
    integer, intent(in) :: ns, nxc, nyc, w_nx, e_nx, s_ny, n_ny
    integer, intent(in) :: we_ny(*), sn_nx(*)
    real(c_double), intent(in) :: we_cy(*), sn_cx(*), w_cx, e_cx, s_cy, n_cy
    integer(c_int), allocatable :: wy(:), sx(:)
    allocate(wy(nyc), sx(nxc))
    wy = we_ny(1:nyc)
    sx = sn_nx(1:nxc)
    call cliffs_gpu_check(cliffs_nest_create(cliffs_gpu_handle, nxc, nyc, wy, we_cy, sx, sn_cx, &
                                             w_nx, w_cx, e_nx, e_cx, s_ny, s_cy, n_ny, n_cy, &
                                             cliffs_gpu_nests(ns)), 'cliffs_nest_create')
  end subroutine nest_create

  subroutine cliffs_gpu_detach()
    integer :: ns
    do ns = 1, size(cliffs_gpu_nests)
      if (c_associated(cliffs_gpu_nests(ns))) then
        call cliffs_gpu_check(cliffs_nest_destroy(cliffs_gpu_nests(ns)), 'cliffs_nest_destroy')
        cliffs_gpu_nests(ns) = c_null_ptr
      end if
    end do
    if (c_associated(cliffs_gpu_handle)) then
      call cliffs_gpu_check(cliffs_destroy(cliffs_gpu_handle), 'cliffs_destroy')
      cliffs_gpu_handle = c_null_ptr
    end if
  end subroutine cliffs_gpu_detach

end module cliffs_dropin

! The reference entry point (cliffs.f:17), advancing line irw of direction
! iu on the device-resident state.
subroutine cliffs(iu, irw, n)
  use cliffs_dropin, only: cliffs_gpu_sweep
  implicit none
  integer iu, irw, n
  call cliffs_gpu_sweep(iu, irw, irw)
end subroutine cliffs

! The reference nest-feed entry point (feed_children_grid.f:10), computing
! the arrival test and the four edge interpolations on the device and
! handing the edges to the driver's output_feed as the reference does
! (feed_children_grid.f:41-42, 50, 81-82, 112-113, 145-146, 176-177).
subroutine feed_children_grid(ns, bn_id, irec, time, hmin, iarv)
  use iso_c_binding
  use cliffs_gpu_mod
  use cliffs_dropin, only: cliffs_gpu_nests, cliffs_gpu_check
  use NESTED, only: nXnst, nYnst, maxlen
  use PARAMETERS, only: cuke
  implicit none
  integer ns, irec, bn_id(4)
  real*8 time, hmin
  logical iarv
  real(c_double), allocatable :: w(:, :), e(:, :), s(:, :), n(:, :), uvq(:, :)
  integer(c_int) :: arv, fed
  integer :: ly, lx
  ly = nYnst(ns)
  lx = nXnst(ns)
  allocate(w(ly, 3), e(ly, 3), s(lx, 3), n(lx, 3), uvq(maxlen, 3))
  arv = 0
  if (iarv) arv = 1
  call cliffs_gpu_check(cliffs_nest_feed(cliffs_gpu_nests(ns), cuke, hmin, arv, fed, w, e, s, n), &
                        'cliffs_nest_feed')
  if ((.not. iarv) .and. arv /= 0) then
    write(9, *) 'reached child grid ', ns, ' at ', time, 'sec'
    call flush(9)
  end if
  iarv = arv /= 0
  if (fed /= 0) then
    irec = irec + 1
    uvq = 0
    uvq(1:ly, :) = w
    call output_feed(bn_id(1), ly, maxlen, uvq, irec, time)
    uvq(1:ly, :) = e
    call output_feed(bn_id(2), ly, maxlen, uvq, irec, time)
    uvq = 0
    uvq(1:lx, :) = s
    call output_feed(bn_id(3), lx, maxlen, uvq, irec, time)
    uvq(1:lx, :) = n
    call output_feed(bn_id(4), lx, maxlen, uvq, irec, time)
  end if
end subroutine feed_children_grid
