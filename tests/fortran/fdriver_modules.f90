! Module variables a Cliffs driver provides to the drop-in (the names of
! global_modules.f:16-55 that cliffs-src_amd/fortran/cliffs_dropin.f90 reads),
! declared here for the test driver fdriver.f90.
module MASTER
  implicit none
  real*8, parameter :: grav = 9.8
  integer nXn, nYn
  real*8, dimension(:, :), allocatable :: cel, xvel, yvel, dep, hmax, velmax
  real*8, dimension(:, :), allocatable :: s1
  real*8, dimension(:), allocatable :: s2, zeta
  real*8, dimension(:, :), allocatable :: edge1, edge2
end module MASTER

module PARAMETERS
  implicit none
  real*8 ground, crough, dt, cuke
end module PARAMETERS


module NESTED
  implicit none
  integer nests
  integer nXnst(99), nYnst(99), maxlen
  integer, dimension(:, :), allocatable :: WE_ny, SN_nx
  integer, dimension(:), allocatable :: W_nx, E_nx, S_ny, N_ny
  real*8, dimension(:, :), allocatable :: WE_cy, SN_cx
  real*8, dimension(:), allocatable :: W_cx, E_cx, S_cy, N_cy
end module NESTED

! Capture of the nest feed for fnest.f90: output_feed (output_feed.f:3-24
! writes NetCDF records) is replaced by a writer of the raw records.
module feed_capture
  implicit none
  character(len=512) :: feed_out_name = 'feed.bin'
  integer :: feed_unit = 0
  logical :: feed_open = .false.   ! newunit= numbers are negative
end module feed_capture

subroutine output_feed(ncid, nn, mm, uvq, it, time)
  use feed_capture
  implicit none
  integer ncid, nn, mm, it
  real*8 uvq(mm, 3), time
  if (.not. feed_open) then
    open(newunit=feed_unit, file=trim(feed_out_name), access='stream', form='unformatted', &
         status='replace')
    feed_open = .true.
  end if
  write(feed_unit) int(ncid, 4), int(nn, 4), int(it, 4), time, uvq(1:nn, 1:3)
end subroutine output_feed

subroutine feed_close(iarv, irec)
  use feed_capture
  implicit none
  logical iarv
  integer irec
  if (.not. feed_open) then
    open(newunit=feed_unit, file=trim(feed_out_name), access='stream', form='unformatted', &
         status='replace')
    feed_open = .true.
  end if
  write(feed_unit) int(-1, 4), merge(1, 0, iarv), int(irec, 4)
  close(feed_unit)
end subroutine feed_close
