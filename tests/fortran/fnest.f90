! Test driver for the nest-feed drop-in: fills the NESTED maps with
! cliffs_nest_maps (NestedGrid.f:98-119), attaches parent and nest, calls the
! reference entry point feed_children_grid(ns,bn_id,irec,time,hmin,iarv)
! once, and lets output_feed (below) capture the four edges.
!
! usage: fnest <input.bin> <output.bin>
program fnest
  use iso_c_binding
  use MASTER
  use PARAMETERS
  use NESTED
  use cliffs_gpu_mod
  use cliffs_dropin
  use feed_capture
  implicit none
  character(len=512) :: fin
  integer :: u, nxc, nyc, irec, bn_id(4)
  integer(4) :: hdr(4)
  real*8 :: par(4), t
  real*8, allocatable :: xp(:), yp(:), xc(:), yc(:)
  logical :: iarv
  integer(c_int) :: wnx, enx, sny, nny
  integer(c_int), allocatable :: wey(:), snx(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, feed_out_name)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nXn = hdr(1); nYn = hdr(2); nxc = hdr(3); nyc = hdr(4)
  read(u) par
  dt = par(1); ground = par(2); crough = par(3); cuke = par(4)
  allocate(dep(nXn, nYn), cel(nXn, nYn), xvel(nXn, nYn), yvel(nXn, nYn))
  allocate(hmax(nXn, nYn), velmax(nXn, nYn), s1(nXn - 1, nYn), s2(nYn - 1), zeta(nYn))
  allocate(xp(nXn), yp(nYn), xc(nxc), yc(nyc))
  read(u) dep, cel, xvel, yvel, s1, s2, zeta, xp, yp, xc, yc
  close(u)

  nests = 1
  nXnst(1) = nxc
  nYnst(1) = nyc
  maxlen = max(nxc, nyc)
  allocate(WE_ny(nyc, 1), WE_cy(nyc, 1), SN_nx(nxc, 1), SN_cx(nxc, 1))
  allocate(W_nx(1), E_nx(1), S_ny(1), N_ny(1), W_cx(1), E_cx(1), S_cy(1), N_cy(1))
  allocate(wey(nyc), snx(nxc))
  call cliffs_gpu_check(cliffs_nest_maps(nXn, xp, nYn, yp, nxc, xc, nyc, yc, wey, WE_cy(:, 1), &
                        snx, SN_cx(:, 1), wnx, W_cx(1), enx, E_cx(1), sny, S_cy(1), nny, &
                        N_cy(1)), 'cliffs_nest_maps')
  WE_ny(:, 1) = wey
  SN_nx(:, 1) = snx
  W_nx(1) = wnx; E_nx(1) = enx; S_ny(1) = sny; N_ny(1) = nny

  open(unit=9, file='/dev/null')
  call cliffs_gpu_attach(0)
  call cliffs_gpu_attach_nests()
  bn_id = (/ 11, 12, 13, 14 /)
  irec = 0
  iarv = .false.
  t = 1.5d0
  call feed_children_grid(1, bn_id, irec, t, ground, iarv)
  call cliffs_gpu_detach()
  call feed_close(iarv, irec)
end program fnest
