! Test-only harness around the reference solver core (global_modules.f +
! cliffs.f, compiled in place from the reference tree by oracle/Makefile).
!
! This file is oracle infrastructure: only tests/ and the golden-vector
! script load the library built from it.  It fills the reference's module
! arrays (MASTER, PARAMETERS; global_modules.f:16-55) from C pointers and
! calls the unmodified `cliffs(iu,irw,n)` (cliffs.f:17) in the order the
! reference time loop does (cliffs_main.f:317-355).  The two loops below
! restate cliffs_main.f:325-329 / 340-344 and the max-wave update
! cliffs_main.f:347-355, which live in a file that needs NetCDF.
module ref_harness
  use iso_c_binding
  use MASTER
  use PARAMETERS
  implicit none
  integer :: h_ii1, h_ii2, h_jj1, h_jj2, h_mx
contains

  ! Set up a grid: depth (nx,ny) column-major, metrics as MasterGrid.f:85-121
  ! leaves them: s1(nx-1,ny), s2(ny-1), zeta(ny).  Line ranges and xrun/yrun
  ! follow cliffs_main.f:140-164; celmin follows cliffs_main.f:227.
  subroutine ref_init(nx, ny, dep_in, s1_in, s2_in, zeta_in, dt_in, &
                      ground_in, crough_in) bind(C, name="ref_init")
    integer(c_int), value :: nx, ny
    real(c_double), intent(in) :: dep_in(nx*ny), s1_in(*), s2_in(*), zeta_in(*)
    real(c_double), value :: dt_in, ground_in, crough_in
    integer :: nc, mc

    call ref_free()
    nXn = nx
    nYn = ny
    nXYn = max(nx, ny)
    allocate(dep(nx, ny), cel(nx, ny), xvel(nx, ny), yvel(nx, ny))
    allocate(hmax(nx, ny), velmax(nx, ny))
    allocate(s1(max(nx-1, 1), ny), s2(max(ny-1, 1)), zeta(ny))
    dep = reshape(dep_in(1:nx*ny), (/nx, ny/))
    s1 = 0
    s2 = 0
    zeta = 0
    if (nx > 2) s1 = reshape(s1_in(1:(nx-1)*ny), (/nx-1, ny/))
    if (ny > 2) then
      s2 = s2_in(1:ny-1)
      zeta = zeta_in(1:ny)
    end if
    cel = 0
    xvel = 0
    yvel = 0
    hmax = 0
    velmax = 0
    dt = dt_in
    ground = ground_in
    crough = crough_in
    celmin = sqrt(grav*ground)

    if (nXn > 2) then
      h_ii1 = 2
      h_ii2 = nXn - 1
      xrun = .true.
      nc = nYn
    else
      h_ii1 = 1
      h_ii2 = 1
      xrun = .false.
      nc = 1
    end if
    if (nYn > 2) then
      h_jj1 = 2
      h_jj2 = nYn - 1
      yrun = .true.
      mc = nXn
    else
      h_jj1 = 1
      h_jj2 = 1
      yrun = .false.
      mc = 1
    end if
    h_mx = max(nc, mc)
    allocate(edge1(h_mx, 3), edge2(h_mx, 3))
    edge1 = 0
    edge2 = 0
  end subroutine ref_init

  subroutine ref_free() bind(C, name="ref_free")
    if (allocated(dep)) deallocate(dep)
    if (allocated(cel)) deallocate(cel)
    if (allocated(xvel)) deallocate(xvel)
    if (allocated(yvel)) deallocate(yvel)
    if (allocated(hmax)) deallocate(hmax)
    if (allocated(velmax)) deallocate(velmax)
    if (allocated(s1)) deallocate(s1)
    if (allocated(s2)) deallocate(s2)
    if (allocated(zeta)) deallocate(zeta)
    if (allocated(edge1)) deallocate(edge1)
    if (allocated(edge2)) deallocate(edge2)
  end subroutine ref_free

  real(c_double) function ref_grav() bind(C, name="ref_grav")
    ref_grav = grav
  end function ref_grav

  real(c_double) function ref_celmin() bind(C, name="ref_celmin")
    ref_celmin = celmin
  end function ref_celmin

  integer(c_int) function ref_mx() bind(C, name="ref_mx")
    ref_mx = h_mx
  end function ref_mx

  subroutine ref_set_state(c_in, u_in, v_in) bind(C, name="ref_set_state")
    real(c_double), intent(in) :: c_in(nXn*nYn), u_in(nXn*nYn), v_in(nXn*nYn)
    cel = reshape(c_in, (/nXn, nYn/))
    xvel = reshape(u_in, (/nXn, nYn/))
    yvel = reshape(v_in, (/nXn, nYn/))
  end subroutine ref_set_state

  subroutine ref_get_state(c_out, u_out, v_out, hm_out, vm_out) bind(C, name="ref_get_state")
    real(c_double), intent(out) :: c_out(nXn*nYn), u_out(nXn*nYn), v_out(nXn*nYn)
    real(c_double), intent(out) :: hm_out(nXn*nYn), vm_out(nXn*nYn)
    c_out = reshape(cel, (/nXn*nYn/))
    u_out = reshape(xvel, (/nXn*nYn/))
    v_out = reshape(yvel, (/nXn*nYn/))
    hm_out = reshape(hmax, (/nXn*nYn/))
    vm_out = reshape(velmax, (/nXn*nYn/))
  end subroutine ref_get_state

  ! edge arrays as the reference holds them: edge1/edge2(mx,3), column-major.
  subroutine ref_set_edges(e1, e2) bind(C, name="ref_set_edges")
    real(c_double), intent(in) :: e1(h_mx*3), e2(h_mx*3)
    edge1 = reshape(e1, (/h_mx, 3/))
    edge2 = reshape(e2, (/h_mx, 3/))
  end subroutine ref_set_edges

  ! one line, exactly the reference call
  subroutine ref_line(iu, irw) bind(C, name="ref_line")
    integer(c_int), value :: iu, irw
    integer :: iu_f, irw_f, n
    iu_f = iu
    irw_f = irw
    if (iu == 1) then
      n = nXn
    else
      n = nYn
    end if
    call cliffs(iu_f, irw_f, n)
  end subroutine ref_line

  ! one directional sweep over lines first..last (cliffs_main.f:325-329, 340-344)
  subroutine ref_sweep(iu, first, last) bind(C, name="ref_sweep")
    integer(c_int), value :: iu, first, last
    integer :: j
    if (iu == 1) then
!$OMP PARALLEL DO
      do j = first, last
        call cliffs(1, j, nXn)
      end do
!$OMP END PARALLEL DO
    else
!$OMP PARALLEL DO
      do j = first, last
        call cliffs(2, j, nYn)
      end do
!$OMP END PARALLEL DO
    end if
  end subroutine ref_sweep

  ! max-wave update (cliffs_main.f:347-355)
  subroutine ref_maxwave() bind(C, name="ref_maxwave")
    integer :: i, j
    real(c_double) :: aa
    do j = 1, nYn
      do i = 1, nXn
        aa = cel(i, j)
        if (aa > hmax(i, j)) hmax(i, j) = aa
        aa = xvel(i, j)**2 + yvel(i, j)**2
        if (aa > velmax(i, j)) velmax(i, j) = aa
      end do
    end do
  end subroutine ref_maxwave

  ! nsteps full steps with fixed edges (cliffs_main.f:317-355, no feed)
  subroutine ref_steps(nsteps) bind(C, name="ref_steps")
    integer(c_int), value :: nsteps
    integer :: istep
    do istep = 1, nsteps
      if (xrun) call ref_sweep(1, h_jj1, h_jj2)
      if (yrun) call ref_sweep(2, h_ii1, h_ii2)
      call ref_maxwave()
    end do
  end subroutine ref_steps

end module ref_harness
