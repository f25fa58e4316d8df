! Test driver for the Fortran boundary (cliffs-src_amd/fortran/).
!
! Declares the module variables a Cliffs driver provides (the names of
! global_modules.f:16-55 that the drop-in reads) and runs the time loop of
! cliffs_main.f:303-355 through the drop-in:
!   - the first `nline` steps with the reference's own loop form,
!       do j = jj1, jj2; call cliffs(1, j, nXn); end do  (serial)
!   - the remaining steps with one cliffs_gpu_sweep per direction,
! then writes the state and maxima.  The Python test compares them with the
! CPU oracle.
!
! usage: fdriver <input.bin> <output.bin>
program fdriver
  use MASTER
  use PARAMETERS
  use cliffs_dropin
  implicit none
  character(len=512) :: fin, fout
  integer :: nsteps, nline, istep, j, ii1, ii2, jj1, jj2, mx, u
  integer(4) :: hdr(4)
  real*8 :: par(3)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nXn = hdr(1)
  nYn = hdr(2)
  nsteps = hdr(3)
  nline = hdr(4)
  read(u) par
  dt = par(1)
  ground = par(2)
  crough = par(3)
  allocate(dep(nXn, nYn), cel(nXn, nYn), xvel(nXn, nYn), yvel(nXn, nYn))
  allocate(hmax(nXn, nYn), velmax(nXn, nYn))
  read(u) dep, cel, xvel, yvel
  if (nXn > 2) then
    allocate(s1(nXn - 1, nYn))
    read(u) s1
  end if
  if (nYn > 2) then
    allocate(s2(nYn - 1), zeta(nYn))
    read(u) s2, zeta
  end if
  close(u)

  ! line ranges (cliffs_main.f:140-164)
  ii1 = 1; ii2 = 1; jj1 = 1; jj2 = 1
  if (nXn > 2) then
    ii1 = 2; ii2 = nXn - 1
  end if
  if (nYn > 2) then
    jj1 = 2; jj2 = nYn - 1
  end if
  mx = max(merge(nYn, 1, nXn > 2), merge(nXn, 1, nYn > 2))
  allocate(edge1(mx, 3), edge2(mx, 3))
  edge1 = 0
  edge2 = 0
  hmax = 0
  velmax = 0

  call cliffs_gpu_attach(0)
  do istep = 1, nsteps
    if (istep <= nline) then
      if (nXn > 2) then
        do j = jj1, jj2
          call cliffs(1, j, nXn)
        end do
      end if
      if (nYn > 2) then
        do j = ii1, ii2
          call cliffs(2, j, nYn)
        end do
      end if
    else
      if (nXn > 2) call cliffs_gpu_sweep(1, jj1, jj2)
      if (nYn > 2) call cliffs_gpu_sweep(2, ii1, ii2)
    end if
    call cliffs_gpu_maxwave()
  end do
  call cliffs_gpu_to_host()
  call cliffs_gpu_detach()

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) cel, xvel, yvel, hmax, velmax
  close(u)
end program fdriver
