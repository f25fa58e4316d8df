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
  real*8 ground, crough, dt
end module PARAMETERS

