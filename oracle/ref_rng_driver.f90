! Test-only driver for the reference's ran2 generator (compiled from the
! reference's own src/mod_nrtype.f90, mod_nrutil.f90, mod_ranstate.f90 by
! oracle/Makefile.ref).  Usage: ref_rng <seed> <n>
! Prints n lines "<rands bits> <ran_sysm bits>" for two interleaved draws per
! line, as 32-bit integers (bit patterns of REAL(SP)), after
! init_rng_by_seed(seed) (src/mod_ranstate.f90:267-283).
program ref_rng
  use ran_state
  implicit none
  integer :: seed, n, i
  character(len=32) :: arg
  real(sp) :: a, b
  call get_command_argument(1, arg)
  read (arg, *) seed
  call get_command_argument(2, arg)
  read (arg, *) n
  call init_rng_by_seed(seed)
  do i = 1, n
    a = rands()
    b = ran_sysm()
    write (*, '(I0,1X,I0)') transfer(a, 0), transfer(b, 0)
  end do
end program ref_rng
