! Host-only check of the drop-in's global-move proposals: runs
! generate_newfield_global (src/mod_update.f90:292-467) on one field
! configuration without opening a device chain.
!
! usage: pqmc_propose <input file> <output file>
! Input (stream): int32 L1 L2 Lt kind seed (kind 1 kspace, 2 rotate,
! 3 kspace_time); float64 M D biased dtau global_update_distance (times
! char_length); float64 fields(Nb, Lt).
! Output: float64 proposed fields(Nb, Lt); float64 -delt*dE.
program pqmc_propose
  use iso_c_binding
  use pqmc_hip
  use pqmc_dropin
  implicit none
  character(len=1024) :: fin, fout
  character(len=16), parameter :: kinds(3) = [character(len=16) :: 'kspace', 'rotate', 'kspace_time']
  integer(c_int32_t) :: hdr(5)
  real(c_double) :: reals(5), ph
  integer :: u_in, u_out

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u_in) hdr, reals
  lat_L1 = hdr(1)
  ncell = hdr(1) * hdr(2)
  nb = 2 * ncell
  ntime = hdr(3)
  pm_M = reals(1); pm_D = reals(2); pm_biased = reals(3); pm_dtau = reals(4)
  global_update_distance = reals(5)
  allocate (bf_cur(nb, ntime), bf_new(nb, ntime))
  read (u_in) bf_cur
  close (u_in)
  call pqmc_ran2_seed(rng_state, hdr(5))
  bf_new = bf_cur
  call generate_newfield_global(trim(kinds(hdr(4))), ph)
  open (newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u_out) bf_new, ph
  close (u_out)
end program pqmc_propose
