! Host-only check of the drop-in's files: measurement bins and the reference's
! output files, the text field format and the checkpoint writer, run on given
! data without opening a device chain.
!
! usage: pqmc_iocheck <input file> <output folder, ending in '/'>
! Input (stream): int32 L1 L2 Lt ngroup ne nbins nmeas meas_number nsamples
! seed; float64 fields(Nb, Lt); then nsamples records of int32 forward,
! int32 time, complex128 sample(total_width).
! Writes into the folder: temp/ and data/ files of the measurement system,
! ph_field_core1.dat (the fields as text), ph_field_read.bin (that text read
! back, float64), ckpt.bin (a checkpoint after the samples, loop 6, ran2
! seeded and drawn 7 times, counters 11 22 3 4, err_fast_max 1.5e-7).
program pqmc_iocheck
  use iso_c_binding
  use pqmc_hip
  use pqmc_dropin
  implicit none
  character(len=1024) :: fin, dout
  integer(c_int32_t) :: hdr(10), rec(2)
  real(c_double), allocatable :: x(:, :), y(:, :)
  complex(c_double_complex), allocatable :: smp(:)
  real(c_float) :: r(7)
  integer :: u_in, u, k

  call get_command_argument(1, fin)
  call get_command_argument(2, dout)
  open (newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u_in) hdr
  lat_L1 = hdr(1)
  ncell = hdr(1) * hdr(2)
  nb = 2 * ncell
  ntime = hdr(3)
  ngroup = hdr(4)
  nelec = hdr(5)
  total_width = 3 + 38 * ncell
  allocate (x(nb, ntime), y(nb, ntime), smp(total_width), meas_sum(total_width))
  read (u_in) x
  call meas_init(int(hdr(6)), int(hdr(7)), int(hdr(8)), trim(dout), 0, 1, 1, total_width)
  meas_sum = (0.0_c_double, 0.0_c_double)
  n_meas = 0
  do k = 1, hdr(9)
    read (u_in) rec, smp
    call record_measurement(rec(1) /= 0, int(rec(2)), smp)
    meas_sum = meas_sum + smp
    n_meas = n_meas + 1
  end do
  close (u_in)
  call data_analyse()

  call write_ph_field(trim(dout)//'ph_field_core1.dat', x)
  call read_ph_field(trim(dout)//'ph_field_core1.dat', y)
  open (newunit=u, file=trim(dout)//'ph_field_read.bin', status='replace', access='stream', form='unformatted')
  write (u) y
  close (u)

  call pqmc_ran2_seed(rng_state, hdr(10))
  call pqmc_ran2_fill(rng_state, 7_c_int64_t, r)
  global_accept = 3
  global_reject = 4
  call write_checkpoint(trim(dout)//'ckpt.bin', 6, .true., 11_c_int64_t, 22_c_int64_t, 1.5e-7_c_double, x)
end program pqmc_iocheck
