! Sweep driver over the drop-in subroutines: the loop body of
! src/Main_PQMC.f90:81-156, called one reference subroutine at a time
! (update_global before a backward sweep when mod(loop + 1,
! global_update_loop) == 0, update_phonon_field, right_/left_evolve on g_h,
! get_g_scratch every ngroup slices, begin_measure around ntime/2 after the
! warm-up, src/Main_PQMC.f90:143-148).
!
! usage: pqmc_driver <input file> <output file> [<output folder>/ [<gradient file>]]
! The output folder (with temp/ and data/ in it) receives the measurement
! bins and files when nbins > 0, and checkpoint.bin when a checkpoint is
! saved or resumed from.
! Input (stream, little-endian): int32 L1 L2 Lt ngroup ne block_bonds nsweeps
! seed; int32 sites[4][npla][4], bfi[4][npla], cross[4][npla]; complex128
! bcph[4][npla][16], kcoe[4], vcoe[4], k12[4][4], v12[4][4], bc[2]; float64
! dtau M D biased jump_distance hop lam; int32 ncopy tr_weight; float64
! fields(Nb, Lt); complex128 slater(Ns, ne); int32 global_update_loop (0 =
! global update off); float64 global_update_distance (times char_length);
! int32 nbins (0 = no bins or files), nmeas_per_bin, meas_number, warmup,
! meas_interval, checkpoint loop (save after it, 0 = none), resume (1 =
! start from checkpoint.bin), proj (1 = projector; 0 = finite temperature,
! src/input.f90 proj; the trial matrix is then read but not used).
! Output: int8 accept flags of every update call in call order
! (nsweeps * Lt * 4 * npla); complex128 G(Ns, Ns) at the end; float64
! fields(Nb, Lt); complex128 summed measurements (total_width); int32 number
! of measurements; int64 positive and negative accepts; int64 global
! accepts and rejects.
! Gradient file (optional): complex128 d ln_cw / dx (Nb, Lt) of the final
! fields, from init_g(1) and one gradient pass after the output is written.
program pqmc_driver
  use iso_c_binding
  use pqmc_hip
  use pqmc_dropin
  implicit none
  character(len=1024) :: fin, fout
  integer :: u_in, u_out
  integer(c_int32_t) :: hdr(8), ncopy_tr(2), glob_loop, mc(8)
  character(len=1024) :: outdir
  character(len=1024) :: fgrad
  complex(c_double_complex), allocatable :: dlnw(:)
  integer :: nbins, nmeas, mnum, warmup, meas_interval, ckpt_loop, loop0
  logical :: resume
  real(c_double) :: glob_dist
  integer :: L1, L2, Lt, ng, ne, nbb, nsweeps, seed, npla_, ns_, nb_
  integer(c_int32_t), allocatable, target :: sites(:), bfi(:), cross(:)
  complex(c_double_complex), allocatable, target :: bcph(:), kcoe(:), vcoe(:), k12(:), v12(:), bc(:)
  real(c_double) :: reals(7)
  real(c_double), allocatable :: fields(:)
  complex(c_double_complex), allocatable :: slater(:)
  type(pqmc_config) :: cfg
  type(pqmc_model) :: model
  type(pf_type) :: ppf_list(4)
  integer(c_int8_t), allocatable :: acc_all(:)
  integer(c_int64_t) :: cnt(3)
  real(c_double) :: efm
  logical :: forward
  integer :: loop, time_count, time, i_pf, k, pos

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u_in) hdr
  L1 = hdr(1); L2 = hdr(2); Lt = hdr(3); ng = hdr(4); ne = hdr(5); nbb = hdr(6); nsweeps = hdr(7); seed = hdr(8)
  ns_ = 2 * L1 * L2
  nb_ = ns_
  npla_ = ns_ / 4
  allocate (sites(4 * npla_ * 4), bfi(4 * npla_), cross(4 * npla_), bcph(4 * npla_ * 16), kcoe(4), vcoe(4), &
            k12(16), v12(16), bc(2), fields(nb_ * Lt), slater(ns_ * ne))
  read (u_in) sites, bfi, cross, bcph, kcoe, vcoe, k12, v12, bc, reals, ncopy_tr, fields, slater, glob_loop, glob_dist, mc
  nbins = mc(1); nmeas = mc(2); mnum = mc(3); warmup = mc(4); meas_interval = mc(5); ckpt_loop = mc(6)
  resume = mc(7) /= 0
  outdir = ''
  if (command_argument_count() >= 3) call get_command_argument(3, outdir)
  if ((nbins > 0 .or. ckpt_loop > 0 .or. resume) .and. len_trim(outdir) == 0) &
    error stop 'pqmc_driver: bins and checkpoints need an output folder'
  close (u_in)

  cfg = pqmc_config(L1, L2, Lt, ng, ne, nbb)
  if (mc(8) == 0) cfg%finite_temperature = 1
  model%sites = c_loc(sites); model%bfi = c_loc(bfi); model%cross = c_loc(cross); model%bcph = c_loc(bcph)
  model%kcoe = c_loc(kcoe); model%vcoe = c_loc(vcoe); model%k12 = c_loc(k12); model%v12 = c_loc(v12)
  model%bc = c_loc(bc)
  model%dtau = reals(1); model%M = reals(2); model%D = reals(3); model%biased = reals(4)
  model%jump_distance = reals(5); model%hop = reals(6); model%lam = reals(7)
  model%ncopy = ncopy_tr(1); model%tr_weight = ncopy_tr(2)
  call pqmc_open(cfg, model, fields, slater, seed, glob_dist)
  do k = 1, 4
    ppf_list(k)%id = k
  end do
  if (nbins > 0) call meas_init(nbins, nmeas, mnum, trim(outdir), 0, 1, 1)
  allocate (acc_all(nsweeps * Lt * 4 * n_pla))
  acc_all = 0
  loop0 = 0
  forward = .true.
  if (resume) then
    call checkpoint_load(trim(outdir)//'checkpoint.bin', loop0, forward)
  else
    call init_g(1)
  end if
  pos = loop0 * Lt * 4 * n_pla
  do loop = loop0 + 1, nsweeps
    forward = .not. forward  ! the first loop is going backwards
    if ((.not. forward) .and. glob_loop > 0) then
      if (mod(loop + 1, int(glob_loop)) == 0) then
        updated = .false.
        call update_global("kspace")
        call update_global("rotate")
        call update_global("kspace_time")
        if (updated) call init_g(1)
      end if
    end if
    do time_count = 1, ntime
      if (.not. forward) then
        time = ntime - time_count + 1
        do i_pf = 4, 1, -1
          call right_evolve(time, g_h, ns, i_pf, two_way=.true.)
          call update_phonon_field(ppf_list(i_pf), time)
          acc_all(pos + 1:pos + n_pla) = last_accepts
          pos = pos + n_pla
        end do
      else
        time = time_count
        do i_pf = 1, 4
          call update_phonon_field(ppf_list(i_pf), time)
          acc_all(pos + 1:pos + n_pla) = last_accepts
          pos = pos + n_pla
          call left_evolve(time, g_h, ns, i_pf, two_way=.true.)
        end do
      end if
      if (mod(time_count, ngroup) == 0) then
        if (forward) then
          call get_g_scratch(time + 1, forward, loop)
        else
          call get_g_scratch(time, forward, loop)
        end if
      end if
      if (loop > warmup .and. mod(loop - warmup, meas_interval) == 0) then
        if (time >= ntime / 2 - meas_number .and. time <= ntime / 2 + meas_number) call begin_measure(forward, time)
      end if
    end do
    if (loop == ckpt_loop) call checkpoint_save(trim(outdir)//'checkpoint.bin', loop, forward)
  end do
  if (nbins > 0) call data_analyse()

  call sync_g_h()
  call pqmc_check(pqmc_get_fields(chain, fields), 'pqmc_get_fields')
  call pqmc_check(pqmc_get_counters(chain, cnt, efm), 'pqmc_get_counters')
  open (newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u_out) acc_all, g_h, fields, meas_sum, int(n_meas, c_int32_t), cnt(1), cnt(2), global_accept, global_reject
  close (u_out)
  if (command_argument_count() >= 4) then
    call get_command_argument(4, fgrad)
    allocate (dlnw(nb_ * Lt))
    call pqmc_check(pqmc_init_g(chain, 1_c_int), 'pqmc_init_g')
    call pqmc_check(pqmc_gradient(chain, dlnw), 'pqmc_gradient')
    open (newunit=u_out, file=trim(fgrad), access='stream', form='unformatted', status='replace')
    write (u_out) dlnw
    close (u_out)
  end if
  call pqmc_close()
end program pqmc_driver
