! Drop-in replacements for the reference's hot-path subroutines, with the
! reference's names and arguments, backed by one GPU-resident chain:
!   update_phonon_field(ppf, time)                (src/mod_update.f90:13-201)
!   left_evolve(time, mat, n, pf_id, two_way)      (src/mod_evolution.f90:236-285)
!   right_evolve(time, mat, n, pf_id, two_way)     (src/mod_evolution.f90:287-328)
!   init_g(start_time)                             (init_g_T0, :46-111)
!   get_g_scratch(time, forward, loop)             (get_g_scratch_T0, :113-195)
!   begin_measure(forward, time) / take_measurement(time)  (src/mod_meas.f90:210-544,729-763)
!   update_global(update_type)                     (src/mod_update.f90:251-290)
!   meas_init / data_analyse: bins and output files (src/mod_meas.f90:577-583,
!   729-763, 850-1017); import_ph_field / export_ph_field: the text field
!   format of src/Main_PQMC.f90:355-368; checkpoint_save / checkpoint_load
! update_global generates the proposal on the host (generate_newfield_global,
! :292-467) and weighs it on the device (cal_new_weight_T0,
! src/mod_evolution.f90:359-408).
! g_h is a host mirror of the device Green's function: passing g_h itself
! as `mat` with two_way = .true. wraps the device copy without any transfer;
! sync_g_h() copies the device G into it on demand.  Any other matrix is
! copied to the device, evolved and copied back.
module pqmc_dropin
  use iso_c_binding
  use pqmc_hip
  ! The bond-family type.  In a reference build (-DPQMC_REFERENCE_TYPES) it
  ! is the reference's own pf_type (src/input.f90:121-137), so the caller's
  ! ppf_list passes straight through; the drop-in reads only its id.  The
  ! stand-alone driver uses the minimal type below instead.
#ifdef PQMC_REFERENCE_TYPES
  use input, only: pf_type
#endif
  implicit none

#ifndef PQMC_REFERENCE_TYPES
  type :: pf_type
    integer :: id = 0
    logical :: V_exist = .true.
  end type pf_type
#endif

  type(c_ptr), save :: chain = c_null_ptr
  type(pqmc_ran2), save :: rng_state
  integer, save :: ns = 0, ncell = 0, nb = 0, n_pla = 0, nblock = 0, total_width = 0
  integer, save :: ntime = 0, ngroup = 0
  complex(c_double_complex), allocatable, target, save :: g_h(:, :)
  complex(c_double_complex), allocatable, save :: meas_sum(:)
  integer, save :: n_meas = 0
  integer(c_int8_t), allocatable, save :: last_accepts(:)
  real(c_double), save :: err_fast = 0.0_c_double

  ! global update (src/input.f90:73,184, src/mod_update.f90:251-467)
  real(c_float), parameter :: pi_sp = 3.141592653589793238462643383279502884197_c_float
  real(c_double), save :: global_update_distance = 0.0_c_double  ! already times char_length
  real(c_double), save :: pm_M = 0, pm_D = 0, pm_biased = 0, pm_dtau = 0
  integer, save :: lat_L1 = 0, nelec = 0
  logical, save :: updated = .false.
  integer(c_int64_t), save :: global_accept = 0, global_reject = 0
  ! ln_cw after an accepted global move: the device value is stale until the
  ! next init_g, which the caller runs when updated is set
  logical, save :: lncw_held = .false.
  complex(c_double_complex), save :: lncw_hold = (0.0_c_double, 0.0_c_double)
  real(c_double), allocatable, save :: bf_cur(:, :), bf_new(:, :)

  ! measurement bins (src/mod_meas.f90:577-583, 729-763) for the observables
  ! of input/model/EPSOCZ.nml:262-283; field kinds get a _k twin
  integer, parameter :: KIND_SCALAR = 1, KIND_FIELD = 2
  character(len=12), parameter :: obs_scalars(3) = [character(len=12) :: 'BF_KE', 'BF_PE', 'BF_X']
  character(len=12), parameter :: obs_fields(19) = [character(len=12) :: &
    'BFx_BFx', 'BFx_BFy', 'BFy_BFy', 'BFxy_BFxy', 'Jxx_Jxx', 'Jxy_Jxy', 'Jyx_Jyx', 'Jyy_Jyy', &
    'Jxx_Jyx', 'Jxy_Jyy', 'Jxx_Jyy', 'Jxy_Jyx', 'cJxy_cJxy', 'cJyx_cJyx', 'den_den', 'SC_SC', &
    'Sz_Sz', 'Sx_Sx', 'Sy_Sy']
  logical, save :: bins_on = .false.
  integer, save :: nbin_per_core = 0, nmeas_per_bin = 0, meas_number = 0, cur_bin = 1, cur_meas = 1, myid = 0
  character(len=512), save :: output_addr = ''
  character(len=20), save :: output_format = 'f18.8'
  complex(c_double_complex), allocatable, save :: meas_acc(:)
  real(c_double), allocatable, save :: bin_sum(:), bin_mean(:, :)

contains

  ! Allocate the chain and load the model, fields (boson_field(Nb, Lt)), trial
  ! Slater matrix (Ns x ne) and the ran2 seed (init_rng_by_seed,
  ! src/mod_ranstate.f90:267-283).  global_distance is global_update_distance
  ! after its scaling by char_length (src/input.f90:184); without it the
  ! reference's default 1.0 times char_length = 1/(M omega), omega = sqrt(D/M),
  ! or sqrt(1/(D lam)) for M < 0.001 (src/input.f90:176-182).
  subroutine pqmc_open(cfg, model, fields, slater, seed, global_distance)
    type(pqmc_config), intent(in) :: cfg
    type(pqmc_model), intent(in) :: model
    real(c_double), intent(in) :: fields(*)
    complex(c_double_complex), intent(in) :: slater(*)
    integer, intent(in) :: seed
    real(c_double), intent(in), optional :: global_distance
    integer(c_int32_t) :: dims(6)
    call pqmc_check(pqmc_create(cfg, chain), 'pqmc_create')
    call pqmc_check(pqmc_set_model(chain, model), 'pqmc_set_model')
    call pqmc_check(pqmc_get_dims(chain, dims), 'pqmc_get_dims')
    ns = dims(1); ncell = dims(2); nb = dims(3); n_pla = dims(4); nblock = dims(5); total_width = dims(6)
    ntime = cfg%Lt
    ngroup = cfg%ngroup
    call pqmc_check(pqmc_set_fields(chain, fields), 'pqmc_set_fields')
    call pqmc_check(pqmc_set_trial(chain, slater), 'pqmc_set_trial')
    call pqmc_ran2_seed(rng_state, int(seed, c_int32_t))
    allocate (g_h(ns, ns), meas_sum(total_width), last_accepts(n_pla), bf_cur(nb, ntime), bf_new(nb, ntime))
    pm_M = model%M; pm_D = model%D; pm_biased = model%biased; pm_dtau = model%dtau
    lat_L1 = cfg%L1
    nelec = cfg%ne
    if (present(global_distance)) then
      global_update_distance = global_distance
    else if (pm_M < 0.001_c_double) then
      global_update_distance = sqrt(1.0_c_double / (pm_D * model%lam))
    else
      global_update_distance = 1.0_c_double / (pm_M * sqrt(pm_D / pm_M))
    end if
    global_accept = 0; global_reject = 0
    updated = .false.; lncw_held = .false.
    g_h = (0.0_c_double, 0.0_c_double)
    meas_sum = (0.0_c_double, 0.0_c_double)
    n_meas = 0
  end subroutine pqmc_open

  subroutine pqmc_close()
    integer(c_int) :: rc
    if (c_associated(chain)) rc = pqmc_destroy(chain)
    chain = c_null_ptr
    if (allocated(g_h)) deallocate (g_h, meas_sum, last_accepts, bf_cur, bf_new)
  end subroutine pqmc_close

  subroutine init_g(start_time)
    integer, intent(in) :: start_time
    call pqmc_check(pqmc_init_g(chain, int(start_time, c_int)), 'init_g')
    lncw_held = .false.
  end subroutine init_g

  ! The draws are taken here, in the reference's call order: for every bond
  ! of the family, ran_sysm() (from one ran2 draw) then rands().
  subroutine update_phonon_field(ppf, time)
    type(pf_type), intent(in) :: ppf
    integer, intent(in) :: time
    real(c_float) :: u(2 * n_pla)
    call pqmc_ran2_fill(rng_state, int(2 * n_pla, c_int64_t), u)
    call pqmc_check(pqmc_update_family(chain, int(ppf%id, c_int), int(time, c_int), u, last_accepts), &
                    'update_phonon_field')
  end subroutine update_phonon_field

  ! mat <- E mat (E^-1 on the right with two_way); mat(ns, n)
  subroutine left_evolve(time, mat, n, pf_id, two_way)
    integer, intent(in) :: time, n, pf_id
    complex(c_double_complex), intent(inout), target :: mat(ns, n)
    logical, intent(in) :: two_way
    if (two_way .and. n /= ns) error stop 'error in left_evolve, dimension not for evolve two-way'
    if (two_way .and. c_associated(c_loc(mat), c_loc(g_h))) then
      call pqmc_check(pqmc_evolve_g(chain, int(time, c_int), int(pf_id, c_int), 1_c_int), 'left_evolve')
    else
      call pqmc_check(pqmc_evolve_matrix(chain, int(time, c_int), int(pf_id, c_int), 1_c_int, &
                                         merge(1_c_int, 0_c_int, two_way), mat, int(n, c_int)), 'left_evolve')
    end if
  end subroutine left_evolve

  ! mat <- mat E (E^-1 on the left with two_way); mat(n, ns)
  subroutine right_evolve(time, mat, n, pf_id, two_way)
    integer, intent(in) :: time, n, pf_id
    complex(c_double_complex), intent(inout), target :: mat(n, ns)
    logical, intent(in) :: two_way
    if (two_way .and. n /= ns) error stop 'error in right_evolve, dimension not for evolve two-way'
    if (two_way .and. c_associated(c_loc(mat), c_loc(g_h))) then
      call pqmc_check(pqmc_evolve_g(chain, int(time, c_int), int(pf_id, c_int), 0_c_int), 'right_evolve')
    else
      call pqmc_check(pqmc_evolve_matrix(chain, int(time, c_int), int(pf_id, c_int), 0_c_int, &
                                         merge(1_c_int, 0_c_int, two_way), mat, int(n, c_int)), 'right_evolve')
    end if
  end subroutine right_evolve

  subroutine get_g_scratch(time, forward, loop)
    integer, intent(in) :: time, loop
    logical, intent(in) :: forward
    call pqmc_check(pqmc_g_scratch(chain, int(time, c_int), merge(1_c_int, 0_c_int, forward), err_fast), &
                    'get_g_scratch')
  end subroutine get_g_scratch

  ! meas_number = 0: one measurement at the requested time
  ! With meas_init called, the measurement also goes into the bins.
  subroutine begin_measure(forward, time)
    logical, intent(in) :: forward
    integer, intent(in) :: time
    complex(c_double_complex) :: out(total_width)
    if (bins_on .and. cur_bin > nbin_per_core) error stop 'begin_measure: no more bins left'
    call pqmc_check(pqmc_measure(chain, int(time, c_int), out), 'take_measurement')
    meas_sum = meas_sum + out
    n_meas = n_meas + 1
    if (bins_on) call record_measurement(forward, time, out)
  end subroutine begin_measure

  subroutine take_measurement(time)
    integer, intent(in) :: time
    complex(c_double_complex) :: out(total_width)
    call pqmc_check(pqmc_measure(chain, int(time, c_int), out), 'take_measurement')
    meas_sum = meas_sum + out
    n_meas = n_meas + 1
  end subroutine take_measurement

  subroutine sync_g_h()
    call pqmc_check(pqmc_get_g(chain, g_h), 'sync_g_h')
  end subroutine sync_g_h

  ! rands, irands and ran_sysm of src/mod_ranstate.f90:193-226: one ran2
  ! draw each, arithmetic in REAL(SP)
  real(c_float) function rands()
    real(c_float) :: r(1)
    call pqmc_ran2_fill(rng_state, 1_c_int64_t, r)
    rands = r(1)
  end function rands

  integer function irands(n)
    integer, intent(in) :: n
    irands = int(rands() * real(n, c_float))
  end function irands

  real(c_float) function ran_sysm()
    ran_sysm = rands() * 2.0_c_float - 1.0_c_float
  end function ran_sysm

  ! cal_prob (src/mod_update.f90:475-500), Metropolis acceptance
  real(c_double) function cal_prob(w)
    complex(c_double_complex), intent(in) :: w
    cal_prob = min(1.0_c_double, exp(min(10.0_c_double, real(w, c_double))))
  end function cal_prob

  ! update_global(update_type) for the kinds the sweep loop calls: "kspace",
  ! "rotate" and "kspace_time".  An accepted move writes its fields to the
  ! device and sets updated; the Green's function and B-strings follow at the
  ! caller's init_g(1), as in src/Main_PQMC.f90:87-95.
  subroutine update_global(update_type)
    character(len=*), intent(in) :: update_type
    complex(c_double_complex) :: w_old, w_new
    real(c_double) :: ri(2), ph, prob
    if (lncw_held) then
      w_old = lncw_hold
    else
      call pqmc_check(pqmc_get_lncw(chain, ri), 'update_global')
      w_old = cmplx(ri(1), ri(2), c_double_complex)
    end if
    call pqmc_check(pqmc_get_fields(chain, bf_cur), 'update_global')
    bf_new = bf_cur
    call generate_newfield_global(update_type, ph)
    call pqmc_check(pqmc_global_weight(chain, bf_new, ri), 'update_global')
    w_new = cmplx(ri(1), ri(2), c_double_complex)
    prob = cal_prob(w_new + ph - w_old)
    if (rands() < prob) then
      global_accept = global_accept + 1
      updated = .true.
      call pqmc_check(pqmc_set_fields(chain, bf_new), 'update_global')
      lncw_hold = w_new
      lncw_held = .true.
    else
      global_reject = global_reject + 1
    end if
  end subroutine update_global

  ! generate_newfield_global (src/mod_update.f90:292-439): bf_new from
  ! bf_cur, ph = -delt * dE.  For "kspace_time" the window [start, end] is
  ! cut at ntime where the reference would index past it.
  subroutine generate_newfield_global(update_type, ph)
    character(len=*), intent(in) :: update_type
    real(c_double), intent(out) :: ph
    real(c_double) :: incell(2), bj(nb), de, kin, ang, r11, r12, r21, r22, a, b
    integer :: i, t, t0, t1, tb, ta, c
    character(len=len(update_type)) :: kind
    kind = adjustl(update_type)
    de = 0.0_c_double
    select case (trim(kind))
    case ('kspace', 'kspace_time')
      do i = 1, 2
        incell(i) = irands(2) * 2.0_c_double - 1.0_c_double
      end do
      call generate_newfield_space(bj, incell)
      t0 = 1
      t1 = ntime
      if (trim(kind) == 'kspace_time') then
        t0 = irands(ntime / 2) + 1
        t1 = min(max(t0 + int(1.0_c_double * pm_M / (pm_dtau * pm_dtau)), ntime), ntime)
      end if
      do t = t0, t1
        bf_new(:, t) = bf_cur(:, t) + bj
      end do
      do t = t0, t1
        do i = 1, nb
          de = de + 0.5_c_double * pm_D * (bf_new(i, t) * bf_new(i, t) - bf_cur(i, t) * bf_cur(i, t))
          de = de + (-pm_biased * (bf_new(i, t) - bf_cur(i, t)))
        end do
      end do
      if (trim(kind) == 'kspace_time') then
        kin = 0.5_c_double * (pm_M / (pm_dtau * pm_dtau))
        tb = modulo(t0 - 2, ntime) + 1   ! modi(t0 - 1, ntime)
        ta = modulo(t1, ntime) + 1       ! modi(t1 + 1, ntime)
        do i = 1, nb
          de = de + kin * ((bf_new(i, t0) - bf_new(i, tb))**2 - (bf_cur(i, t0) - bf_cur(i, tb))**2)
          de = de + kin * ((bf_new(i, t1) - bf_new(i, ta))**2 - (bf_cur(i, t1) - bf_cur(i, ta))**2)
        end do
      end if
    case ('rotate')
      ang = real(ran_sysm(), c_double) * 0.1_c_double * real(pi_sp, c_double)
      r11 = cos(ang); r12 = -sin(ang); r21 = sin(ang); r22 = cos(ang)
      if (ran_sysm() < 0.0_c_float) then
        r12 = -r12
        r22 = -r22
      end if
      do t = 1, ntime
        do c = 1, ncell
          a = bf_cur(2 * c - 1, t)
          b = bf_cur(2 * c, t)
          bf_new(2 * c - 1, t) = r11 * a + r12 * b
          bf_new(2 * c, t) = r21 * a + r22 * b
        end do
      end do
    case default
      error stop 'update_global: only kspace, rotate and kspace_time are on the sweep path'
    end select
    ph = -de * pm_dtau
  end subroutine generate_newfield_global

  ! generate_newfield_space (src/mod_update.f90:441-467) with
  ! K_vec = 2*PI*[0.5d0, 0.5d0] (PI in REAL(SP)), phi = 0 and rpos the
  ! 0-based cell position (lat_vec = identity)
  subroutine generate_newfield_space(bj, incell)
    real(c_double), intent(out) :: bj(nb)
    real(c_double), intent(in) :: incell(2)
    real(c_double) :: dist, amp, kv, cs
    integer :: c, i
    dist = 2 * global_update_distance / ncell
    amp = dist * real(rands(), c_double)
    kv = real(2 * pi_sp, c_double) * 0.5_c_double
    do c = 1, ncell
      cs = cos((kv * real(mod(c - 1, lat_L1), c_double) + kv * real((c - 1) / lat_L1, c_double)) + 0.0_c_double)
      do i = 1, 2
        bj(2 * (c - 1) + i) = incell(i) * amp * cs
      end do
    end do
  end subroutine generate_newfield_space

  ! ------------------------------------------------------------------
  ! measurement bins and output files

  ! init_meas + init_storage (src/mod_meas.f90:577-583, 709-718): nbins bins
  ! of nmeas measurements, each the average over the 2*meas_num + 1 slices
  ! around ntime/2.  out_addr ends with '/' and holds the temp/ and data/
  ! folders; core id 0 writes the reading guide.  Needs total_width, so
  ! pqmc_open comes first.
  subroutine meas_init(nbins, nmeas, meas_num, out_addr, id, n_cores, mpi_nblock, total_w)
    integer, intent(in) :: nbins, nmeas, meas_num, id, n_cores, mpi_nblock
    character(len=*), intent(in) :: out_addr
    integer, intent(in), optional :: total_w
    if (nbins <= 0 .or. nmeas <= 0) error stop 'init_from_namelist: nbins/nmeas must be >0'
    if (present(total_w)) total_width = total_w
    nbin_per_core = nbins
    nmeas_per_bin = nmeas
    meas_number = meas_num
    output_addr = out_addr
    myid = id
    cur_bin = 1
    cur_meas = 1
    if (allocated(meas_acc)) deallocate (meas_acc, bin_sum, bin_mean)
    allocate (meas_acc(total_width), bin_sum(total_width), bin_mean(total_width, nbins))
    meas_acc = (0.0_c_double, 0.0_c_double)
    bin_sum = 0.0_c_double
    bin_mean = 0.0_c_double
    bins_on = .true.
    if (id == 0) call output_reading_guide(n_cores, mpi_nblock)
  end subroutine meas_init

  ! The bookkeeping of begin_measure (src/mod_meas.f90:736-755) for one
  ! slice's measurement, k-space companions included.
  subroutine record_measurement(forward, time, sample)
    logical, intent(in) :: forward
    integer, intent(in) :: time
    complex(c_double_complex), intent(in) :: sample(:)
    real(c_double) :: d
    if (cur_bin > nbin_per_core) error stop 'begin_measure: no more bins left'
    meas_acc = meas_acc + sample
    if ((forward .and. time == ntime / 2 + meas_number) .or. &
        ((.not. forward) .and. time == ntime / 2 - meas_number)) then
      d = real(2 * meas_number + 1, c_double)
      meas_acc = cmplx(real(meas_acc, c_double) / d, aimag(meas_acc) / d, c_double_complex)
      bin_sum = bin_sum + real(meas_acc, c_double)
      meas_acc = (0.0_c_double, 0.0_c_double)
      cur_meas = cur_meas + 1
      if (cur_meas > nmeas_per_bin) then
        bin_mean(:, cur_bin) = bin_sum / real(nmeas_per_bin, c_double)
        bin_sum = 0.0_c_double
        call output_bin_data(cur_bin)
        cur_meas = 1
        cur_bin = cur_bin + 1
      end if
    end if
  end subroutine record_measurement

  ! data_analyse (src/mod_meas.f90:843-852): every bin's mean (unfinished
  ! bins stay zero) to data/out_core<id>.bin
  subroutine data_analyse()
    integer :: u, ios
    if (.not. bins_on) error stop 'data_analyse: storage not initialized'
    open (newunit=u, file=trim(output_addr)//'data/out_core'//trim(i0str(myid))//'.bin', &
          status='replace', access='stream', form='unformatted', iostat=ios)
    if (ios /= 0) error stop 'output_all_data: cannot open file'
    write (u) bin_mean
    close (u)
  end subroutine data_analyse

  ! output_bin_data + output_bin_count (src/mod_meas.f90:886-942)
  subroutine output_bin_data(b)
    integer, intent(in) :: b
    integer :: u, ios
    character(len=600) :: path
    logical :: ex
    path = trim(output_addr)//'temp/tempout_core'//trim(i0str(myid))//'.bin'
    inquire (file=trim(path), exist=ex)
    if (ex) then
      open (newunit=u, file=trim(path), status='old', action='write', position='append', &
            access='stream', form='unformatted', iostat=ios)
    else
      open (newunit=u, file=trim(path), status='new', action='write', access='stream', &
            form='unformatted', iostat=ios)
    end if
    if (ios /= 0) error stop 'output_bin_data: cannot open file'
    write (u) bin_mean(:, b)
    close (u)
    path = trim(output_addr)//'temp/tempout_core_count.dat'
    inquire (file=trim(path), exist=ex)
    if (ex) then
      open (newunit=u, file=trim(path), status='old', action='write', position='append', iostat=ios)
    else
      open (newunit=u, file=trim(path), status='new', action='write', iostat=ios)
    end if
    if (ios /= 0) error stop 'output_bin_count: cannot open file'
    write (u, '(A,1i4,A,1i4)') 'C', myid, '   |    B', b
    close (u)
  end subroutine output_bin_data

  ! output_reading_guide (src/mod_meas.f90:973-1017), the MPI build's
  ! fields, in fixed formats
  subroutine output_reading_guide(n_cores, mpi_nblock)
    integer, intent(in) :: n_cores, mpi_nblock
    integer :: u, ios, i, lo
    open (newunit=u, file=trim(output_addr)//'data/reading_guide.nml', status='replace', action='write', &
          iostat=ios)
    if (ios /= 0) error stop 'output_reading_guide: cannot open file'
    write (u, '(a)') ' &basics'
    write (u, '(a,i12)') ' n_obs=', 3 + 2 * size(obs_fields)
    write (u, '(a,i12)') ' total_width=', total_width
    write (u, '(a,i12)') ' n_bins=', nbin_per_core
    write (u, '(a,i12)') ' n_meas_per_bin=', nmeas_per_bin
    write (u, '(a,i12)') ' field_width=', ncell
    write (u, '(a,i12)') ' size_of_dp=', storage_size(1.0_c_double) / 8
    write (u, '(a,i12)') ' n_cores=', n_cores
    write (u, '(a,i12)') ' MPI_nblock=', mpi_nblock
    write (u, '(a,i12)') ' MPI_one_block=', n_cores / mpi_nblock
    write (u, '(a)') ' output_format="'//trim(output_format)//'"'
    write (u, '(a)') ' /'
    write (u, '(a)') ' '
    lo = 1
    do i = 1, size(obs_scalars)
      call guide_obs(u, obs_scalars(i), KIND_SCALAR, 1, lo)
    end do
    do i = 1, size(obs_fields)
      call guide_obs(u, obs_fields(i), KIND_FIELD, ncell, lo)
      call guide_obs(u, trim(obs_fields(i))//'_k', KIND_FIELD, ncell, lo)
    end do
    close (u)
  end subroutine output_reading_guide

  subroutine guide_obs(u, name, kind, width, lo)
    integer, intent(in) :: u, kind, width
    character(len=*), intent(in) :: name
    integer, intent(inout) :: lo
    write (u, '(a)') ' &obs'
    write (u, '(a)') ' name="'//trim(name)//'"'
    write (u, '(a,i12)') ' kind=', kind
    write (u, '(a,i12)') ' width=', width
    write (u, '(a,i12)') ' offset_lo=', lo
    write (u, '(a,i12)') ' offset_hi=', lo + width - 1
    write (u, '(a)') ' /'
    write (u, '(a)') ' '
    lo = lo + width
  end subroutine guide_obs

  function i0str(v) result(s)
    integer, intent(in) :: v
    character(len=16) :: s
    write (s, '(i0)') v
  end function i0str

  ! ------------------------------------------------------------------
  ! phonon-field text files (src/Main_PQMC.f90:355-368) and checkpoints

  ! boson_field(:, t), t = 1..size(x, 2), one (Nb)F18.6 line per slice
  subroutine read_ph_field(path, x)
    character(len=*), intent(in) :: path
    real(c_double), intent(out) :: x(:, :)
    integer :: u, t, ios
    open (newunit=u, file=path, status='old', position='rewind', action='read', iostat=ios)
    if (ios /= 0) error stop 'import_ph_field: cannot open file'
    do t = 1, size(x, 2)
      read (u, '('//trim(i0str(size(x, 1)))//'f18.6)', iostat=ios) x(:, t)
      if (ios /= 0) error stop 'import_ph_field: cannot read a slice'
    end do
    close (u)
  end subroutine read_ph_field

  subroutine write_ph_field(path, x)
    character(len=*), intent(in) :: path
    real(c_double), intent(in) :: x(:, :)
    integer :: u, t, ios
    open (newunit=u, file=path, status='replace', action='write', iostat=ios)
    if (ios /= 0) error stop 'export_ph_field: cannot open file'
    do t = 1, size(x, 2)
      write (u, '('//trim(i0str(size(x, 1)))//'f18.6)') x(:, t)
    end do
    close (u)
  end subroutine write_ph_field

  ! The chain's fields from / to a text field file; init_g follows an import.
  subroutine import_ph_field(path)
    character(len=*), intent(in) :: path
    call read_ph_field(path, bf_cur)
    call pqmc_check(pqmc_set_fields(chain, bf_cur), 'import_ph_field')
  end subroutine import_ph_field

  subroutine export_ph_field(path)
    character(len=*), intent(in) :: path
    call pqmc_check(pqmc_get_fields(chain, bf_cur), 'export_ph_field')
    call write_ph_field(path, bf_cur)
  end subroutine export_ph_field

  ! Checkpoint at a loop boundary before a backward sweep (binary stream:
  ! magic, sizes, loop and direction, ran2 state, counters, fields, the
  ! measurement sums and bins).  loop is the last completed loop, forward its
  ! direction.
  subroutine checkpoint_save(path, loop, forward)
    character(len=*), intent(in) :: path
    integer, intent(in) :: loop
    logical, intent(in) :: forward
    integer(c_int64_t) :: cnt(3)
    real(c_double) :: efm
    if (.not. forward) error stop 'checkpoint_save: only after a forward sweep (resume runs init_g(1))'
    call pqmc_check(pqmc_get_fields(chain, bf_cur), 'checkpoint_save')
    call pqmc_check(pqmc_get_counters(chain, cnt, efm), 'checkpoint_save')
    call write_checkpoint(path, loop, forward, cnt(1), cnt(2), efm, bf_cur)
  end subroutine checkpoint_save

  ! Resume: fields, init_g(1), counters, ran2 state and bins as saved.  The
  ! caller continues with loop + 1.
  subroutine checkpoint_load(path, loop, forward)
    character(len=*), intent(in) :: path
    integer, intent(out) :: loop
    logical, intent(out) :: forward
    integer(c_int64_t) :: pa, na
    real(c_double) :: efm
    call read_checkpoint(path, loop, forward, pa, na, efm, bf_cur)
    call pqmc_check(pqmc_set_fields(chain, bf_cur), 'checkpoint_load')
    call init_g(1)
    call pqmc_check(pqmc_set_counters(chain, [pa, na], efm), 'checkpoint_load')
  end subroutine checkpoint_load

  subroutine write_checkpoint(path, loop, forward, pa, na, efm, x)
    character(len=*), intent(in) :: path
    integer, intent(in) :: loop
    logical, intent(in) :: forward
    integer(c_int64_t), intent(in) :: pa, na
    real(c_double), intent(in) :: efm, x(:, :)
    integer :: u, ios
    open (newunit=u, file=path, status='replace', access='stream', form='unformatted', action='write', iostat=ios)
    if (ios /= 0) error stop 'checkpoint_save: cannot open file'
    write (u) 'PQMCCKP1'
    write (u) int([lat_L1, ncell / lat_L1, ntime, ngroup, nelec, nb, loop, merge(1, 0, forward)], c_int32_t)
    write (u) rng_state%i, rng_state%j, rng_state%k, rng_state%m, rng_state%n
    write (u) pa, na, global_accept, global_reject
    write (u) efm
    write (u) x
    write (u) int(total_width, c_int32_t), int(n_meas, c_int32_t)
    write (u) meas_sum
    if (bins_on) then
      write (u) int([1, nbin_per_core, nmeas_per_bin, meas_number, cur_bin, cur_meas], c_int32_t)
      write (u) meas_acc, bin_sum, bin_mean
    else
      write (u) int([0, 0, 0, 0, 0, 0], c_int32_t)
    end if
    close (u)
  end subroutine write_checkpoint

  subroutine read_checkpoint(path, loop, forward, pa, na, efm, x)
    character(len=*), intent(in) :: path
    integer, intent(out) :: loop
    logical, intent(out) :: forward
    integer(c_int64_t), intent(out) :: pa, na
    real(c_double), intent(out) :: efm, x(:, :)
    character(len=8) :: magic
    integer(c_int32_t) :: h(8), r(5), b(6), tw, nm
    integer :: u, ios
    open (newunit=u, file=path, status='old', access='stream', form='unformatted', action='read', iostat=ios)
    if (ios /= 0) error stop 'checkpoint_load: cannot open file'
    read (u) magic
    if (magic /= 'PQMCCKP1') error stop 'checkpoint_load: not a checkpoint of this library'
    read (u) h
    if (h(1) /= lat_L1 .or. h(1) * h(2) /= ncell .or. h(3) /= ntime .or. h(4) /= ngroup .or. h(5) /= nelec &
        .or. h(6) /= nb) error stop 'checkpoint_load: sizes differ from the open chain'
    loop = h(7)
    forward = h(8) /= 0
    read (u) r
    rng_state%i = r(1); rng_state%j = r(2); rng_state%k = r(3); rng_state%m = r(4); rng_state%n = r(5)
    read (u) pa, na, global_accept, global_reject
    read (u) efm
    read (u) x
    read (u) tw, nm
    if (tw /= total_width) error stop 'checkpoint_load: measurement width differs'
    n_meas = nm
    read (u) meas_sum
    read (u) b
    if (b(1) /= 0) then
      if (.not. bins_on) error stop 'checkpoint_load: the checkpoint has bins; call meas_init first'
      if (b(2) /= nbin_per_core .or. b(3) /= nmeas_per_bin .or. b(4) /= meas_number) &
        error stop 'checkpoint_load: bins differ from meas_init'
      cur_bin = b(5)
      cur_meas = b(6)
      read (u) meas_acc, bin_sum, bin_mean
    end if
    close (u)
    lncw_held = .false.
  end subroutine read_checkpoint

end module pqmc_dropin
