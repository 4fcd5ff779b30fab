! Drop-in replacements for the reference's hot-path subroutines, with the
! reference's names and arguments, backed by one GPU-resident chain:
!   update_phonon_field(ppf, time)                (src/mod_update.f90:13-201)
!   left_evolve(time, mat, n, pf_id, two_way)      (src/mod_evolution.f90:236-285)
!   right_evolve(time, mat, n, pf_id, two_way)     (src/mod_evolution.f90:287-328)
!   init_g(start_time)                             (init_g_T0, :46-111)
!   get_g_scratch(time, forward, loop)             (get_g_scratch_T0, :113-195)
!   begin_measure(forward, time) / take_measurement(time)  (src/mod_meas.f90:210-544,729-763)
!   update_global(update_type)                     (src/mod_update.f90:251-290)
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
  implicit none

  type :: pf_type
    integer :: id = 0
    logical :: V_exist = .true.
  end type pf_type

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
  integer, save :: lat_L1 = 0
  logical, save :: updated = .false.
  integer(c_int64_t), save :: global_accept = 0, global_reject = 0
  ! ln_cw after an accepted global move: the device value is stale until the
  ! next init_g, which the caller runs when updated is set
  logical, save :: lncw_held = .false.
  complex(c_double_complex), save :: lncw_hold = (0.0_c_double, 0.0_c_double)
  real(c_double), allocatable, save :: bf_cur(:, :), bf_new(:, :)

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
  subroutine begin_measure(forward, time)
    logical, intent(in) :: forward
    integer, intent(in) :: time
    call take_measurement(time)
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

end module pqmc_dropin
