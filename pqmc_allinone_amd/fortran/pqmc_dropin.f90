! Drop-in replacements for the reference's hot-path subroutines, with the
! reference's names and arguments, backed by one GPU-resident chain:
!   update_phonon_field(ppf, time)                (src/mod_update.f90:13-201)
!   left_evolve(time, mat, n, pf_id, two_way)      (src/mod_evolution.f90:236-285)
!   right_evolve(time, mat, n, pf_id, two_way)     (src/mod_evolution.f90:287-328)
!   init_g(start_time)                             (init_g_T0, :46-111)
!   get_g_scratch(time, forward, loop)             (get_g_scratch_T0, :113-195)
!   begin_measure(forward, time) / take_measurement(time)  (src/mod_meas.f90:210-544,729-763)
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

contains

  ! Allocate the chain and load the model, fields (boson_field(Nb, Lt)), trial
  ! Slater matrix (Ns x ne) and the ran2 seed (init_rng_by_seed,
  ! src/mod_ranstate.f90:267-283).
  subroutine pqmc_open(cfg, model, fields, slater, seed)
    type(pqmc_config), intent(in) :: cfg
    type(pqmc_model), intent(in) :: model
    real(c_double), intent(in) :: fields(*)
    complex(c_double_complex), intent(in) :: slater(*)
    integer, intent(in) :: seed
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
    allocate (g_h(ns, ns), meas_sum(total_width), last_accepts(n_pla))
    g_h = (0.0_c_double, 0.0_c_double)
    meas_sum = (0.0_c_double, 0.0_c_double)
    n_meas = 0
  end subroutine pqmc_open

  subroutine pqmc_close()
    integer(c_int) :: rc
    if (c_associated(chain)) rc = pqmc_destroy(chain)
    chain = c_null_ptr
    if (allocated(g_h)) deallocate (g_h, meas_sum, last_accepts)
  end subroutine pqmc_close

  subroutine init_g(start_time)
    integer, intent(in) :: start_time
    call pqmc_check(pqmc_init_g(chain, int(start_time, c_int)), 'init_g')
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

end module pqmc_dropin
