! iso_c_binding interfaces to libpqmc_hip.so.  One interface per function of
! include/pqmc_hip.h, with the same argument meaning; pqmc_check turns a
! non-zero status into `error stop` (the reference stops on a failed LAPACK
! call, src/mod_matrixlib.f90:366-430).
module pqmc_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PQMC_OK = 0, PQMC_EINVAL = 1, PQMC_EHIP = 2, PQMC_ENUMERIC = 3, &
                               PQMC_ESYNC = 4, PQMC_ENODEV = 5

  type, bind(C) :: pqmc_config
    integer(c_int) :: L1, L2, Lt, ngroup, ne, block_bonds
    integer(c_int) :: helper_bonds = 0
    integer(c_int) :: finite_temperature = 0
  end type pqmc_config

  type, bind(C) :: pqmc_model
    type(c_ptr) :: sites, bfi, cross, bcph, kcoe, vcoe, k12, v12, bc
    real(c_double) :: dtau, M, D, biased, jump_distance, hop, lam
    integer(c_int32_t) :: ncopy, tr_weight
  end type pqmc_model

  type, bind(C) :: pqmc_ran2
    integer(c_int32_t) :: i, j, k, m, n
  end type pqmc_ran2

  interface
    integer(c_int) function pqmc_create(cfg, h) bind(C, name='pqmc_create')
      import :: c_int, c_ptr, pqmc_config
      type(pqmc_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function pqmc_create

    integer(c_int) function pqmc_destroy(h) bind(C, name='pqmc_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function pqmc_destroy

    type(c_ptr) function pqmc_last_error() bind(C, name='pqmc_last_error')
      import :: c_ptr
    end function pqmc_last_error

    integer(c_int) function pqmc_get_dims(h, out6) bind(C, name='pqmc_get_dims')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: out6(6)
    end function pqmc_get_dims

    integer(c_int) function pqmc_set_model(h, m) bind(C, name='pqmc_set_model')
      import :: c_int, c_ptr, pqmc_model
      type(c_ptr), value :: h
      type(pqmc_model), intent(in) :: m
    end function pqmc_set_model

    integer(c_int) function pqmc_set_fields(h, x) bind(C, name='pqmc_set_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
    end function pqmc_set_fields

    integer(c_int) function pqmc_get_fields(h, x) bind(C, name='pqmc_get_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: x(*)
    end function pqmc_get_fields

    integer(c_int) function pqmc_set_trial(h, slater) bind(C, name='pqmc_set_trial')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: slater(*)
    end function pqmc_set_trial

    integer(c_int) function pqmc_init_g(h, start_time) bind(C, name='pqmc_init_g')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: start_time
    end function pqmc_init_g

    integer(c_int) function pqmc_update_family(h, pf_id, time, u, acc) bind(C, name='pqmc_update_family')
      import :: c_int, c_ptr, c_float, c_int8_t
      type(c_ptr), value :: h
      integer(c_int), value :: pf_id, time
      real(c_float), intent(in) :: u(*)
      integer(c_int8_t), intent(out) :: acc(*)
    end function pqmc_update_family

    integer(c_int) function pqmc_evolve_g(h, time, pf_id, forward) bind(C, name='pqmc_evolve_g')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: time, pf_id, forward
    end function pqmc_evolve_g

    integer(c_int) function pqmc_evolve_matrix(h, time, pf_id, left, two_way, mat, n) &
        bind(C, name='pqmc_evolve_matrix')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: time, pf_id, left, two_way, n
      complex(c_double_complex), intent(inout) :: mat(*)
    end function pqmc_evolve_matrix

    integer(c_int) function pqmc_g_scratch(h, time, forward, err_fast) bind(C, name='pqmc_g_scratch')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: time, forward
      real(c_double), intent(out) :: err_fast
    end function pqmc_g_scratch

    integer(c_int) function pqmc_measure(h, time, out) bind(C, name='pqmc_measure')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: time
      complex(c_double_complex), intent(out) :: out(*)
    end function pqmc_measure

    integer(c_int) function pqmc_sweep(h, forward, u, ntc, measure, meas_out) bind(C, name='pqmc_sweep')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: forward, ntc, measure
      real(c_float), intent(in) :: u(*)
      type(c_ptr), value :: meas_out
    end function pqmc_sweep

    ! hs, u and meas_out are arrays of c_ptr (c_loc of each chain's arrays)
    integer(c_int) function pqmc_sweep_batch(hs, n, forward, u, measure, meas_out) &
        bind(C, name='pqmc_sweep_batch')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int), value :: n, measure
      integer(c_int), intent(in) :: forward(*)
      type(c_ptr), intent(in) :: u(*)
      type(c_ptr), value :: meas_out
    end function pqmc_sweep_batch

    integer(c_int) function pqmc_sweep_many(hs, n, forward, u, measure, meas_out, max_concurrent) &
        bind(C, name='pqmc_sweep_many')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: hs(*), u(*)
      integer(c_int), value :: n, measure, max_concurrent
      integer(c_int), intent(in) :: forward(*)
      type(c_ptr), value :: meas_out
    end function pqmc_sweep_many

    integer(c_int) function pqmc_get_g(h, g) bind(C, name='pqmc_get_g')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(out) :: g(*)
    end function pqmc_get_g

    integer(c_int) function pqmc_set_g(h, g) bind(C, name='pqmc_set_g')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: g(*)
    end function pqmc_set_g

    integer(c_int) function pqmc_get_qstring(h, block, q) bind(C, name='pqmc_get_qstring')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: block
      complex(c_double_complex), intent(out) :: q(*)
    end function pqmc_get_qstring

    integer(c_int) function pqmc_get_dstring(h, d) bind(C, name='pqmc_get_dstring')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(out) :: d(*)
    end function pqmc_get_dstring

    integer(c_int) function pqmc_get_lncw(h, re_im) bind(C, name='pqmc_get_lncw')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: re_im(2)
    end function pqmc_get_lncw

    integer(c_int) function pqmc_global_weight(h, x, re_im) bind(C, name='pqmc_global_weight')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: re_im(2)
    end function pqmc_global_weight

    ! d ln_cw / dx for every field, complex, boson_field layout (Nb, Lt)
    integer(c_int) function pqmc_gradient(h, grad) bind(C, name='pqmc_gradient')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(out) :: grad(*)
    end function pqmc_gradient

    ! the same pass with the accumulator in the library's fixed device array
    ! (no device allocation; at most 1,048,576 fields)
    integer(c_int) function pqmc_gradient_noalloc(h, grad) bind(C, name='pqmc_gradient_noalloc')
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      complex(c_double_complex), intent(out) :: grad(*)
    end function pqmc_gradient_noalloc

    ! B-string products between QR checkpoints: 0 FP64, 1 FP32x3, 2 BF16x3
    integer(c_int) function pqmc_set_precision(h, mode) bind(C, name='pqmc_set_precision')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function pqmc_set_precision

    ! stabilization QR: 0 one launch (default), 1 panel by panel
    integer(c_int) function pqmc_set_qr_mode(h, mode) bind(C, name='pqmc_set_qr_mode')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function pqmc_set_qr_mode

    ! the latest sweep's largest G and Re ln_cw drifts
    integer(c_int) function pqmc_get_drift(h, out2) bind(C, name='pqmc_get_drift')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out2(2)
    end function pqmc_get_drift

    ! every stabilization's (G, Re ln_cw) drift pair of the latest sweep
    integer(c_int) function pqmc_get_drift_history(h, out, cap, n) bind(C, name='pqmc_get_drift_history')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: n
    end function pqmc_get_drift_history

    ! get_expKV_epsoc (src/mod_phonon_field.f90:572-604) for n (bond, slice)
    ! pairs: out(2, 2, n) holds E or (inv /= 0) E^-1; bcph may be c_null_ptr
    ! with cross when no bond crosses
    integer(c_int) function pqmc_expkv_epsoc(dtau, nfam, kcoe, vcoe, k12, v12, n, fam, x, cross, bcph, inv, out) &
        bind(C, name='pqmc_expkv_epsoc')
      import :: c_int, c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr
      real(c_double), value :: dtau
      integer(c_int32_t), value :: nfam, inv
      complex(c_double_complex), intent(in) :: kcoe(*), vcoe(*), k12(*), v12(*)
      integer(c_int64_t), value :: n
      integer(c_int32_t), intent(in) :: fam(*)
      real(c_double), intent(in) :: x(*)
      type(c_ptr), value :: cross, bcph
      complex(c_double_complex), intent(out) :: out(*)
    end function pqmc_expkv_epsoc

    integer(c_int) function pqmc_get_accepts(h, acc, n) bind(C, name='pqmc_get_accepts')
      import :: c_int, c_ptr, c_int8_t, c_int64_t
      type(c_ptr), value :: h
      integer(c_int8_t), intent(out) :: acc(*)
      integer(c_int64_t), value :: n
    end function pqmc_get_accepts

    integer(c_int) function pqmc_get_counters(h, out3, err_fast_max) bind(C, name='pqmc_get_counters')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(out) :: out3(3)
      real(c_double), intent(out) :: err_fast_max
    end function pqmc_get_counters

    integer(c_int) function pqmc_set_counters(h, in2, err_fast_max) bind(C, name='pqmc_set_counters')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), intent(in) :: in2(2)
      real(c_double), value :: err_fast_max
    end function pqmc_set_counters

    integer(c_int) function pqmc_synchronize(h) bind(C, name='pqmc_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function pqmc_synchronize

    subroutine pqmc_ran2_seed(s, seed) bind(C, name='pqmc_ran2_seed')
      import :: pqmc_ran2, c_int32_t
      type(pqmc_ran2), intent(inout) :: s
      integer(c_int32_t), value :: seed
    end subroutine pqmc_ran2_seed

    subroutine pqmc_ran2_fill(s, n, out) bind(C, name='pqmc_ran2_fill')
      import :: pqmc_ran2, c_int64_t, c_float
      type(pqmc_ran2), intent(inout) :: s
      integer(c_int64_t), value :: n
      real(c_float), intent(out) :: out(*)
    end subroutine pqmc_ran2_fill

    integer(c_int) function pqmc_update_profile(h, enable, out, p) bind(C, name='pqmc_update_profile')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, out, p
      integer(c_int), value :: enable
    end function pqmc_update_profile

    integer(c_int) function pqmc_mfma_selftest(max_abs_err) bind(C, name='pqmc_mfma_selftest')
      import :: c_int, c_double
      real(c_double), intent(out) :: max_abs_err
    end function pqmc_mfma_selftest

    integer(c_int) function pqmc_timing(h, enable, kind, ms, launches) bind(C, name='pqmc_timing')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: h
      integer(c_int), value :: enable, kind
      real(c_double), intent(out) :: ms
      integer(c_int64_t), intent(out) :: launches
    end function pqmc_timing
  end interface

contains

  ! error stop with the library's message when rc is not PQMC_OK
  subroutine pqmc_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == PQMC_OK) return
    text = ''
    if (c_associated(pqmc_last_error())) then
      call c_f_pointer(pqmc_last_error(), msg, [512])
      do i = 1, 512
        if (msg(i) == c_null_char) exit
        text(i:i) = msg(i)
      end do
    end if
    write (0, '(a,a,a,i0,a,a)') 'pqmc: ', where, ' failed with status ', rc, ': ', trim(text)
    error stop 1
  end subroutine pqmc_check

end module pqmc_hip
