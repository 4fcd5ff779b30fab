/*
 * pqmc_hip.h -- C ABI of the MI355X projector-DQMC library (libpqmc_hip.so).
 *
 * The reference (LEOMMM816/PQMC_allinone) has no FFI: its boundary is the
 * set of Fortran module subroutines the sweep loop of src/Main_PQMC.f90:81-156
 * calls.  Each entry point below replaces one of them; a Fortran host binds
 * them with iso_c_binding (pqmc_allinone_amd/fortran/pqmc_hip_mod.f90).
 *
 * Conventions: plain pointers and sizes only.  Complex numbers are pairs of
 * doubles (re, im), matrices column-major as complex(dp) arrays in Fortran.
 * `time` and `pf_id` are 1-based exactly as in the reference.  Device state
 * (G, fields, B-string checkpoints, expKV table) is owned by the handle;
 * host arrays are copied in or out only by the set_/get_ calls.
 * Every function returns 0 on success or a pqmc_status code; the message of
 * the last failure is available from pqmc_last_error().
 */
#ifndef PQMC_HIP_H
#define PQMC_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum pqmc_status {
  PQMC_OK = 0,
  PQMC_EINVAL = 1,     /* bad argument or call order */
  PQMC_EHIP = 2,       /* HIP runtime error */
  PQMC_ENUMERIC = 3,   /* non-finite Green's function or log-weight */
  PQMC_ESYNC = 4,      /* a bounded device-side wait expired */
  PQMC_ENODEV = 5      /* no usable gfx950 device */
};

typedef struct pqmc_handle pqmc_handle;

/* Sizes of one chain.  L1 x L2 cells with two sites each (Ns = 2 L1 L2),
 * four checkerboard bond families of L1 L2 / 2 bonds, Nb = 2 L1 L2 fields.
 * Mirrors src/input.f90:167-205 (ntime, ngroup, nblock, nelec). */
typedef struct pqmc_config {
  int L1, L2;
  int Lt;             /* ntime */
  int ngroup;         /* stabilization interval */
  int ne;             /* nelec */
  int block_bonds;    /* bonds per decision block of a family pass (1..8), 0 = automatic */
  int helper_bonds;   /* bonds per column-owning workgroup of a family pass (1, 2 or 4), 0 = automatic:
                         1 when the pass fits the device.  2 and 4 divide the compute units a pass
                         keeps resident, so that more chains fit a batch (pqmc_sweep_batch). */
  int finite_temperature;  /* 0: projector (proj = .true.); 1: finite temperature, G = (1 + B(beta,0))^-1
                              from Ns x Ns UDT strings (init_g_FT / get_g_scratch_FT,
                              src/mod_evolution.f90:411-785).  Needs Lt % ngroup == 0; no trial
                              matrix; the projector string readers refuse it. */
} pqmc_config;

/* Model tables as built by set_typed_pf / init_pf_para
 * (src/mod_phonon_field.f90:249-421, src/input.f90:207-216). */
typedef struct pqmc_model {
  const int32_t *sites;  /* [4][npla][4] pla_site_list, 0-based sites */
  const int32_t *bfi;    /* [4][npla]    bf_list, 0-based field index */
  const int32_t *cross;  /* [4][npla]    boundary_crossing */
  const double *bcph;    /* [4][npla][4x4] complex BC_phases, column-major */
  const double *kcoe;    /* [4] complex K_coe (already times -hop) */
  const double *vcoe;    /* [4] complex V_coe (already times ep_parameter) */
  const double *k12;     /* [4][2x2] complex Kmatrix(1:2,3:4), column-major */
  const double *v12;     /* [4][2x2] complex Vmatrix(1:2,3:4), column-major */
  const double *bc;      /* [2] complex Lat%BC_phase */
  double dtau, M, D, biased, jump_distance, hop, lam;
  int32_t ncopy, tr_weight;
} pqmc_model;

/* Allocate a chain on the current device.  No reference counterpart (the
 * reference allocates module globals in init / init_g_T0). */
int pqmc_create(const pqmc_config *cfg, pqmc_handle **out);
int pqmc_destroy(pqmc_handle *h);
const char *pqmc_last_error(void);
/* Derived sizes: Ns, Ncell, Nb, npla, nblock, total_width. */
int pqmc_get_dims(const pqmc_handle *h, int32_t *out6);

/* Model tables; replaces set_pf_main + init_pf_para (mod_phonon_field.f90:22-57). */
int pqmc_set_model(pqmc_handle *h, const pqmc_model *m);
/* boson_field(Nb, Lt) in/out (src/input.f90:86). */
int pqmc_set_fields(pqmc_handle *h, const double *x);
int pqmc_get_fields(pqmc_handle *h, double *x);
/* Trial Slater matrix (Ns x ne complex); QDR on the device as
 * set_nelec_occupied (src/Main_PQMC.f90:301-302). */
int pqmc_set_trial(pqmc_handle *h, const double *slater);

/* Precision of the B-string products between QR checkpoints, the one-way
 * left_evolve / right_evolve loops of init_g_T0 and get_g_scratch_T0
 * (src/mod_evolution.f90:88-102, 137-162) and of cal_new_weight_T0
 * (:359-408): 0 = FP64 (default); 1 = the slices' product P formed in FP64
 * and P X taken on FP32 MFMA with three-term split operands; 2 = the same on
 * BF16 MFMA with BF16x3 operands.  QR, D, the (L R)^-1 solve, the updates and
 * every determinant ratio stay FP64.  Projector chains only.  No reference
 * counterpart (the reference computes in FP64 throughout). */
int pqmc_set_precision(pqmc_handle *h, int mode);

/* Householder QR of the stabilizations (ZGEQRF + ZUNGQR in qdr / ldq,
 * src/mod_matrixlib.f90:323-360): 0 = one launch per factorization, a
 * workgroup per 16-column panel, each panel's reflector handed to the panels
 * right of it through flags (default); 1 = a panel launch and a
 * block-reflector launch per panel.  Same conventions and outputs, equal to
 * rounding.  No reference counterpart (LAPACK's blocking is internal). */
int pqmc_set_qr_mode(pqmc_handle *h, int mode);

/* init_MC: init_expKV + set_pf_expKV_mat + init_g(start_time)
 * (src/Main_PQMC.f90:185-203, src/mod_evolution.f90:46-111). */
int pqmc_init_g(pqmc_handle *h, int start_time);

/* update_phonon_field(ppf_list(pf_id), time) (src/mod_update.f90:13-201).
 * u: 2*npla ran2 draws in call order (ran_sysm source, rands) per bond.
 * acc (npla bytes, may be NULL) receives the accept flags. */
int pqmc_update_family(pqmc_handle *h, int pf_id, int time, const float *u, uint8_t *acc);

/* left_evolve / right_evolve on g_h with two_way = .true.
 * (src/mod_evolution.f90:236-328): forward != 0 -> G <- E G E^-1, else
 * G <- E^-1 G E, for family pf_id at slice time. */
int pqmc_evolve_g(pqmc_handle *h, int time, int pf_id, int forward);

/* left_evolve(time, mat(Ns,n), n, pf_id, two_way) / right_evolve(time,
 * mat(n,Ns), ...) applied to a caller matrix (host memory, in place). */
int pqmc_evolve_matrix(pqmc_handle *h, int time, int pf_id, int left, int two_way, double *mat, int n);

/* get_g_scratch(time, forward, loop) (src/mod_evolution.f90:113-195).
 * err_fast (may be NULL) receives max|G_fast - G_scratch|.  Returns
 * PQMC_ENUMERIC when the G or ln_cw drift is not finite. */
int pqmc_g_scratch(pqmc_handle *h, int time, int forward, double *err_fast);

/* Meas_sys%take_measurement(time) + k_space_obs (src/mod_meas.f90:210-566)
 * for a begin_measure call with meas_number = 0; out: total_width complex. */
int pqmc_measure(pqmc_handle *h, int time, double *out);

/* One iteration of the sweep loop body (src/Main_PQMC.f90:85-154) with the
 * global update off: Lt slices x 4 families of updates and two-way wraps,
 * get_g_scratch every ngroup slices, and (measure != 0) one measurement at
 * time = Lt/2.  u: 2*Nb*Lt draws in call order.  ntc limits the number of
 * time_count iterations (0 = all).  meas_out (total_width complex) may be
 * NULL.  Accept flags stay on the device (pqmc_get_accepts).  The drift of
 * every stabilization of the sweep is kept on the device and read once at
 * the end: PQMC_ENUMERIC if any of them was not finite (the sweep's later
 * results are then invalid). */
int pqmc_sweep(pqmc_handle *h, int forward, const float *u, int ntc, int measure, double *meas_out);

/* n chains, one sweep each, advanced at the same time (the batched form):
 * one host thread per chain runs pqmc_sweep(hs[i], forward[i], u[i], 0,
 * measure, meas_out[i]) on the chain's own stream.  Chains may differ in
 * lattice and Lt.  The family passes of all n chains must be resident on the
 * device together; a batch that does not fit is refused with PQMC_EINVAL.
 * meas_out may be NULL (or hold NULL entries).  Returns the first failing
 * chain's status.  No reference counterpart: the reference runs one chain
 * per MPI rank (src/Main_PQMC.f90:380-400). */
int pqmc_sweep_batch(pqmc_handle *const *hs, int n, const int *forward, const float *const *u, int measure,
                     double *const *meas_out);

/* Any number of chains, one sweep each (the batched form for more chains
 * than fit together, e.g. 128 chains of uneven lattice and Lt): up to
 * max_concurrent host threads (0 = 16) take the chains in order, and a
 * chain's sweep starts once the compute units its family passes need fit
 * beside those of the sweeps already running.  A chain that needs more
 * compute units than the device has is refused with PQMC_EINVAL before
 * anything runs.  Returns the first failing chain's status.  No reference
 * counterpart (src/Main_PQMC.f90:380-400 runs one chain per MPI rank). */
int pqmc_sweep_many(pqmc_handle *const *hs, int n, const int *forward, const float *const *u, int measure,
                    double *const *meas_out, int max_concurrent);

/* State readers/writers (parity and checkpointing). */
int pqmc_get_g(pqmc_handle *h, double *g);
int pqmc_set_g(pqmc_handle *h, const double *g);
int pqmc_get_qstring(pqmc_handle *h, int block, double *q);   /* Q_string(:,:,block) */
int pqmc_get_dstring(pqmc_handle *h, double *d);              /* D_string(ne, nblock) */
/* ln_cw: the last stabilization's value plus the accepts' log(real(R_elec))
 * since then (src/mod_update.f90:142, src/mod_evolution.f90:330-357). */
int pqmc_get_lncw(pqmc_handle *h, double *re_im);
int pqmc_get_accepts(pqmc_handle *h, uint8_t *acc, int64_t n); /* last sweep, call order */
/* positive_accept, negative_accept, sync timeouts; err_fast_max, the running
 * maximum of both drift checks of get_g_scratch_T0 (max|G_fast - G_scratch|
 * and |Re(ln_cw - ln_cw_fast)|, src/mod_evolution.f90:179-193) over every
 * stabilization, starting from 1e-6 as the reference's module variable. */
int pqmc_get_counters(pqmc_handle *h, int64_t *out3, double *err_fast_max);
/* out2: the largest G drift and the largest Re ln_cw drift over the
 * stabilizations of the latest pqmc_sweep, pqmc_gradient or pqmc_g_scratch
 * call (not floored at 1e-6).  Diagnostics; the reference prints each new
 * maximum instead (src/mod_evolution.f90:180-192). */
int pqmc_get_drift(pqmc_handle *h, double *out2);
/* The same two drifts for each stabilization of the latest pqmc_sweep,
 * pqmc_gradient or pqmc_g_scratch call, in order: out[2 k], out[2 k + 1] for
 * k < min(cap, *n); *n receives the number of stabilizations.  Diagnostics,
 * as the reference's per-event prints (src/mod_evolution.f90:180-192). */
int pqmc_get_drift_history(pqmc_handle *h, double *out, int32_t cap, int32_t *n);
/* Sets positive_accept, negative_accept (in2) and err_fast_max, for a run
 * resumed from a checkpoint.  No reference counterpart: the reference keeps
 * these as module globals (src/mod_update.f90, src/Main_PQMC.f90:222-224)
 * and has no resume. */
int pqmc_set_counters(pqmc_handle *h, const int64_t *in2, double err_fast_max);
int pqmc_synchronize(pqmc_handle *h);

/* Global moves: the full-determinant weight of the field configuration x
 * (Nb*Lt, boson_field layout), replacing cal_new_weight_T0
 * (src/mod_evolution.f90:359-408), or cal_new_weight_FT (:743-785) for a
 * finite-temperature chain, as called by update_global
 * (src/mod_update.f90:251-290).  The chain's fields, Green's function and
 * B-strings are left unchanged; after an accepted move the caller sets the
 * fields and runs pqmc_init_g(1), as src/Main_PQMC.f90:94 does. */
int pqmc_global_weight(pqmc_handle *h, const double *x, double *re_im);

/* Field gradient of the fermionic log-weight: grad (complex, Nb*Lt,
 * boson_field layout) = d ln_cw / dx for every field, as Tr[G dE E^-1] along
 * a backward pass over the slices with no updates, G recomputed from the
 * checkpointed B-strings every ngroup slices.  No counterpart in the
 * reference; it is the adjoint a hybrid-Monte-Carlo global move needs
 * beside update_global (src/mod_update.f90:251-290).  Starts where a
 * backward pqmc_sweep starts (after pqmc_init_g(1) or a forward sweep) and
 * leaves the chain where a backward sweep that accepts nothing leaves it.
 * Projector and finite-temperature chains alike.  Refused (PQMC_EINVAL)
 * when G is known to sit at the bottom: after a full backward sweep or a
 * gradient pass, with no init_g or forward sweep since. */
int pqmc_gradient(pqmc_handle *h, double *grad);

/* pqmc_gradient without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the device accumulator is the library's fixed 16 MiB
 * array, one per device, instead of a buffer allocated on the chain's first
 * gradient pass.  Same arguments, result and call-order rules as
 * pqmc_gradient; PQMC_EINVAL when Nb * Lt exceeds 1,048,576 fields.  Calls
 * on one device share the array: calls that overlap on a device race.  No
 * counterpart in the reference. */
int pqmc_gradient_noalloc(pqmc_handle *h, double *grad);

/* The ran2 generator of src/mod_ranstate.f90:171-191 seeded as
 * init_rng_by_seed (:267-283); host-side, for drivers without their own. */
typedef struct pqmc_ran2 { int32_t i, j, k, m, n; } pqmc_ran2;
void pqmc_ran2_seed(pqmc_ran2 *s, int32_t seed);
void pqmc_ran2_fill(pqmc_ran2 *s, int64_t n, float *out);

/* Debug phase timers of the family-pass kernel: enable allocates and zeroes
 * 8 accumulators per row; out (8*P doubles, microseconds) and P (the row
 * count) may be NULL.  Rows: the deciding wave of workgroup 0, then one per
 * column-owning workgroup, then workgroup 0's B-building and tile-loading
 * waves.  No reference counterpart (diagnostics). */
int pqmc_update_profile(pqmc_handle *h, int enable, double *out, int32_t *P);

/* FP64 MFMA operand-layout self check: C(16x16) = A(16x4) B(4x16). */
int pqmc_mfma_selftest(double *max_abs_err);

/* The stabilization's product and triangular solve on host arrays
 * (column-major, interleaved complex doubles): which = 0 gives
 * C = alpha op(A) op(B) + beta C (the ZGEMM calls of cal_g_and_det_T0,
 * src/mod_evolution.f90:344-350; alpha_beta = {Re a, Im a, Re b, Im b});
 * which = 1 gives C <- R^-1 C with R = A upper triangular (M x M).  A test
 * hook: no reference counterpart. */
int pqmc_dense_check(int which, int opa, int opb, int M, int N, int K, const double *alpha_beta, const double *A,
                     int lda, const double *B, int ldb, double *C, int ldc);

/* get_expKV_epsoc (src/mod_phonon_field.f90:572-604, selected through the
 * get_expKV pointer for the EPSOC model, :472-496): the spin-up 2x2 bond
 * exponential exp(-dtau H_b), H_b = [[0, h], [conj(h), 0]],
 * h = K_coe Kmatrix(1,2) + V_coe x Vmatrix(1,2) times BC_phases(1,2) of a
 * bond that crosses the boundary, in closed form as the reference evaluates
 * it.  For pair i < n: family fam[i] in [0, nfam), field x[i], and (when
 * cross is not null and cross[i] != 0) phase bcph[i]; out[8 i .. 8 i + 7] is
 * E (inv = 0) or E^-1 (inv = 1), 2x2 complex column-major.  kcoe, vcoe,
 * k12, v12: [nfam] complex (re, im pairs) as the reference's pf_type holds
 * them.  Host arrays in and out; computed on the current device. */
int pqmc_expkv_epsoc(double dtau, int32_t nfam, const double *kcoe, const double *vcoe, const double *k12,
                     const double *v12, int64_t n, const int32_t *fam, const double *x, const int32_t *cross,
                     const double *bcph, int32_t inv, double *out);

/* Kernel timing hook for benchmarks: HIP events bracket every launch of
 * the named kernel class on the handle's stream when enabled; returns the
 * accumulated milliseconds and launch count. kind: 0 = family update,
 * 1 = standalone wrap, 2 = stabilization (whole get_g_scratch).  Inside
 * pqmc_sweep the wraps are fused into the family passes (kind 0). */
int pqmc_timing(pqmc_handle *h, int enable, int kind, double *ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
