/*
 * pqmc_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference projector-DQMC (PQMC) sweep of
 * LEOMMM816/PQMC_allinone, EPSOCZ model.  It is the checker for the HIP
 * library in pqmc_allinone_amd/ and the "port" CPU baseline timed by
 * bench.py.  Nothing in the product links, loads or calls it.
 *
 * Conventions follow the reference: complex(dp) matrices are column-major
 * with leading dimension = rows; `time` and `pf_id` arguments are 1-based
 * exactly as in src/Main_PQMC.f90; sites, bonds and fields are 0-based
 * inside arrays (reference index - 1).
 */
#ifndef PQMC_ORACLE_H
#define PQMC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

typedef struct {
  int L1, L2;        /* lattice cells per direction (lat_size) */
  int Lt;            /* ntime = nint(beta/delt) */
  int ngroup;        /* stabilization interval */
  int ne;            /* nelec */
  double dtau;       /* delt */
  double lam;        /* ep_parameter (V_coe scale) */
  double omega;      /* phonon frequency: M = D*omega^-2 */
  double jump;       /* jump_distance in units of the characteristic length */
  double hop;        /* hop (K_coe scale is -hop) */
  double bc_phase1;  /* BC_phase per direction, in units of pi */
  double bc_phase2;
  double biased;     /* biased_phonon */
  int ncopy;         /* ncopy */
  int tr_weight;     /* TR_weight */
} orc_params;

/* sizes derived by the model setup (src/input.f90:167-205) */
typedef struct {
  int Ns, Ncell, Nb, nfam, npla, nblock, total_width;
  double M, D, char_length, jump_distance;
} orc_dims;

orc_model *orc_create(const orc_params *p);
void orc_destroy(orc_model *m);
void orc_get_dims(const orc_model *m, orc_dims *d);

/* model tables: sites[f][b][4], bfi[f][b], cross[f][b], bcph[f][b][4x4 col-major] */
void orc_get_model(const orc_model *m, int *sites, int *bfi, int *cross, double *bcph);
void orc_get_coeffs(const orc_model *m, double *kcoe, double *vcoe, double *k12, double *v12);

/* fields boson_field(Nb, Lt) column-major */
void orc_set_fields(orc_model *m, const double *x);
void orc_get_fields(const orc_model *m, double *x);

/* trial wave function slater(Ns, ne): QDR -> slater_Q/D/R (Main_PQMC.f90:301-302) */
void orc_set_trial(orc_model *m, const double *slater_c128);

/* expKV tables (mod_phonon_field.f90:423-449, 605-628) */
void orc_init_expkv(orc_model *m);
void orc_get_expkv(const orc_model *m, int pf_id, int pla, double x, int inv, double *out_c128);
/* get_expKV_epsoc (mod_phonon_field.f90:572-604): 2x2, complex pairs */
void orc_expkv_epsoc(double dtau, const double *kcoe, const double *vcoe, const double *k12, const double *v12,
                     const double *ph, double x, int inv, double *out_c128);

/* G and B-string */
void orc_init_g(orc_model *m, int start_time);
/* finite temperature (proj = .false.): init_g and get_g_scratch build
 * Ns x Ns strings and G = (1 + B(tau,0) B(beta,tau))^-1 (g_h = 1 - G) */
void orc_set_ft(orc_model *m, int on);
void orc_get_g(const orc_model *m, double *g_c128);
void orc_set_g(orc_model *m, const double *g_c128);
void orc_get_qstring(const orc_model *m, int block, double *q_c128);
void orc_get_dstring(const orc_model *m, double *d_c128);
void orc_get_lncw(const orc_model *m, double *re_im);

void orc_left_evolve(orc_model *m, int time, double *mat_c128, int n, int pf_id, int two_way);
void orc_right_evolve(orc_model *m, int time, double *mat_c128, int n, int pf_id, int two_way);
/* two-way wrap of the model's own g_h */
void orc_wrap_g(orc_model *m, int time, int pf_id, int forward);

/* one family update: u = 2*npla raw ran2 draws in call order (ran_sysm source, rands) */
void orc_update_phonon_field(orc_model *m, int pf_id, int time, const float *u, uint8_t *acc);

void orc_get_g_scratch(orc_model *m, int time, int forward, double *err_fast);

/* one equal-time measurement (take_measurement + k-space), total_width complex */
void orc_measure(orc_model *m, int time, double *out_c128);

/* one iteration of the Main_PQMC sweep loop body.  u holds 2*Nb*Lt draws in
 * call order.  ntc limits the number of time_count iterations (0 = all).
 * acc (Nb*Lt bytes, may be NULL) receives accept flags in proposal order.
 * If meas_out != NULL, the measurement at time == ntime/2 is taken. */
void orc_sweep(orc_model *m, int forward, const float *u, int ntc, uint8_t *acc,
               double *meas_out, int *n_meas_taken);

void orc_get_counters(const orc_model *m, long *pos_acc, long *neg_acc, double *err_fast_max);

/* the ran2 generator of src/mod_ranstate.f90:171-191, seeded as :267-283 */
typedef struct { int32_t i, j, k, m, n; } orc_ran2;
void orc_ran2_seed(orc_ran2 *s, int32_t seed);
void orc_ran2_fill(orc_ran2 *s, int64_t n, float *out);

/* global moves: cal_new_weight_T0 (src/mod_evolution.f90:359-408) on the
 * current fields (sets ln_cw), and update_global (src/mod_update.f90:251-290)
 * for the kinds the sweep loop calls (src/Main_PQMC.f90:88-94), drawing from
 * rng in the reference's order; returns 1 if the move was accepted. */
enum { ORC_GU_KSPACE = 0, ORC_GU_ROTATE = 1, ORC_GU_KSPACE_TIME = 2 };
typedef struct {
  double distance; /* global_update_distance before the char_length scaling */
} orc_global;
void orc_cal_new_weight(orc_model *m, double *re_im);
/* complex d ln_cw / dx for every field, grad[i + Nb (t-1)], from a backward
 * pass with no updates (the chain ends where a backward sweep leaves it);
 * d expkv / dx for one bond */
void orc_gradient(orc_model *m, double *grad);
void orc_get_dexpkv(const orc_model *m, int pf_id, int pla, double xv, double *out);
int orc_update_global(orc_model *m, int kind, const orc_global *gp, orc_ran2 *rng);
void orc_get_global_counters(const orc_model *m, long *accepted, long *rejected);
/* the proposal of update_global alone (fields and -delt*dE), model left unchanged */
void orc_propose_global(orc_model *m, int kind, const orc_global *gp, orc_ran2 *rng, double *x_out, double *ph_out);

/* dense kernels (exposed for tests) */
void orc_zqdr(int m, int n, double *a_c128, double *r_c128, double *d_c128);
void orc_zldq(int m, int n, double *a_c128, double *l_c128, double *d_c128);
void orc_zinverse(int n, double *a_c128);
void orc_zdet(int n, const double *a_c128, double *det_c128);

#ifdef __cplusplus
}
#endif
#endif
