// C ABI and host orchestration of the projector-DQMC sweep on one MI355X.
// One handle = one Markov chain; every call enqueues on the handle's stream
// and returns without synchronizing unless it hands data back to the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pqmc_hip.h"
#include "pqmc_common.h"
#include "k_pass.h"
#include "k_qrflow.h"

// ---- kernel launchers (k_*.hip) ----
struct prop_args {
  model_t md;
  zd *X;
  int ncols;
  const int *sites;
  const ekv_t *ekv;
  int t_first, nslices, backward, fam_only;
};
extern "C" {
hipError_t pqmc_launch_flush_wrap(const fw_args *a, hipStream_t s);
hipError_t pqmc_launch_propagate(const prop_args *a, hipStream_t s);
hipError_t pqmc_launch_ekv_epsoc(double dtau, const zd *kcoe, const zd *vcoe, const zd *k12, const zd *v12, int nfam,
                                 long n, const int *fam, const double *x, const int *cross, const zd *ph, int inv,
                                 zd *out, hipStream_t s);
hipError_t pqmc_launch_ekv_build(const model_t *md, const double *x, const int *bfi, const int *cross,
                                 const zd *ph, ekv_t *ekv, int t_lo, int t_hi, hipStream_t s);
hipError_t pqmc_zgemm(int opA, int opB, int M, int N, int K, zd alpha, const zd *A, int lda, const zd *B, int ldb,
                      zd beta, zd *C, int ldc, hipStream_t s);
hipError_t pqmc_qr_panel(zd *A, int lda, int m, int j0, int nb, zd *V, int ldv, zd *T, zd *tau, double *vn2,
                         zd *beta_out, hipStream_t s);
hipError_t pqmc_apply_refl(zd *A, int lda, int mp, int ncols, const zd *V, int ldv, const zd *T, int nb,
                           int trans, hipStream_t s);
hipError_t pqmc_eye(zd *Q, int ldq, int m, int n, hipStream_t s);
hipError_t pqmc_qdr_finish(const zd *A, int lda, int n, zd *Rout, zd *D, int conjt, hipStream_t s);
hipError_t pqmc_trsm_upper(const zd *R, int ldr, int n, zd *Z, int ldz, int ncols, zd *work, hipStream_t s);
hipError_t pqmc_lncw(int ne, const zd *betaM, const zd *tauM, const double *vn2M, const zd *dlog, int nblk,
                     int tr_weight, int ncopy, zd *out, const double *acc, unsigned long long *drift_call,
                     unsigned long long *drift_run, int conjq, hipStream_t s);
hipError_t pqmc_dlog(const zd *D, int n, zd *out, hipStream_t s);
hipError_t pqmc_dlog_conj(const zd *D, int n, zd *out, hipStream_t s);
hipError_t pqmc_maxdiff(const zd *a, const zd *b, long n, unsigned long long *out, unsigned long long *run,
                        hipStream_t s);
hipError_t pqmc_conjt(const zd *A, int lda, int m, int n, zd *C, int ldc, hipStream_t s);
hipError_t pqmc_mfma_probe(const double *A, const double *B, double *C, hipStream_t s);
hipError_t pqmc_launch_measure(const model_t *md, const zd *G, const double *x, int time, const zd *phase,
                               zd *out, hipStream_t s);
hipError_t pqmc_ft_rowdiv(zd *X, int n, const zd *D, hipStream_t s);
hipError_t pqmc_ft_prep(const zd *X, int n, const zd *D, int *perm, int *ep, hipStream_t s);
hipError_t pqmc_ft_gather_scale(const zd *X, zd *Xp, int n, const zd *D, const int *perm, const int *ep, int conj,
                                hipStream_t s);
hipError_t pqmc_ft_unscale(zd *T, zd *D, int n, const int *ep, int conjt, hipStream_t s);
hipError_t pqmc_ft_permute(const zd *src, zd *dst, int n, const int *perm, int rows, hipStream_t s);
hipError_t pqmc_ft_mix(zd *M, const zd *QQ, const zd *TT, const zd *Dr, const zd *Dl, int n, hipStream_t s);
hipError_t pqmc_ft_dlog(const zd *Dr, const zd *Dl, const zd *Lr, const zd *Ll, int n, zd *out, hipStream_t s);
hipError_t pqmc_ft_qlog(const zd *tau, const double *vn2, int n, int conj, zd *out, hipStream_t s);
hipError_t pqmc_ft_ones(zd *D, int n, hipStream_t s);
hipError_t pqmc_ft_finish(zd *G, int n, hipStream_t s);
hipError_t pqmc_mx_product(int mode, int M, int N, int K, const zd *P, int lda, const zd *Q, int ldb, zd *C, int ldc,
                           double *scales, hipStream_t s);
}

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(PQMC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

struct qr_ws {  // workspace of one blocked QR
  zd *V = nullptr, *T = nullptr, *tau = nullptr, *beta = nullptr;
  double *vn2 = nullptr;
  // one-launch form: ticket counter, abort word and panel flags; the host
  // counts launches (epoch) and the tickets they took
  unsigned long long *sync = nullptr;
  unsigned long long epoch = 0, tickets = 0;
  unsigned long long ycount = 0;  // the fused launch's column counts so far (per panel)
};

struct ev_pool {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
};

struct pqmc_handle {
  int L1, L2, Ncell, Ns, ne, Nb, npla, Lt, ngroup, nblock, total_width, nbq;
  int nb, nblk, bph, H;  // family pass: bonds per decision block, blocks, bonds per helper, helpers
  int dev = 0;           // device the chain lives on
  model_t md{};
  hipStream_t st = nullptr;
  bool model_set = false, trial_set = false, fields_set = false, g_ready = false;
  int g_at = 0;  // where G sits: 0 not tracked, 1 top (after init_g or a full forward sweep), 2 bottom
  // device model tables
  int *d_sites = nullptr, *d_bfi = nullptr, *d_cross = nullptr, *d_pos = nullptr;
  zd *d_ph = nullptr;
  double *d_x = nullptr;
  ekv_t *d_ekv = nullptr;
  // Green's function and delayed-update factors
  zd *d_G = nullptr, *d_Gn = nullptr, *d_scr = nullptr;
  // B-string
  zd *d_Q = nullptr, *d_D = nullptr, *d_R = nullptr;
  zd *d_dlog = nullptr;  // sum_i log D_string(i, b) per block b
  zd *d_rinv = nullptr;  // inverses of the 16 x 16 diagonal blocks of R_M
  zd *d_slQ = nullptr, *d_slD = nullptr, *d_slR = nullptr;
  zd *d_W = nullptr, *d_L = nullptr, *d_M = nullptr, *d_QM = nullptr, *d_Z = nullptr;
  qr_ws qa, qm;
  // family-pass state
  float *d_u = nullptr;
  size_t u_cap = 0;
  uint8_t *d_acc = nullptr;
  zd *d_slot = nullptr;
  zd *d_tslot = nullptr;
  ps_rec *d_rec = nullptr;
  ps_prop *d_pre = nullptr;  // [2][PS_NB] block-0 proposals handed from one pass of a sweep to the next
  unsigned long long *d_psync = nullptr;
  double *d_lncw_acc = nullptr;
  zd *d_grad = nullptr;  // field gradient, [Lt][Nb]
  zd *d_lncw = nullptr;
  unsigned long long *d_stats = nullptr;
  // drift of the stabilizations, uint64 bits of nonnegative doubles (NaN as
  // +inf): [0] max|G_fast - G_scratch| and [1] |Re(ln_cw - ln_cw_fast)| of
  // the latest get_g_scratch, [2] and [3] their running maxima over every
  // stabilization since the host last read them
  unsigned long long *d_err = nullptr;
  double last_drift[2] = {0.0, 0.0};  // [2] and [3] as of the latest read
  // [0] and [1] of every stabilization of the latest sweep, gradient pass or
  // get_g_scratch call, in order (room for Lt / ngroup + 1)
  unsigned long long *d_hist = nullptr;
  int nhist = 0, hist_cap = 0;
  zd *d_meas = nullptr, *d_kphase = nullptr;
  unsigned long long *d_prof = nullptr;  // family-pass phase timers (debug)
  // global-move weight (allocated on first use): proposed fields and their exponentials
  double *d_gx = nullptr;
  ekv_t *d_gekv = nullptr;
  zd *d_glog = nullptr;  // [nblock + 4] log terms of the weight
  zd *d_gw = nullptr;    // the weight
  zd *d_gft = nullptr;   // finite temperature: two rolling Ns x Ns UDT strings
  double err_fast_max = 1e-6;
  // finite temperature: per block b an Ns x Ns string B = Q D T (right
  // blocks, FQ = Q) or T D Q (left blocks, FQ = Q^H), the scales D, the
  // triangular factor T and log det of the unitary factor
  int ft = 0;
  zd *d_FQ = nullptr, *d_FT = nullptr, *d_FD = nullptr, *d_FL = nullptr;
  zd *d_fw = nullptr;    // five Ns x Ns work matrices
  zd *d_flog = nullptr;  // [4] log terms of ln det(1 + B)
  int *d_fperm = nullptr;  // [2 Ns] pre-pivoting order and scale exponents of a string step
  // precision of the B-string products between checkpoints: 0 FP64, 1 FP32
  // three-term split, 2 BF16x3 (k_mixed.hip); the dense group product P
  // (Ns x Ns), the product's output (Ns x ne) and the split scales
  int prec = 0;
  int qr_mode = 0;  // 0: one-launch QR (k_qr_flow), 1: panel by panel (k_qr_panel + k_apply_refl)
  zd *d_mxP = nullptr, *d_mxX = nullptr;
  double *d_mxs = nullptr;
  qr_ws qf;              // QR workspace for Ns x Ns
  // timing
  bool timing[3] = {false, false, false};
  ev_pool pools[3];
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[3];
};

template <class T>
static hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
}

static hipEvent_t pool_get(ev_pool &p) {
  if (p.used == p.ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p.ev.push_back(e);
  }
  return p.ev[p.used++];
}

struct span_guard {  // HIP events around one kernel launch when timing is on
  pqmc_handle *h;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  span_guard(pqmc_handle *hh, int k) : h(hh), kind(k) {
    if (h->timing[kind]) {
      a = pool_get(h->pools[kind]);
      b = pool_get(h->pools[kind]);
      if (a && b) hipEventRecord(a, h->st);
    }
  }
  ~span_guard() {
    if (a && b) {
      hipEventRecord(b, h->st);
      h->spans[kind].push_back({a, b});
    }
  }
};

// ---------------------------------------------------------------------------

extern "C" const char *pqmc_last_error(void) { return g_err.c_str(); }

extern "C" int pqmc_create(const pqmc_config *cfg, pqmc_handle **out) {
  if (!cfg || !out) return fail(PQMC_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->L1 < 2 || cfg->L2 < 2 || (cfg->L1 * cfg->L2) % 2 || cfg->Lt < 1 || cfg->ngroup < 1 || cfg->ne < 1)
    return fail(PQMC_EINVAL, "invalid sizes");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PQMC_ENODEV, "no HIP device");
  int dev = 0;
  {
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
      return fail(PQMC_ENODEV, "cannot query the current HIP device");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return fail(PQMC_ENODEV, std::string("current device is ") + prop.gcnArchName + ", the library is built for gfx950");
  }
  pqmc_handle *h = new pqmc_handle();
  h->dev = dev;
  h->L1 = cfg->L1;
  h->L2 = cfg->L2;
  h->Ncell = cfg->L1 * cfg->L2;
  h->Ns = 2 * h->Ncell;
  h->ne = cfg->ne;
  h->Nb = 2 * h->Ncell;
  h->npla = h->Ns / 4;
  h->Lt = cfg->Lt;
  h->ngroup = cfg->ngroup;
  h->nblock = cfg->Lt / cfg->ngroup + 2;
  h->total_width = 3 + 19 * 2 * h->Ncell;
  if (h->ne > h->Ns) { delete h; return fail(PQMC_EINVAL, "ne > Ns"); }
  if (h->Ns > 1024) { delete h; return fail(PQMC_EINVAL, "Ns > 1024 not supported"); }
  if (cfg->block_bonds < 0 || cfg->block_bonds > PS_NB) { delete h; return fail(PQMC_EINVAL, "block_bonds out of range"); }
  h->ft = cfg->finite_temperature ? 1 : 0;
  if (h->ft && cfg->Lt % cfg->ngroup) { delete h; return fail(PQMC_EINVAL, "finite temperature needs Lt % ngroup == 0"); }
  h->nb = cfg->block_bonds > 0 ? cfg->block_bonds : PS_NB;
  h->nblk = (h->npla + h->nb - 1) / h->nb;
  {
    // every workgroup of a family pass must be resident at once: one leader
    // plus ceil(npla / bph) helpers
    if (cfg->helper_bonds != 0 && cfg->helper_bonds != 1 && cfg->helper_bonds != 2 && cfg->helper_bonds != 4) {
      delete h;
      return fail(PQMC_EINVAL, "helper_bonds must be 0, 1, 2 or 4");
    }
    h->bph = 0;
    for (int bph = cfg->helper_bonds ? cfg->helper_bonds : 1; bph <= 4 && !h->bph; bph *= 2) {
      int maxg = 0;
      if (pqmc_pass_max_groups(h->Ns, bph, &maxg) != hipSuccess) { delete h; return fail(PQMC_EHIP, "occupancy query failed"); }
      if ((h->npla + bph - 1) / bph + 1 <= maxg) h->bph = bph;
      if (cfg->helper_bonds) break;
    }
    if (!h->bph) { delete h; return fail(PQMC_EINVAL, "lattice too large for a resident family pass"); }
    h->H = (h->npla + h->bph - 1) / h->bph;
  }
  h->nbq = (h->Ns <= 512) ? 16 : 8;
  const int Ns = h->Ns, ne = h->ne, npla = h->npla;
  h->md.Ns = Ns; h->md.ne = ne; h->md.npla = npla; h->md.Nb = h->Nb; h->md.Lt = h->Lt;
  h->md.Ncell = h->Ncell; h->md.L1 = h->L1; h->md.L2 = h->L2;
#define A_(p, n)                                                                   \
  if (dalloc(&(p), (n)) != hipSuccess) {                                           \
    pqmc_destroy(h);                                                               \
    return fail(PQMC_EHIP, "device allocation failed: " #p);                       \
  }
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(PQMC_EHIP, "stream creation failed");
  }
  A_(h->d_sites, 4 * npla * 4);
  A_(h->d_bfi, 4 * npla);
  A_(h->d_cross, 4 * npla);
  A_(h->d_pos, 4 * Ns);
  A_(h->d_ph, 4 * npla * 4);
  A_(h->d_x, (size_t)h->Nb * h->Lt);
  A_(h->d_ekv, (size_t)h->Lt * 4 * npla);
  A_(h->d_G, (size_t)Ns * Ns);
  A_(h->d_Gn, (size_t)Ns * Ns);
  A_(h->d_scr, (size_t)Ns * Ns);
  A_(h->d_Q, (size_t)Ns * ne * h->nblock);
  A_(h->d_D, (size_t)ne * h->nblock);
  A_(h->d_R, (size_t)ne * ne * h->nblock);
  A_(h->d_dlog, (size_t)h->nblock);
  A_(h->d_rinv, (size_t)256 * (((h->ft ? Ns : ne) + 15) / 16));
  A_(h->d_slQ, (size_t)Ns * ne);
  A_(h->d_slD, ne);
  A_(h->d_slR, (size_t)ne * ne);
  A_(h->d_W, (size_t)Ns * ne);
  A_(h->d_L, (size_t)Ns * ne);
  A_(h->d_M, (size_t)ne * ne);
  A_(h->d_QM, (size_t)ne * ne);
  A_(h->d_Z, (size_t)ne * Ns);
  A_(h->qa.V, (size_t)Ns * ne);
  A_(h->qa.T, (size_t)16 * 16 * ((ne + 7) / 8));
  A_(h->qa.tau, ne);
  A_(h->qa.beta, ne);
  A_(h->qa.vn2, ne);
  A_(h->qm.V, (size_t)ne * ne);
  A_(h->qm.T, (size_t)16 * 16 * ((ne + 7) / 8));
  A_(h->qm.tau, ne);
  A_(h->qm.beta, ne);
  A_(h->qm.vn2, ne);
  A_(h->qa.sync, QF_SYNC_WORDS);
  A_(h->qm.sync, QF_SYNC_WORDS);
  A_(h->d_acc, (size_t)h->Nb * h->Lt);
  A_(h->d_slot, pqmc_pass_slot_elems(Ns, h->nblk));
  A_(h->d_tslot, pqmc_pass_tslot_elems(h->nblk));
  A_(h->d_rec, (size_t)h->nblk);
  A_(h->d_pre, 2 * (size_t)PS_NB);
  A_(h->d_psync, pqmc_pass_sync_words(h->nblk));
  A_(h->d_lncw_acc, 1);
  A_(h->d_lncw, 1);
  A_(h->d_stats, 4);
  A_(h->d_err, 4);
  h->hist_cap = h->Lt / h->ngroup + 1;
  A_(h->d_hist, 2 * (size_t)h->hist_cap);
  A_(h->d_meas, h->total_width);
  A_(h->d_kphase, (size_t)h->Ncell * h->Ncell);
  if (h->ft) {
    const size_t nn = (size_t)Ns * Ns;
    A_(h->d_FQ, nn * h->nblock);
    A_(h->d_FT, nn * h->nblock);
    A_(h->d_FD, (size_t)Ns * h->nblock);
    A_(h->d_FL, h->nblock);
    A_(h->d_fw, 5 * nn);
    A_(h->d_flog, 4);
    A_(h->d_fperm, 2 * Ns);
    A_(h->qf.V, nn);
    A_(h->qf.T, (size_t)16 * 16 * ((Ns + 7) / 8));
    A_(h->qf.tau, Ns);
    A_(h->qf.beta, Ns);
    A_(h->qf.vn2, Ns);
    A_(h->qf.sync, QF_SYNC_WORDS);
    if (hipMemsetAsync(h->qf.sync, 0, QF_SYNC_WORDS * sizeof(unsigned long long), h->st) != hipSuccess) {
      pqmc_destroy(h);
      return fail(PQMC_EHIP, "initial memset failed");
    }
  }
#undef A_
  if (hipMemsetAsync(h->d_dlog, 0, h->nblock * sizeof(zd), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_psync, 0, pqmc_pass_sync_words(h->nblk) * sizeof(unsigned long long), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_stats, 0, 4 * sizeof(unsigned long long), h->st) != hipSuccess ||
      hipMemsetAsync(h->qa.sync, 0, QF_SYNC_WORDS * sizeof(unsigned long long), h->st) != hipSuccess ||
      hipMemsetAsync(h->qm.sync, 0, QF_SYNC_WORDS * sizeof(unsigned long long), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_lncw_acc, 0, sizeof(double), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_lncw, 0, sizeof(zd), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_err, 0, 4 * sizeof(unsigned long long), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_acc, 0, (size_t)h->Nb * h->Lt, h->st) != hipSuccess) {
    pqmc_destroy(h);
    return fail(PQMC_EHIP, "initial memset failed");
  }
  // k-space phases exp(i FBZ(k).r_i) (src/mod_lattice.f90:164-205, unit square lattice)
  {
    std::vector<zd> ph((size_t)h->Ncell * h->Ncell);
    const double two_pi = 2.0 * std::acos(-1.0);
    double kx0 = (h->L1 % 2) * two_pi / h->L1 / 2.0 - 0.5 * two_pi;
    double ky0 = (h->L2 % 2) * two_pi / h->L2 / 2.0 - 0.5 * two_pi;
    for (int k = 0; k < h->Ncell; ++k) {
      double f0 = two_pi / h->L1 * (k % h->L1) + kx0, f1 = two_pi / h->L2 * (k / h->L1) + ky0;
      for (int i = 0; i < h->Ncell; ++i) {
        double arg = f0 * (i % h->L1) + f1 * (i / h->L1);
        ph[k + (size_t)h->Ncell * i] = make_double2(std::cos(arg), std::sin(arg));
      }
    }
    if (hipMemcpyAsync(h->d_kphase, ph.data(), sizeof(zd) * ph.size(), hipMemcpyHostToDevice, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess) {
      pqmc_destroy(h);
      return fail(PQMC_EHIP, "phase upload failed");
    }
  }
  *out = h;
  return PQMC_OK;
}

extern "C" int pqmc_destroy(pqmc_handle *h) {
  if (!h) return PQMC_OK;
  if (h->st) hipStreamSynchronize(h->st);
  void *ptrs[] = {h->d_sites, h->d_bfi, h->d_cross, h->d_pos, h->d_ph, h->d_x, h->d_ekv, h->d_G, h->d_Gn, h->d_scr,
                  h->d_Q, h->d_D, h->d_R, h->d_dlog, h->d_rinv, h->d_slQ, h->d_slD, h->d_slR, h->d_W, h->d_L, h->d_M, h->d_QM,
                  h->d_Z, h->qa.V, h->qa.T, h->qa.tau, h->qa.beta, h->qa.vn2, h->qm.V, h->qm.T, h->qm.tau,
                  h->qm.beta, h->qm.vn2, h->d_u, h->d_acc, h->d_slot, h->d_tslot, h->d_rec, h->d_pre, h->d_psync, h->d_lncw_acc,
                  h->d_lncw, h->d_stats, h->d_err, h->d_hist, h->d_meas, h->d_kphase, h->d_prof,
                  h->d_gx, h->d_gekv, h->d_glog, h->d_gw, h->d_FQ, h->d_FT, h->d_FD, h->d_FL, h->d_fw,
                  h->d_flog, h->d_fperm, h->qf.V, h->qf.T, h->qf.tau, h->qf.beta, h->qf.vn2, h->d_gft, h->d_grad,
                  h->d_mxP, h->d_mxX, h->d_mxs, h->qa.sync, h->qm.sync, h->qf.sync};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (auto &pool : h->pools)
    for (auto e : pool.ev) hipEventDestroy(e);
  if (h->st) hipStreamDestroy(h->st);
  delete h;
  return PQMC_OK;
}

extern "C" int pqmc_get_dims(const pqmc_handle *h, int32_t *o) {
  if (!h || !o) return fail(PQMC_EINVAL, "null argument");
  o[0] = h->Ns; o[1] = h->Ncell; o[2] = h->Nb; o[3] = h->npla; o[4] = h->nblock; o[5] = h->total_width;
  return PQMC_OK;
}

extern "C" int pqmc_set_model(pqmc_handle *h, const pqmc_model *m) {
  if (!h || !m || !m->sites || !m->bfi || !m->cross || !m->bcph || !m->kcoe || !m->vcoe || !m->k12 || !m->v12 ||
      !m->bc)
    return fail(PQMC_EINVAL, "null argument");
  const int npla = h->npla, Ns = h->Ns;
  std::vector<int> pos(4 * Ns, -1);
  for (int f = 0; f < 4; ++f)
    for (int q = 0; q < npla * 4; ++q) {
      int s = m->sites[f * npla * 4 + q];
      if (s < 0 || s >= Ns || pos[f * Ns + s] != -1) return fail(PQMC_EINVAL, "bond family does not partition the sites");
      pos[f * Ns + s] = q;
    }
  for (int q = 0; q < 4 * npla; ++q)
    if (m->bfi[q] < 0 || m->bfi[q] >= h->Nb) return fail(PQMC_EINVAL, "field index out of range");
  std::vector<zd> ph(4 * npla * 4);
  for (int fb = 0; fb < 4 * npla; ++fb)
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) {
        const double *p = &m->bcph[2 * (fb * 16 + a + 4 * (2 + c))];
        ph[fb * 4 + a + 2 * c] = make_double2(p[0], p[1]);
      }
  model_t &md = h->md;
  md.dtau = m->dtau; md.M = m->M; md.D = m->D; md.biased = m->biased; md.jump = m->jump_distance;
  md.hop = m->hop; md.lam = m->lam; md.ncopy = m->ncopy; md.tr_weight = m->tr_weight;
  for (int f = 0; f < 4; ++f) {
    md.kcoe[f] = make_double2(m->kcoe[2 * f], m->kcoe[2 * f + 1]);
    md.vcoe[f] = make_double2(m->vcoe[2 * f], m->vcoe[2 * f + 1]);
    for (int q = 0; q < 4; ++q) {
      md.k12[f][q] = make_double2(m->k12[2 * (4 * f + q)], m->k12[2 * (4 * f + q) + 1]);
      md.v12[f][q] = make_double2(m->v12[2 * (4 * f + q)], m->v12[2 * (4 * f + q) + 1]);
    }
  }
  md.bc[0] = make_double2(m->bc[0], m->bc[1]);
  md.bc[1] = make_double2(m->bc[2], m->bc[3]);
  if (md.ncopy < 1) return fail(PQMC_EINVAL, "ncopy < 1");
  HIPCHK(hipMemcpyAsync(h->d_sites, m->sites, sizeof(int) * 4 * npla * 4, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_bfi, m->bfi, sizeof(int) * 4 * npla, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_cross, m->cross, sizeof(int) * 4 * npla, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_pos, pos.data(), sizeof(int) * 4 * Ns, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_ph, ph.data(), sizeof(zd) * ph.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->model_set = true;
  return PQMC_OK;
}

extern "C" int pqmc_set_fields(pqmc_handle *h, const double *x) {
  if (!h || !x) return fail(PQMC_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(h->d_x, x, sizeof(double) * h->Nb * h->Lt, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->fields_set = true;
  h->g_ready = false;
  return PQMC_OK;
}

extern "C" int pqmc_get_fields(pqmc_handle *h, double *x) {
  if (!h || !x) return fail(PQMC_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(x, h->d_x, sizeof(double) * h->Nb * h->Lt, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

// ---------------------------------------------------------------------------
// blocked Householder QR of A (m x n, lda): Q (m x n, ld m) explicit,
// D = diag(R), Rout = D^-1 R (or its conjugate transpose).

static int run_qr(pqmc_handle *h, qr_ws &w, zd *A, int lda, int m, int n, zd *Q, zd *D, zd *Rout, int conjt) {
  const int nb = h->nbq;
  const int np = (n + nb - 1) / nb;
  if (h->qr_mode == 0) {
    // one launch: panel b waits for panels 0..b-1 (k_qrflow.hip)
    qf_args a{};
    a.A = A;
    a.lda = lda;
    a.m = m;
    a.n = n;
    a.nb = nb;
    a.np = np;
    a.Q = Q;
    a.V = w.V;
    a.T = w.T;
    a.tau = w.tau;
    a.beta = w.beta;
    a.vn2 = w.vn2;
    a.sync = w.sync;
    a.timeouts = h->d_stats + 2;
    a.epoch = w.epoch + 1;
    a.ticket_base = w.tickets;
    HIPCHK(pqmc_qr_flow(&a, h->st));
    w.epoch += 1;
    w.tickets += np;
    if (D && Rout) HIPCHK(pqmc_qdr_finish(A, lda, n, Rout, D, conjt, h->st));
    return PQMC_OK;
  }
  for (int p = 0; p < np; ++p) {
    const int j0 = p * nb, nbp = std::min(nb, n - j0);
    HIPCHK(pqmc_qr_panel(A, lda, m, j0, nbp, w.V, m, w.T + (size_t)256 * p, w.tau + j0, w.vn2 + j0, w.beta + j0,
                         h->st));
    if (j0 + nbp < n)
      HIPCHK(pqmc_apply_refl(A + j0 + (size_t)lda * (j0 + nbp), lda, m - j0, n - j0 - nbp, w.V + j0 + (size_t)m * j0,
                             m, w.T + (size_t)256 * p, nbp, 1, h->st));
  }
  if (Q) {
    HIPCHK(pqmc_eye(Q, m, m, n, h->st));
    for (int p = np - 1; p >= 0; --p) {
      const int j0 = p * nb, nbp = std::min(nb, n - j0);
      HIPCHK(pqmc_apply_refl(Q + j0 + (size_t)m * j0, m, m - j0, n - j0, w.V + j0 + (size_t)m * j0, m,
                             w.T + (size_t)256 * p, nbp, 0, h->st));
    }
  }
  if (D && Rout) HIPCHK(pqmc_qdr_finish(A, lda, n, Rout, D, conjt, h->st));
  return PQMC_OK;
}

// The stabilization's two QRs in one launch (k_qr_pair): the fresh string W
// (Ns x ne) with its explicit Q into Qout, and beside it the stored string S,
// so that Y = (Q_full^H S)[0:ne, :]^H = S^H Qout lands in d_M column block by
// column block and is factored there (Q_Y in d_QM, R_Y in d_M) as it lands.
// Forward (fresh right string, S the left one) Y = QL^H QR = M; backward
// (fresh left string, S the right one) Y = M^H.
static int qr_pair(pqmc_handle *h, zd *W, zd *Qout, const zd *S) {
  const int Ns = h->Ns, ne = h->ne, nb = h->nbq, np = (ne + nb - 1) / nb;
  qf_pair_args P{};
  qf_args &a1 = P.a1, &a2 = P.a2;
  a1.A = W;
  a1.lda = Ns;
  a1.m = Ns;
  a1.n = ne;
  a1.nb = nb;
  a1.np = np;
  a1.Q = Qout;
  a1.V = h->qa.V;
  a1.T = h->qa.T;
  a1.tau = h->qa.tau;
  a1.beta = h->qa.beta;
  a1.vn2 = h->qa.vn2;
  a1.sync = h->qa.sync;
  a1.timeouts = h->d_stats + 2;
  a1.epoch = h->qa.epoch + 1;
  a1.ticket_base = h->qa.tickets;
  a2 = a1;
  a2.A = h->d_M;
  a2.lda = ne;
  a2.m = ne;
  a2.Q = h->d_QM;
  a2.V = h->qm.V;
  a2.T = h->qm.T;
  a2.tau = h->qm.tau;
  a2.beta = h->qm.beta;
  a2.vn2 = h->qm.vn2;
  a2.sync = h->qm.sync;
  a2.epoch = h->qm.epoch + 1;
  a2.ticket_base = 0;  // the tickets come from a1's counter
  P.S = S;
  P.ytarget = h->qm.ycount + np;
  HIPCHK(pqmc_qr_pair(&P, h->st));
  h->qa.epoch += 1;
  h->qa.tickets += 3 * (unsigned long long)np;
  h->qm.epoch += 1;
  h->qm.ycount += np;
  return PQMC_OK;
}

// The rest of G after qr_pair: with QLt, QRt the strings such that Y =
// QLt^H QRt (the left and right strings forward, the right and left ones
// backward), Gt = QRt Y^-1 QLt^H is G forward and G^H backward, and
// ln det Y is ln det M, or its conjugate (cal_g_and_det_T0,
// src/mod_evolution.f90:330-357).
static int cal_g_pair(pqmc_handle *h, const zd *QLt, const zd *QRt, zd *Gout, bool backward) {
  const int Ns = h->Ns, ne = h->ne;
  const zd one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  // Z = Q_Y^H QLt^H  (ne x Ns), then R_Y^-1 Z
  HIPCHK(pqmc_zgemm(1, 1, ne, Ns, ne, one, h->d_QM, ne, QLt, Ns, zero, h->d_Z, ne, h->st));
  HIPCHK(pqmc_trsm_upper(h->d_M, ne, ne, h->d_Z, ne, Ns, h->d_rinv, h->st));
  if (backward) {
    HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, ne, one, QRt, Ns, h->d_Z, ne, zero, h->d_scr, Ns, h->st));
    HIPCHK(pqmc_conjt(h->d_scr, Ns, Ns, Ns, Gout, Ns, h->st));
  } else {
    HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, ne, one, QRt, Ns, h->d_Z, ne, zero, Gout, Ns, h->st));
  }
  HIPCHK(pqmc_lncw(ne, h->qm.beta, h->qm.tau, h->qm.vn2, h->d_dlog, h->nblock, h->md.tr_weight, h->md.ncopy,
                   h->d_lncw, h->d_lncw_acc, h->d_err + 1, h->d_err + 3, backward ? 1 : 0, h->st));
  HIPCHK(hipMemsetAsync(h->d_lncw_acc, 0, sizeof(double), h->st));
  return PQMC_OK;
}

// G = QR (QL^H QR)^-1 QL^H with ln_cw (cal_g_and_det_T0, src/mod_evolution.f90:330-357).
// QL is the left B-string stored as lmat^H (Ns x ne); QR is the right one.
// drift: a stabilization (get_g_scratch), whose ln_cw is compared with the
// fast one into the drift words.
static int cal_g(pqmc_handle *h, const zd *QL, const zd *QRm, zd *Gout, bool drift) {
  const int Ns = h->Ns, ne = h->ne;
  const zd one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  // M = QL^H QR
  HIPCHK(pqmc_zgemm(1, 0, ne, ne, Ns, one, QL, Ns, QRm, Ns, zero, h->d_M, ne, h->st));
  int rc = run_qr(h, h->qm, h->d_M, ne, ne, ne, h->d_QM, nullptr, nullptr, 0);
  if (rc) return rc;
  // Z = QM^H QL^H  (ne x Ns)
  HIPCHK(pqmc_zgemm(1, 1, ne, Ns, ne, one, h->d_QM, ne, QL, Ns, zero, h->d_Z, ne, h->st));
  // Z <- R_M^-1 Z
  HIPCHK(pqmc_trsm_upper(h->d_M, ne, ne, h->d_Z, ne, Ns, h->d_rinv, h->st));
  // G = QR Z
  HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, ne, one, QRm, Ns, h->d_Z, ne, zero, Gout, Ns, h->st));
  HIPCHK(pqmc_lncw(ne, h->qm.beta, h->qm.tau, h->qm.vn2, h->d_dlog, h->nblock, h->md.tr_weight, h->md.ncopy,
                   h->d_lncw, h->d_lncw_acc, drift ? h->d_err + 1 : nullptr, h->d_err + 3, 0, h->st));
  // the accepts' log(real(R_elec)) terms since the last stabilization are
  // overwritten with it (src/mod_update.f90:142)
  HIPCHK(hipMemsetAsync(h->d_lncw_acc, 0, sizeof(double), h->st));
  return PQMC_OK;
}

static zd *QS(pqmc_handle *h, int block) { return h->d_Q + (size_t)h->Ns * h->ne * (block - 1); }
static zd *DS(pqmc_handle *h, int block) { return h->d_D + (size_t)h->ne * (block - 1); }
static zd *RS(pqmc_handle *h, int block) { return h->d_R + (size_t)h->ne * h->ne * (block - 1); }

static int propagate(pqmc_handle *h, zd *X, int t_first, int nslices, int backward, const ekv_t *ekv = nullptr,
                     int ncols = 0) {
  prop_args a;
  a.md = h->md;
  a.X = X;
  a.ncols = ncols > 0 ? ncols : h->ne;
  a.sites = h->d_sites;
  a.ekv = ekv ? ekv : h->d_ekv;
  a.t_first = t_first;
  a.nslices = nslices;
  a.backward = backward;
  a.fam_only = -1;
  HIPCHK(pqmc_launch_propagate(&a, h->st));
  return PQMC_OK;
}

// One B-string step X (Ns x ne) <- P X over nslices slices, P the product of
// their family layers (the one-way left_evolve / right_evolve loops of
// init_g_T0 and get_g_scratch_T0): applied layer by layer in FP64, or, in a
// mixed-precision mode, P formed in FP64 (the layers applied to the
// identity) and the product P X taken on FP32 or BF16 MFMA with split
// operands (k_mixed.hip).
static int string_step(pqmc_handle *h, zd *X, int t_first, int nslices, int backward, const ekv_t *ekv = nullptr) {
  if (h->prec == 0) return propagate(h, X, t_first, nslices, backward, ekv);
  const int Ns = h->Ns, ne = h->ne;
  HIPCHK(pqmc_eye(h->d_mxP, Ns, Ns, Ns, h->st));
  int rc = propagate(h, h->d_mxP, t_first, nslices, backward, ekv, Ns);
  if (rc) return rc;
  HIPCHK(pqmc_mx_product(h->prec, Ns, ne, Ns, h->d_mxP, Ns, X, Ns, h->d_mxX, Ns, h->d_mxs, h->st));
  HIPCHK(hipMemcpyAsync(X, h->d_mxX, sizeof(zd) * Ns * ne, hipMemcpyDeviceToDevice, h->st));
  return PQMC_OK;
}

extern "C" int pqmc_set_precision(pqmc_handle *h, int mode) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  if (mode < 0 || mode > 2) return fail(PQMC_EINVAL, "precision mode must be 0 (FP64), 1 (FP32 split) or 2 (BF16x3)");
  if (mode && h->ft) return fail(PQMC_EINVAL, "mixed-precision B-string products are for projector chains");
  if (mode && !h->d_mxP) {
    const size_t Ns = h->Ns;
    HIPCHK(dalloc(&h->d_mxP, Ns * Ns));
    HIPCHK(dalloc(&h->d_mxX, Ns * h->ne));
    HIPCHK(dalloc(&h->d_mxs, Ns + h->ne));
  }
  h->prec = mode;
  return PQMC_OK;
}

extern "C" int pqmc_set_qr_mode(pqmc_handle *h, int mode) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  if (mode < 0 || mode > 1) return fail(PQMC_EINVAL, "QR mode must be 0 (one launch) or 1 (panel by panel)");
  HIPCHK(hipStreamSynchronize(h->st));
  h->qr_mode = mode;
  return PQMC_OK;
}

extern "C" int pqmc_set_trial(pqmc_handle *h, const double *slater) {
  if (!h || !slater) return fail(PQMC_EINVAL, "null argument");
  const size_t n = (size_t)h->Ns * h->ne;
  HIPCHK(hipMemcpyAsync(h->d_W, slater, sizeof(zd) * n, hipMemcpyHostToDevice, h->st));
  int rc = run_qr(h, h->qa, h->d_W, h->Ns, h->Ns, h->ne, h->d_slQ, h->d_slD, h->d_slR, 0);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  h->trial_set = true;
  return PQMC_OK;
}

// standalone two-way wrap of G (wrap 1: E G E^-1, 2: E^-1 G E)
static int wrap_g(pqmc_handle *h, int wrap, int f_wrap, int t_wrap, zd *grad = nullptr) {
  fw_args a{};
  a.md = h->md;
  a.G = h->d_G;
  a.U = nullptr;
  a.VT = nullptr;
  a.rank = nullptr;
  a.pos_ffl = nullptr;
  a.tsites = h->d_sites + (size_t)f_wrap * h->npla * 4;
  a.wrap = wrap;
  a.ekv_w = h->d_ekv + ((size_t)(t_wrap - 1) * 4 + f_wrap) * h->npla;
  if (grad) {
    a.grad = grad;
    a.x = h->d_x;
    a.bfi = h->d_bfi;
    a.cross = h->d_cross;
    a.ph = h->d_ph;
    a.f_wrap = f_wrap;
    a.t_wrap = t_wrap;
  }
  span_guard sg(h, 1);
  HIPCHK(pqmc_launch_flush_wrap(&a, h->st));
  return PQMC_OK;
}

// ---------------------------------------------------------------------------
// Finite temperature (init_g_FT / get_g_scratch_FT / cal_g_and_det_FT,
// src/mod_evolution.f90:411-785).  The reference's branch is marked as not in
// use and is not self-consistent (its init forms (1 + B B^H)^-1 where the
// commented call beside it forms (1 + B)^-1, and its left string keeps the
// unitary and triangular factors in swapped roles); this computes what the
// branch is built for, with its structure: G = (1 + B(beta,0))^-1 from a
// right string B(t,0) = Q_r D_r T_r and a left string B(beta,t) = T_l D_l Q_l,
// each advanced by one column- (row-) scaled QR (LQ) per ngroup slices, the
// scales split at |d| = 1; g_h = 1 - G and ln_cw = ln det(1 + B(beta,0)).
// A right block keeps Q_r, a left block keeps Q_l^H, so that both strings are
// advanced by the same one-way propagation and column-pivot-free QR as the
// projector strings.

static zd *FQS(pqmc_handle *h, int b) { return h->d_FQ + (size_t)h->Ns * h->Ns * (b - 1); }
static zd *FTS(pqmc_handle *h, int b) { return h->d_FT + (size_t)h->Ns * h->Ns * (b - 1); }
static zd *FDS(pqmc_handle *h, int b) { return h->d_FD + (size_t)h->Ns * (b - 1); }
static zd *FW(pqmc_handle *h, int k) { return h->d_fw + (size_t)h->Ns * h->Ns * k; }

static int ft_identity(pqmc_handle *h, int b) {
  const int Ns = h->Ns;
  HIPCHK(pqmc_eye(FQS(h, b), Ns, Ns, Ns, h->st));
  HIPCHK(pqmc_eye(FTS(h, b), Ns, Ns, Ns, h->st));
  HIPCHK(pqmc_ft_ones(FDS(h, b), Ns, h->st));
  HIPCHK(hipMemsetAsync(h->d_FL + (b - 1), 0, sizeof(zd), h->st));
  return PQMC_OK;
}

// One string step over slices t0..t1 with the exponentials ekv.
//  right: B Q D P = Q' D' T'';            T' = T'' P^T T
//  left (kept as conjugate transposes): (P^T D Q B)^H = Q'^H D'^H L''^H;
//         T' = T P L''  (L'' unit lower)
// Scales reach exp(400) at beta = 100: the QR runs on X Dh with
// D = Dh 2^e, |Dh| in [0.5, 1), and the powers of two go back exactly into
// D' and T'' (Householder QR is invariant under power-of-two column
// scaling).  P orders the columns of X D by decreasing norm before the
// unpivoted QR (pre-pivoting, Bai, Lee, Li, Xu, Linear Algebra Appl. 435
// (2011) 659-673), so that D' is graded and T'' bounded, as the sorted
// scales of the reference's SVD-based udv (src/mod_matrixlib.f90:704-743)
// are.  L' = log det of the new unitary factor.
struct ft_udt {
  zd *Q, *D, *T, *L;
};

static int ft_step(pqmc_handle *h, ft_udt in, ft_udt out, int t0, int t1, const ekv_t *ekv, int left) {
  const int Ns = h->Ns;
  const zd one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  zd *X = FW(h, 0), *Tp = FW(h, 1), *Xp = FW(h, 2), *Tg = FW(h, 3);
  int *perm = h->d_fperm, *ep = h->d_fperm + Ns;
  HIPCHK(hipMemcpyAsync(X, in.Q, sizeof(zd) * Ns * Ns, hipMemcpyDeviceToDevice, h->st));
  int rc = propagate(h, X, left ? t1 : t0, t1 - t0 + 1, left, ekv, Ns);
  if (rc) return rc;
  HIPCHK(pqmc_ft_prep(X, Ns, in.D, perm, ep, h->st));
  HIPCHK(pqmc_ft_gather_scale(X, Xp, Ns, in.D, perm, ep, left, h->st));
  rc = run_qr(h, h->qf, Xp, Ns, Ns, Ns, out.Q, out.D, Tp, left);
  if (rc) return rc;
  HIPCHK(pqmc_ft_qlog(h->qf.tau, h->qf.vn2, Ns, left, out.L, h->st));
  HIPCHK(pqmc_ft_unscale(Tp, out.D, Ns, ep, left, h->st));
  HIPCHK(pqmc_ft_permute(in.T, Tg, Ns, perm, !left, h->st));
  if (left)
    HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, Ns, one, Tg, Ns, Tp, Ns, zero, out.T, Ns, h->st));
  else
    HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, Ns, one, Tp, Ns, Tg, Ns, zero, out.T, Ns, h->st));
  return PQMC_OK;
}

static ft_udt ft_block(pqmc_handle *h, int b) { return {FQS(h, b), FDS(h, b), FTS(h, b), h->d_FL + (b - 1)}; }

// right string of block b from block b-1 and slices t0..t1
static int ft_right(pqmc_handle *h, int b, int t0, int t1) {
  return ft_step(h, ft_block(h, b - 1), ft_block(h, b), t0, t1, h->d_ekv, 0);
}

// left string of block b from block b+1 and slices t1 down to t0
static int ft_left(pqmc_handle *h, int b, int t0, int t1) {
  return ft_step(h, ft_block(h, b + 1), ft_block(h, b), t0, t1, h->d_ekv, 1);
}

// M = Db_r^-1 Q_r^H Q_l^H Db_l^-1 + Ds_r T_r T_l Ds_l in FW(2), factored by
// a QR (explicit Q into QM unless null), and
// ln det(1 + B) = ln det M + ln det Q_r + ln det Q_l + sum ln Db_r + sum ln Db_l
// into *lncw (cal_g_and_det_FT, src/mod_evolution.f90:494-648)
static int ft_weight(pqmc_handle *h, ft_udt r, ft_udt l, zd *QM, zd *lncw, bool drift = false) {
  const int Ns = h->Ns;
  const zd one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  zd *QQ = FW(h, 0), *TT = FW(h, 1), *M = FW(h, 2);
  HIPCHK(pqmc_zgemm(1, 0, Ns, Ns, Ns, one, r.Q, Ns, l.Q, Ns, zero, QQ, Ns, h->st));
  HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, Ns, one, r.T, Ns, l.T, Ns, zero, TT, Ns, h->st));
  HIPCHK(pqmc_ft_mix(M, QQ, TT, r.D, l.D, Ns, h->st));
  HIPCHK(pqmc_ft_dlog(r.D, l.D, r.L, l.L, Ns, h->d_flog, h->st));
  int rc = run_qr(h, h->qf, M, Ns, Ns, Ns, QM, nullptr, nullptr, 0);
  if (rc) return rc;
  HIPCHK(pqmc_lncw(Ns, h->qf.beta, h->qf.tau, h->qf.vn2, h->d_flog, 4, h->md.tr_weight, h->md.ncopy, lncw,
                   h->d_lncw_acc, drift ? h->d_err + 1 : nullptr, h->d_err + 3, 0, h->st));
  return PQMC_OK;
}

// G = (1 + Q_r D_r T_r T_l D_l Q_l)^-1 = Q_l^H Db_l^-1 M^-1 Db_r^-1 Q_r^H;
// Gout = 1 - G; ln_cw = ln det(1 + B); drift as in cal_g
static int ft_cal_g(pqmc_handle *h, int br, int bl, zd *Gout, bool drift) {
  const int Ns = h->Ns;
  const zd one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  zd *TT = FW(h, 1), *M = FW(h, 2), *QM = FW(h, 3), *Z = FW(h, 4);
  const zd *Qr = FQS(h, br), *Ul = FQS(h, bl), *Dr = FDS(h, br), *Dl = FDS(h, bl);
  int rc = ft_weight(h, ft_block(h, br), ft_block(h, bl), QM, h->d_lncw, drift);
  if (rc) return rc;
  // TT <- Db_r^-1 Q_r^H;  Z = R_M^-1 Q_M^H TT;  Z <- Db_l^-1 Z;  G = Q_l^H Z
  HIPCHK(pqmc_conjt(Qr, Ns, Ns, Ns, TT, Ns, h->st));
  HIPCHK(pqmc_ft_rowdiv(TT, Ns, Dr, h->st));
  HIPCHK(pqmc_zgemm(1, 0, Ns, Ns, Ns, one, QM, Ns, TT, Ns, zero, Z, Ns, h->st));
  HIPCHK(pqmc_trsm_upper(M, Ns, Ns, Z, Ns, Ns, h->d_rinv, h->st));
  HIPCHK(pqmc_ft_rowdiv(Z, Ns, Dl, h->st));
  HIPCHK(pqmc_zgemm(0, 0, Ns, Ns, Ns, one, Ul, Ns, Z, Ns, zero, Gout, Ns, h->st));
  HIPCHK(pqmc_ft_finish(Gout, Ns, h->st));
  HIPCHK(hipMemsetAsync(h->d_lncw_acc, 0, sizeof(double), h->st));
  return PQMC_OK;
}

// init_g_FT: the identity at blocks 1 and nblock, right strings of blocks
// 2..nblock-1, G at the end of the slices
static int ft_init(pqmc_handle *h) {
  const int ng = h->ngroup, nb = h->nblock;
  int rc = ft_identity(h, 1);
  if (!rc) rc = ft_identity(h, nb);
  for (int b = 2; b <= nb - 1 && !rc; ++b) rc = ft_right(h, b, (b - 2) * ng + 1, (b - 1) * ng);
  if (!rc) rc = ft_cal_g(h, nb - 1, nb, h->d_G, false);
  return rc;
}

// get_g_scratch_FT (src/mod_evolution.f90:651-740): the new G into d_Gn
static int ft_scratch(pqmc_handle *h, int time, int forward) {
  const int ng = h->ngroup;
  if (forward) {
    const int block = time / ng + 1;
    if (block - 1 < 1 || block + 1 > h->nblock) return fail(PQMC_EINVAL, "stabilization block out of range");
    int rc = ft_right(h, block, time - ng, time - 1);
    return rc ? rc : ft_cal_g(h, block, block + 1, h->d_Gn, true);
  }
  int block = time / ng + 2;
  if (ng == 1) block -= 1;
  if (block - 1 < 1 || block + 1 > h->nblock) return fail(PQMC_EINVAL, "stabilization block out of range");
  int rc = ft_left(h, block, time, time + ng - 1);
  return rc ? rc : ft_cal_g(h, block - 1, block, h->d_Gn, true);
}

extern "C" int pqmc_init_g(pqmc_handle *h, int start_time) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  if (!h->model_set || !h->fields_set || (!h->trial_set && !h->ft))
    return fail(PQMC_EINVAL, h->ft ? "model and fields must be set" : "model, trial and fields must be set");
  if (start_time < 1 || start_time > h->Lt) return fail(PQMC_EINVAL, "start_time out of range");
  const int Ns = h->Ns, ne = h->ne, ng = h->ngroup, nb = h->nblock;
  HIPCHK(pqmc_launch_ekv_build(&h->md, h->d_x, h->d_bfi, h->d_cross, h->d_ph, h->d_ekv, 1, h->Lt, h->st));
  if (h->ft) {
    int rc = ft_init(h);
    if (rc) return rc;
    unsigned long long stats[4];
    HIPCHK(hipMemcpyAsync(stats, h->d_stats, sizeof(stats), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (stats[2]) return fail(PQMC_ESYNC, "device-side wait expired in a QR");
    h->g_ready = true;
    h->g_at = 1;
    return PQMC_OK;
  }
  const size_t qb = sizeof(zd) * Ns * ne;
  HIPCHK(hipMemcpyAsync(QS(h, 1), h->d_slQ, qb, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(QS(h, nb), h->d_slQ, qb, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(DS(h, 1), h->d_slD, sizeof(zd) * ne, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(DS(h, nb), h->d_slD, sizeof(zd) * ne, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(pqmc_dlog(DS(h, 1), ne, h->d_dlog, h->st));
  HIPCHK(pqmc_dlog(DS(h, nb), ne, h->d_dlog + nb - 1, h->st));
  HIPCHK(hipMemcpyAsync(RS(h, 1), h->d_slR, sizeof(zd) * ne * ne, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(RS(h, nb), h->d_slR, sizeof(zd) * ne * ne, hipMemcpyDeviceToDevice, h->st));
  const int start_block = start_time / ng + 1;
  const int stb = (start_block - 1) * ng + 1;
  HIPCHK(hipMemcpyAsync(h->d_W, QS(h, start_block), qb, hipMemcpyDeviceToDevice, h->st));
  const zd *rmat = h->d_W;
  int p = stb;
  while (p <= h->Lt) {
    int pend = std::min(h->Lt, ((p - 1) / ng + 1) * ng);
    int rc = string_step(h, h->d_W, p, pend - p + 1, 0);
    if (rc) return rc;
    rmat = h->d_W;
    if (pend % ng == 0) {
      const int block = (pend - 1) / ng + 2;
      rc = run_qr(h, h->qa, h->d_W, Ns, Ns, ne, QS(h, block), DS(h, block), RS(h, block), 0);
      if (rc) return rc;
      HIPCHK(pqmc_dlog(DS(h, block), ne, h->d_dlog + block - 1, h->st));
      rmat = QS(h, block);
      if (pend < h->Lt) HIPCHK(hipMemcpyAsync(h->d_W, rmat, qb, hipMemcpyDeviceToDevice, h->st));
    }
    p = pend + 1;
  }
  int rc = cal_g(h, QS(h, nb), rmat, h->d_G, false);
  if (rc) return rc;
  unsigned long long stats[4];
  HIPCHK(hipMemcpyAsync(stats, h->d_stats, sizeof(stats), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (stats[2]) return fail(PQMC_ESYNC, "device-side wait expired in a QR");
  h->g_ready = true;
  h->g_at = 1;
  return PQMC_OK;
}

static int ensure_u(pqmc_handle *h, size_t n) {
  if (n <= h->u_cap) return PQMC_OK;
  if (h->d_u) hipFree(h->d_u);
  h->d_u = nullptr;
  HIPCHK(dalloc(&h->d_u, n));
  h->u_cap = n;
  return PQMC_OK;
}

// One family pass (update, optionally fused with the family's two-way wrap:
// wrap 2 = E^-1 G E before the update, 1 = E G E^-1 after it); d_u points at
// its 2*npla draws on the device.
// A pass of a sweep may take block 0's proposals from the pass before it
// (pre) and form those of the pass after it (next: family, slice, draws).
struct pass_link {
  const ps_prop *pre = nullptr;
  ps_prop *nxt = nullptr;
  int nf = 0, nt = 0;
  const float *nu = nullptr;
};

static int launch_pass(pqmc_handle *h, int f, int t, int wrap, const float *d_u, uint8_t *d_acc,
                       const pass_link &lk = pass_link()) {
  ps_args a;
  a.md = h->md;
  a.f = f;
  a.t = t;
  a.wrap = wrap;
  a.nb = h->nb;
  a.nblk = h->nblk;
  a.bph = h->bph;
  a.H = h->H;
  a.fsites = h->d_sites + (size_t)f * h->npla * 4;
  a.fbfi = h->d_bfi + (size_t)f * h->npla;
  a.fcross = h->d_cross + (size_t)f * h->npla;
  a.fph = h->d_ph + (size_t)f * h->npla * 4;
  a.x = h->d_x;
  a.ekv = h->d_ekv;
  a.u = d_u;
  a.G = h->d_G;
  a.slot = h->d_slot;
  a.lds = pqmc_pass_slot_ld(h->Ns);
  a.tslot = h->d_tslot;
  a.rec = h->d_rec;
  a.sync = h->d_psync;
  a.pre = lk.pre;
  a.nxt = lk.nxt;
  a.nx_f = lk.nf;
  a.nx_t = lk.nt;
  a.nx_u = lk.nu;
  a.nx_fbfi = h->d_bfi + (size_t)lk.nf * h->npla;
  a.nx_fcross = h->d_cross + (size_t)lk.nf * h->npla;
  a.nx_fph = h->d_ph + (size_t)lk.nf * h->npla * 4;
  a.acc = d_acc;
  a.lncw_acc = h->d_lncw_acc;
  a.stats = h->d_stats;
  a.prof = h->d_prof;
  span_guard sg(h, 0);
  HIPCHK(pqmc_launch_pass(&a, h->st));
  return PQMC_OK;
}

extern "C" int pqmc_update_family(pqmc_handle *h, int pf_id, int time, const float *u, uint8_t *acc) {
  if (!h || !u) return fail(PQMC_EINVAL, "null argument");
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  if (pf_id < 1 || pf_id > 4 || time < 1 || time > h->Lt) return fail(PQMC_EINVAL, "pf_id or time out of range");
  int rc = ensure_u(h, 2 * (size_t)h->npla);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h->d_u, u, sizeof(float) * 2 * h->npla, hipMemcpyHostToDevice, h->st));
  rc = launch_pass(h, pf_id - 1, time, 0, h->d_u, h->d_acc);
  if (rc) return rc;
  if (acc) {
    HIPCHK(hipMemcpyAsync(acc, h->d_acc, h->npla, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  return PQMC_OK;
}

extern "C" int pqmc_evolve_g(pqmc_handle *h, int time, int pf_id, int forward) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  if (pf_id < 1 || pf_id > 4 || time < 1 || time > h->Lt) return fail(PQMC_EINVAL, "pf_id or time out of range");
  h->g_at = 0;
  return wrap_g(h, forward ? 1 : 2, pf_id - 1, time);
}

extern "C" int pqmc_evolve_matrix(pqmc_handle *h, int time, int pf_id, int left, int two_way, double *mat, int n) {
  if (!h || !mat) return fail(PQMC_EINVAL, "null argument");
  if (pf_id < 1 || pf_id > 4 || time < 1 || time > h->Lt || n < 1) return fail(PQMC_EINVAL, "argument out of range");
  if (two_way && n != h->Ns) return fail(PQMC_EINVAL, "two-way evolve needs n == Ns");
  if (n > h->Ns) return fail(PQMC_EINVAL, "n > Ns");
  const int Ns = h->Ns;
  // ensure the expKV table exists (init_g builds it)
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  int rc = PQMC_OK;
  if (two_way) {
    // run the two-way wrap on a scratch copy in d_Gn
    HIPCHK(hipMemcpyAsync(h->d_Gn, mat, sizeof(zd) * Ns * Ns, hipMemcpyHostToDevice, h->st));
    fw_args a{};
    a.md = h->md;
    a.G = h->d_Gn;
    a.U = a.VT = nullptr;
    a.rank = nullptr;
    a.pos_ffl = nullptr;
    a.tsites = h->d_sites + (size_t)(pf_id - 1) * h->npla * 4;
    a.wrap = left ? 1 : 2;
    a.ekv_w = h->d_ekv + ((size_t)(time - 1) * 4 + (pf_id - 1)) * h->npla;
    HIPCHK(pqmc_launch_flush_wrap(&a, h->st));
    HIPCHK(hipMemcpyAsync(mat, h->d_Gn, sizeof(zd) * Ns * Ns, hipMemcpyDeviceToHost, h->st));
  } else {
    // one family: left: mat(Ns,n) <- E mat; right: mat(n,Ns) <- mat E = (E mat^H)^H
    // (E is Hermitian)
    zd *X = h->d_Gn;
    if (left) {
      HIPCHK(hipMemcpyAsync(X, mat, sizeof(zd) * Ns * n, hipMemcpyHostToDevice, h->st));
    } else {
      HIPCHK(hipMemcpyAsync(h->d_scr, mat, sizeof(zd) * Ns * n, hipMemcpyHostToDevice, h->st));
      HIPCHK(pqmc_conjt(h->d_scr, n, n, Ns, X, Ns, h->st));
    }
    prop_args a;
    a.md = h->md;
    a.X = X;
    a.ncols = n;
    a.sites = h->d_sites;
    a.ekv = h->d_ekv;
    a.t_first = time;
    a.nslices = 1;
    a.backward = 0;
    a.fam_only = pf_id - 1;
    HIPCHK(pqmc_launch_propagate(&a, h->st));
    if (left) {
      HIPCHK(hipMemcpyAsync(mat, X, sizeof(zd) * Ns * n, hipMemcpyDeviceToHost, h->st));
    } else {
      HIPCHK(pqmc_conjt(X, Ns, Ns, n, h->d_scr, n, h->st));
      HIPCHK(hipMemcpyAsync(mat, h->d_scr, sizeof(zd) * Ns * n, hipMemcpyDeviceToHost, h->st));
    }
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

// get_g_scratch_T0 (src/mod_evolution.f90:113-195)
static int g_scratch(pqmc_handle *h, int time, int forward) {
  const int Ns = h->Ns, ne = h->ne, ng = h->ngroup;
  int rc = PQMC_OK;
  span_guard sg(h, 2);
  // this stabilization's drift words; the running maxima stay
  HIPCHK(hipMemsetAsync(h->d_err, 0, 2 * sizeof(unsigned long long), h->st));
  const size_t qb = sizeof(zd) * Ns * ne;
  const zd *lmatH = nullptr, *rmat = nullptr;
  bool paired = false;  // the two QRs ran as one launch (the default one-launch QR mode)
  if (h->ft) {
    rc = ft_scratch(h, time, forward);
    if (rc) return rc;
  } else if (forward) {
    const int block = time / ng + 1;
    if (block - 1 < 1 || block + 1 > h->nblock) return fail(PQMC_EINVAL, "stabilization block out of range");
    HIPCHK(hipMemcpyAsync(h->d_W, QS(h, block - 1), qb, hipMemcpyDeviceToDevice, h->st));
    rc = string_step(h, h->d_W, time - ng, ng, 0);
    if (rc) return rc;
    lmatH = QS(h, block + 1);
    rmat = QS(h, block);
    if (h->qr_mode == 0) {
      rc = qr_pair(h, h->d_W, QS(h, block), lmatH);
      if (rc) return rc;
      HIPCHK(pqmc_qdr_finish(h->d_W, Ns, ne, RS(h, block), DS(h, block), 0, h->st));
      HIPCHK(pqmc_dlog(DS(h, block), ne, h->d_dlog + block - 1, h->st));
      rc = cal_g_pair(h, lmatH, rmat, h->d_Gn, false);
      if (rc) return rc;
      paired = true;
    } else {
      rc = run_qr(h, h->qa, h->d_W, Ns, Ns, ne, QS(h, block), DS(h, block), RS(h, block), 0);
      if (rc) return rc;
      HIPCHK(pqmc_dlog(DS(h, block), ne, h->d_dlog + block - 1, h->st));
    }
  } else {
    int block = time / ng + 2;
    if (ng == 1) block -= 1;
    if (block - 1 < 1 || block + 1 > h->nblock) return fail(PQMC_EINVAL, "stabilization block out of range");
    HIPCHK(hipMemcpyAsync(h->d_W, QS(h, block + 1), qb, hipMemcpyDeviceToDevice, h->st));
    rc = string_step(h, h->d_W, time + ng - 1, ng, 1);
    if (rc) return rc;
    lmatH = QS(h, block);
    rmat = QS(h, block - 1);
    if (h->qr_mode == 0) {
      rc = qr_pair(h, h->d_W, QS(h, block), rmat);
      if (rc) return rc;
      HIPCHK(pqmc_qdr_finish(h->d_W, Ns, ne, RS(h, block), DS(h, block), 1, h->st));
      HIPCHK(pqmc_dlog(DS(h, block), ne, h->d_dlog + block - 1, h->st));
      rc = cal_g_pair(h, rmat, lmatH, h->d_Gn, true);
      if (rc) return rc;
      paired = true;
    } else {
      rc = run_qr(h, h->qa, h->d_W, Ns, Ns, ne, QS(h, block), DS(h, block), RS(h, block), 1);
      if (rc) return rc;
      HIPCHK(pqmc_dlog(DS(h, block), ne, h->d_dlog + block - 1, h->st));
    }
  }
  if (!h->ft && !paired) {
    rc = cal_g(h, lmatH, rmat, h->d_Gn, true);
    if (rc) return rc;
  }
  HIPCHK(pqmc_maxdiff(h->d_G, h->d_Gn, (long)Ns * Ns, h->d_err, h->d_err + 2, h->st));
  if (h->nhist < h->hist_cap)
    HIPCHK(hipMemcpyAsync(h->d_hist + 2 * h->nhist++, h->d_err, 2 * sizeof(unsigned long long),
                          hipMemcpyDeviceToDevice, h->st));
  std::swap(h->d_G, h->d_Gn);
  return PQMC_OK;
}

// Reads the drift words once (after a sweep, a gradient pass or one
// get_g_scratch): every stabilization since the previous read is in the
// running maxima, which are cleared for the next ones.  err_fast_max keeps
// the largest G or Re ln_cw drift, as get_g_scratch_T0 does
// (src/mod_evolution.f90:179-193); a non-finite one fails the call.
static int read_err(pqmc_handle *h, double *err) {
  unsigned long long bits[4] = {0, 0, 0, 0}, stats[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(bits, h->d_err, sizeof(bits), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(stats, h->d_stats, sizeof(stats), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemsetAsync(h->d_err + 2, 0, 2 * sizeof(unsigned long long), h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (stats[2]) return fail(PQMC_ESYNC, "device-side wait expired in a family pass or a QR");
  double e[4];
  std::memcpy(e, bits, sizeof(e));
  if (err) *err = e[0];
  h->last_drift[0] = e[2];
  h->last_drift[1] = e[3];
  for (int i = 2; i < 4; ++i)
    if (e[i] > h->err_fast_max) h->err_fast_max = e[i];
  if (!std::isfinite(e[2]))
    return fail(PQMC_ENUMERIC, "non-finite Green's function at a stabilization");
  if (!std::isfinite(e[3])) return fail(PQMC_ENUMERIC, "non-finite log-weight at a stabilization");
  return PQMC_OK;
}

extern "C" int pqmc_g_scratch(pqmc_handle *h, int time, int forward, double *err_fast) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  h->nhist = 0;
  int rc = g_scratch(h, time, forward);
  if (rc) return rc;
  return read_err(h, err_fast);
}

extern "C" int pqmc_measure(pqmc_handle *h, int time, double *out) {
  if (!h || !out) return fail(PQMC_EINVAL, "null argument");
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  if (time < 1 || time >= h->Lt) return fail(PQMC_EINVAL, "measurement time must be in [1, Lt-1]");
  int rc = PQMC_OK;
  HIPCHK(pqmc_launch_measure(&h->md, h->d_G, h->d_x, time, h->d_kphase, h->d_meas, h->st));
  HIPCHK(hipMemcpyAsync(out, h->d_meas, sizeof(zd) * h->total_width, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

extern "C" int pqmc_sweep(pqmc_handle *h, int forward, const float *u, int ntc, int measure, double *meas_out) {
  if (!h || !u) return fail(PQMC_EINVAL, "null argument");
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  const int Lt = h->Lt, npla = h->npla;
  if (ntc <= 0 || ntc > Lt) ntc = Lt;
  const size_t nu = 2 * (size_t)h->Nb * Lt;
  int rc = ensure_u(h, nu);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h->d_u, u, sizeof(float) * nu, hipMemcpyHostToDevice, h->st));
  size_t off = 0;
  bool measured = false;
  h->nhist = 0;
  // passes in sweep order; each hands block 0's proposals of the next one on
  int ip = 0;
  const int npass = 4 * ntc;
  auto pass = [&](int f, int time, int wrap) -> int {
    pass_link lk;
    if (ip > 0) lk.pre = h->d_pre + (size_t)(ip & 1) * PS_NB;
    if (ip + 1 < npass) {
      lk.nxt = h->d_pre + (size_t)((ip + 1) & 1) * PS_NB;
      lk.nf = forward ? (f + 1) % 4 : (f + 3) % 4;
      lk.nt = forward ? (f == 3 ? time + 1 : time) : (f == 0 ? time - 1 : time);
      lk.nu = h->d_u + off + 2 * (size_t)npla;
    }
    const int r = launch_pass(h, f, time, wrap, h->d_u + off, h->d_acc + off / 2, lk);
    off += 2 * (size_t)npla;
    ++ip;
    return r;
  };
  for (int tc = 1; tc <= ntc; ++tc) {
    int time;
    if (!forward) {
      time = Lt - tc + 1;
      for (int f = 3; f >= 0; --f) {
        rc = pass(f, time, 2);
        if (rc) return rc;
      }
    } else {
      time = tc;
      for (int f = 0; f < 4; ++f) {
        rc = pass(f, time, 1);
        if (rc) return rc;
      }
    }
    if (tc % h->ngroup == 0) {
      rc = g_scratch(h, forward ? time + 1 : time, forward);
      if (rc) return rc;
    }
    if (measure && time == Lt / 2) {
      HIPCHK(pqmc_launch_measure(&h->md, h->d_G, h->d_x, time, h->d_kphase, h->d_meas, h->st));
      measured = true;
    }
  }
  if (measured && meas_out)
    HIPCHK(hipMemcpyAsync(meas_out, h->d_meas, sizeof(zd) * h->total_width, hipMemcpyDeviceToHost, h->st));
  if (ntc >= h->ngroup) {
    rc = read_err(h, nullptr);
    if (rc) return rc;
  }
  unsigned long long stats[4];
  HIPCHK(hipMemcpyAsync(stats, h->d_stats, sizeof(stats), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (stats[2]) return fail(PQMC_ESYNC, "device-side wait expired in a family pass or a QR");
  h->g_at = ntc == Lt ? (forward ? 1 : 2) : 0;
  return PQMC_OK;
}

// Field gradient of ln_cw (no counterpart in the reference): a backward pass
// without updates; each family's two-way wrap E^-1 G E first adds the
// family's gradient from the G(S, S) blocks it holds (k_flush_wrap's diagonal
// tiles), and G is recomputed every ngroup slices as in a backward sweep.
// acc: device accumulator of Nb*Lt complex values, cleared here.
static int gradient_pass(pqmc_handle *h, zd *acc, double *grad) {
  const int Lt = h->Lt;
  const size_t n = (size_t)h->Nb * Lt;
  HIPCHK(hipMemsetAsync(acc, 0, sizeof(zd) * n, h->st));
  int rc = PQMC_OK;
  h->nhist = 0;
  for (int tc = 1; tc <= Lt; ++tc) {
    const int time = Lt - tc + 1;
    for (int f = 3; f >= 0; --f) {
      rc = wrap_g(h, 2, f, time, acc);
      if (rc) return rc;
    }
    if (tc % h->ngroup == 0) {
      rc = g_scratch(h, time, 0);
      if (rc) return rc;
    }
  }
  if (Lt >= h->ngroup) {
    rc = read_err(h, nullptr);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpyAsync(grad, acc, sizeof(zd) * n, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->g_at = 2;
  return PQMC_OK;
}

static int gradient_check(pqmc_handle *h, double *grad) {
  if (!h || !grad) return fail(PQMC_EINVAL, "null argument");
  if (!h->g_ready) return fail(PQMC_EINVAL, "init_g has not run");
  if (h->g_at == 2)
    return fail(PQMC_EINVAL, "G is at the bottom (after a backward sweep or a gradient pass): the gradient pass "
                             "starts where a backward sweep starts, after init_g or a forward sweep");
  return PQMC_OK;
}

extern "C" int pqmc_gradient(pqmc_handle *h, double *grad) {
  int rc = gradient_check(h, grad);
  if (rc) return rc;
  if (!h->d_grad) HIPCHK(dalloc(&h->d_grad, (size_t)h->Nb * h->Lt));
  return gradient_pass(h, h->d_grad, grad);
}

// The same pass for callers that may not allocate device memory: the
// accumulator is the library's fixed device array `scratch` (16 MiB, one per
// device, held until the process ends) instead of a per-chain buffer.  Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the device (taken once per device, with that device current); the array's
// name is a host-side stand-in.  Calls on one device share the array, so
// calls that overlap on a device race.
#define PQMC_NOALLOC_GRAD_MAX (1 << 20)  // complex entries: Nb * Lt up to 1,048,576
static __device__ __align__(256) zd scratch[PQMC_NOALLOC_GRAD_MAX];
#define PQMC_MAX_DEVICES 64
static std::mutex g_noalloc_mu;
static zd *g_noalloc_ptr[PQMC_MAX_DEVICES];

static int noalloc_scratch(int dev, zd **out) {
  if (dev < 0 || dev >= PQMC_MAX_DEVICES) return fail(PQMC_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(PQMC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[dev] = static_cast<zd *>(p);
  }
  *out = g_noalloc_ptr[dev];
  return PQMC_OK;
}

extern "C" int pqmc_gradient_noalloc(pqmc_handle *h, double *grad) {
  if (!h || !grad) return fail(PQMC_EINVAL, "null argument");
  if ((size_t)h->Nb * h->Lt > (size_t)PQMC_NOALLOC_GRAD_MAX)
    return fail(PQMC_EINVAL, "Nb * Lt = " + std::to_string((size_t)h->Nb * h->Lt) +
                                 " fields exceed the allocation-free gradient's fixed array of " +
                                 std::to_string(PQMC_NOALLOC_GRAD_MAX));
  int rc = gradient_check(h, grad);
  if (rc) return rc;
  zd *acc = nullptr;
  rc = noalloc_scratch(h->dev, &acc);
  if (rc) return rc;
  return gradient_pass(h, acc, grad);
}

// Compute units one chain's family passes keep resident: every workgroup of
// a pass must be resident at once, and each holds the leader's LDS.
static int pass_cus(pqmc_handle *h, int ncu, int *need) {
  int maxg = 0;
  HIPCHK(pqmc_pass_max_groups(h->Ns, h->bph, &maxg));
  const int per_cu = maxg / (ncu > 0 ? ncu : 1);
  if (per_cu < 1) return fail(PQMC_EINVAL, "a family pass workgroup does not fit one compute unit");
  *need = (h->H + 1 + per_cu - 1) / per_cu;
  return PQMC_OK;
}

static int check_chains(pqmc_handle *const *hs, int n, const float *const *u, int *ncu) {
  if (!hs || !u || n < 1) return fail(PQMC_EINVAL, "null argument or empty batch");
  for (int i = 0; i < n; ++i) {
    if (!hs[i] || !u[i]) return fail(PQMC_EINVAL, "null handle or draws in the batch");
    for (int j = 0; j < i; ++j)
      if (hs[j] == hs[i]) return fail(PQMC_EINVAL, "a chain appears twice in the batch");
    if (hs[i]->dev != hs[0]->dev) return fail(PQMC_EINVAL, "the chains of a batch must share one device");
  }
  HIPCHK(hipSetDevice(hs[0]->dev));
  HIPCHK(hipDeviceGetAttribute(ncu, hipDeviceAttributeMultiprocessorCount, hs[0]->dev));
  return PQMC_OK;
}

// Several chains, one sweep each, at the same time: one host thread per
// chain drives pqmc_sweep on the chain's own stream.  A family pass spins
// until all of its workgroups are resident, so the passes of every chain of
// the batch must fit the device together; the batch is refused otherwise.
extern "C" int pqmc_sweep_batch(pqmc_handle *const *hs, int n, const int *forward, const float *const *u,
                                int measure, double *const *meas_out) {
  if (!forward) return fail(PQMC_EINVAL, "null argument");
  int ncu = 0;
  int rc0 = check_chains(hs, n, u, &ncu);
  if (rc0) return rc0;
  long need = 0;  // compute units the passes occupy together
  for (int i = 0; i < n; ++i) {
    int c = 0;
    rc0 = pass_cus(hs[i], ncu, &c);
    if (rc0) return rc0;
    need += c;
  }
  if (need > ncu)
    return fail(PQMC_EINVAL, "the batch's family passes (" + std::to_string(need) + " compute units) do not fit the " +
                                 std::to_string(ncu) + " compute units of the device together");
  std::vector<int> rc(n, PQMC_OK);
  std::vector<std::string> msg(n);
  auto one = [&](int i) {
    if (hipSetDevice(hs[i]->dev) != hipSuccess) {
      rc[i] = PQMC_EHIP;
      msg[i] = "hipSetDevice failed";
      return;
    }
    rc[i] = pqmc_sweep(hs[i], forward[i], u[i], 0, measure, meas_out ? meas_out[i] : nullptr);
    if (rc[i]) msg[i] = g_err;
  };
  std::vector<std::thread> th;
  th.reserve(n - 1);
  for (int i = 1; i < n; ++i) th.emplace_back(one, i);
  one(0);
  for (auto &t : th) t.join();
  for (int i = 0; i < n; ++i)
    if (rc[i]) return fail(rc[i], "chain " + std::to_string(i) + " of the batch: " + msg[i]);
  return PQMC_OK;
}

// Any number of chains, one sweep each.  Up to max_concurrent host threads
// take the chains in order; a thread starts a chain's sweep once the compute
// units its family passes need fit beside those of the sweeps already
// running (a budget of the device's compute units), and returns them when
// the sweep ends.  So the passes in flight always fit together, whatever
// the mix of lattices; a chain that needs more than the device is refused
// before anything runs.
extern "C" int pqmc_sweep_many(pqmc_handle *const *hs, int n, const int *forward, const float *const *u,
                               int measure, double *const *meas_out, int max_concurrent) {
  if (!forward) return fail(PQMC_EINVAL, "null argument");
  int ncu = 0;
  int rc0 = check_chains(hs, n, u, &ncu);
  if (rc0) return rc0;
  std::vector<int> need(n);
  for (int i = 0; i < n; ++i) {
    rc0 = pass_cus(hs[i], ncu, &need[i]);
    if (rc0) return rc0;
    if (need[i] > ncu)
      return fail(PQMC_EINVAL, "chain " + std::to_string(i) + " needs " + std::to_string(need[i]) +
                                   " compute units for its family passes, more than the device has");
  }
  const int nthr = std::max(1, std::min(n, max_concurrent > 0 ? max_concurrent : 16));
  std::mutex mu;
  std::condition_variable cv;
  int free_cus = ncu;
  std::atomic<int> next{0};
  std::vector<int> rc(n, PQMC_OK);
  std::vector<std::string> msg(n);
  auto worker = [&]() {
    if (hipSetDevice(hs[0]->dev) != hipSuccess) {
      std::lock_guard<std::mutex> lk(mu);
      for (int i = next.exchange(n); i < n; ++i) {
        rc[i] = PQMC_EHIP;
        msg[i] = "hipSetDevice failed";
      }
      return;
    }
    for (int i = next++; i < n; i = next++) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return free_cus >= need[i]; });
        free_cus -= need[i];
      }
      rc[i] = pqmc_sweep(hs[i], forward[i], u[i], 0, measure, meas_out ? meas_out[i] : nullptr);
      if (rc[i]) msg[i] = g_err;
      {
        std::lock_guard<std::mutex> lk(mu);
        free_cus += need[i];
      }
      cv.notify_all();
    }
  };
  std::vector<std::thread> th;
  th.reserve(nthr - 1);
  for (int t = 1; t < nthr; ++t) th.emplace_back(worker);
  worker();
  for (auto &t : th) t.join();
  for (int i = 0; i < n; ++i)
    if (rc[i]) return fail(rc[i], "chain " + std::to_string(i) + ": " + msg[i]);
  return PQMC_OK;
}

extern "C" int pqmc_get_g(pqmc_handle *h, double *g) {
  if (!h || !g) return fail(PQMC_EINVAL, "null argument");
  int rc = PQMC_OK;
  HIPCHK(hipMemcpyAsync(g, h->d_G, sizeof(zd) * h->Ns * h->Ns, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

extern "C" int pqmc_set_g(pqmc_handle *h, const double *g) {
  if (!h || !g) return fail(PQMC_EINVAL, "null argument");
  HIPCHK(hipMemcpyAsync(h->d_G, g, sizeof(zd) * h->Ns * h->Ns, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->g_at = 0;
  return PQMC_OK;
}

extern "C" int pqmc_get_qstring(pqmc_handle *h, int block, double *q) {
  if (!h || !q) return fail(PQMC_EINVAL, "null argument");
  if (h->ft) return fail(PQMC_EINVAL, "no projector B-string at finite temperature");
  if (block < 1 || block > h->nblock) return fail(PQMC_EINVAL, "block out of range");
  HIPCHK(hipMemcpyAsync(q, QS(h, block), sizeof(zd) * h->Ns * h->ne, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

extern "C" int pqmc_get_dstring(pqmc_handle *h, double *d) {
  if (!h || !d) return fail(PQMC_EINVAL, "null argument");
  if (h->ft) return fail(PQMC_EINVAL, "no projector B-string at finite temperature");
  HIPCHK(hipMemcpyAsync(d, h->d_D, sizeof(zd) * h->ne * h->nblock, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

extern "C" int pqmc_get_lncw(pqmc_handle *h, double *ri) {
  if (!h || !ri) return fail(PQMC_EINVAL, "null argument");
  zd v;
  double acc = 0.0;
  HIPCHK(hipMemcpyAsync(&v, h->d_lncw, sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(&acc, h->d_lncw_acc, sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  ri[0] = v.x + acc;
  ri[1] = v.y;
  return PQMC_OK;
}

static int global_fields(pqmc_handle *h, const double *x) {
  const size_t nx = (size_t)h->Nb * h->Lt;
  if (!h->d_gx) {
    HIPCHK(dalloc(&h->d_gx, nx));
    HIPCHK(dalloc(&h->d_gekv, (size_t)h->Lt * 4 * h->npla));
    HIPCHK(dalloc(&h->d_glog, (size_t)h->nblock + 4));
    HIPCHK(dalloc(&h->d_gw, 1));
  }
  HIPCHK(hipMemcpyAsync(h->d_gx, x, sizeof(double) * nx, hipMemcpyHostToDevice, h->st));
  HIPCHK(pqmc_launch_ekv_build(&h->md, h->d_gx, h->d_bfi, h->d_cross, h->d_ph, h->d_gekv, 1, h->Lt, h->st));
  return PQMC_OK;
}

// cal_new_weight_FT (src/mod_evolution.f90:743-785) for the fields x: a
// right string from the identity over all slices (two rolling UDT buffers),
// then ln det(1 + B) against the identity as the left string of block
// nblock.  The reference scales the rows of Q by D where its comment says
// Q' = Q*D (:771-773); this takes Q*D, the step get_g_scratch takes.  The
// chain's strings and G are not touched, and need not exist (update_global
// may run with fields set by an accepted move, before init_g).
static int ft_global_weight(pqmc_handle *h, const double *x, double *ri) {
  if (!h->model_set) return fail(PQMC_EINVAL, "model must be set");
  const int Ns = h->Ns, ng = h->ngroup, nb = h->nblock;
  const size_t nn = (size_t)Ns * Ns;
  int rc = global_fields(h, x);
  if (rc) return rc;
  ft_udt buf[3];  // two rolling strings and the identity
  if (!h->d_gft) HIPCHK(dalloc(&h->d_gft, 6 * nn + 3 * (size_t)Ns + 3));
  for (int k = 0; k < 3; ++k)
    buf[k] = {h->d_gft + 2 * k * nn, h->d_gft + 6 * nn + k * Ns, h->d_gft + (2 * k + 1) * nn,
              h->d_gft + 6 * nn + 3 * Ns + k};
  const ft_udt id = buf[2];
  HIPCHK(pqmc_eye(id.Q, Ns, Ns, Ns, h->st));
  HIPCHK(pqmc_eye(id.T, Ns, Ns, Ns, h->st));
  HIPCHK(pqmc_ft_ones(id.D, Ns, h->st));
  HIPCHK(hipMemsetAsync(id.L, 0, sizeof(zd), h->st));
  ft_udt cur = id;
  for (int b = 2; b <= nb - 1; ++b) {
    rc = ft_step(h, cur, buf[b & 1], (b - 2) * ng + 1, (b - 1) * ng, h->d_gekv, 0);
    if (rc) return rc;
    cur = buf[b & 1];
  }
  rc = ft_weight(h, cur, id, nullptr, h->d_gw);
  if (rc) return rc;
  zd v;
  HIPCHK(hipMemcpyAsync(&v, h->d_gw, sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  ri[0] = v.x;
  ri[1] = v.y;
  return PQMC_OK;
}

// cal_new_weight_T0 (src/mod_evolution.f90:359-408) for the fields x: the
// left trial string propagated down through all slices with an LQ every
// ngroup slices (the QR of its conjugate transpose, whose real diagonal is
// the same), then log det of its product with the right trial string, plus
// the trial's log D at both ends.  The chain's own state is not touched.
extern "C" int pqmc_global_weight(pqmc_handle *h, const double *x, double *ri) {
  if (!h || !x || !ri) return fail(PQMC_EINVAL, "null argument");
  if (h->ft) return ft_global_weight(h, x, ri);
  if (!h->model_set || !h->trial_set) return fail(PQMC_EINVAL, "model and trial must be set");
  const int Ns = h->Ns, ne = h->ne, ng = h->ngroup, Lt = h->Lt;
  int rc = global_fields(h, x);
  if (rc) return rc;
  const size_t qb = sizeof(zd) * Ns * ne;
  zd *W = h->d_W, *Qn = h->d_scr;  // string (as lmat^H) and the QR's explicit Q
  HIPCHK(hipMemcpyAsync(W, h->d_slQ, qb, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(pqmc_dlog(h->d_slD, ne, h->d_glog, h->st));      // D_string(:, nblock)
  HIPCHK(pqmc_dlog(h->d_slD, ne, h->d_glog + 1, h->st));  // D_string(:, 1)
  int nlog = 2;
  for (int p = Lt; p >= 1;) {
    const int n = std::min(ng, p);
    rc = string_step(h, W, p, n, 1, h->d_gekv);
    if (rc) return rc;
    p -= n;
    if (n < ng) break;  // a partial last group has no LQ
    rc = run_qr(h, h->qa, W, Ns, Ns, ne, Qn, nullptr, nullptr, 0);
    if (rc) return rc;
    HIPCHK(pqmc_dlog_conj(h->qa.beta, ne, h->d_glog + nlog++, h->st));
    std::swap(W, Qn);
  }
  const zd one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  HIPCHK(pqmc_zgemm(1, 0, ne, ne, Ns, one, W, Ns, h->d_slQ, Ns, zero, h->d_M, ne, h->st));
  rc = run_qr(h, h->qm, h->d_M, ne, ne, ne, nullptr, nullptr, nullptr, 0);
  if (rc) return rc;
  HIPCHK(pqmc_lncw(ne, h->qm.beta, h->qm.tau, h->qm.vn2, h->d_glog, nlog, h->md.tr_weight, h->md.ncopy, h->d_gw,
                   nullptr, nullptr, nullptr, 0, h->st));
  zd v;
  HIPCHK(hipMemcpyAsync(&v, h->d_gw, sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  ri[0] = v.x;
  ri[1] = v.y;
  return PQMC_OK;
}

extern "C" int pqmc_get_accepts(pqmc_handle *h, uint8_t *acc, int64_t n) {
  if (!h || !acc) return fail(PQMC_EINVAL, "null argument");
  if (n < 0 || n > (int64_t)h->Nb * h->Lt) return fail(PQMC_EINVAL, "n out of range");
  HIPCHK(hipMemcpyAsync(acc, h->d_acc, (size_t)n, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

extern "C" int pqmc_get_counters(pqmc_handle *h, int64_t *o3, double *efm) {
  if (!h || !o3) return fail(PQMC_EINVAL, "null argument");
  unsigned long long s[4];
  HIPCHK(hipMemcpyAsync(s, h->d_stats, sizeof(s), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  o3[0] = (int64_t)s[0];
  o3[1] = (int64_t)s[1];
  o3[2] = (int64_t)s[2];
  if (efm) *efm = h->err_fast_max;
  return PQMC_OK;
}

extern "C" int pqmc_get_drift(pqmc_handle *h, double *o2) {
  if (!h || !o2) return fail(PQMC_EINVAL, "null argument");
  o2[0] = h->last_drift[0];
  o2[1] = h->last_drift[1];
  return PQMC_OK;
}

extern "C" int pqmc_get_drift_history(pqmc_handle *h, double *out, int32_t cap, int32_t *n) {
  if (!h || !n || (cap > 0 && !out)) return fail(PQMC_EINVAL, "null argument");
  *n = h->nhist;
  const int k = std::min(cap, h->nhist);
  if (k > 0) {
    HIPCHK(hipMemcpyAsync(out, h->d_hist, 2 * sizeof(double) * k, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  return PQMC_OK;
}

// Restores positive_accept, negative_accept and err_fast_max (resume from a
// checkpoint); the sync-timeout counter is left as it is.
extern "C" int pqmc_set_counters(pqmc_handle *h, const int64_t *in2, double err_fast_max) {
  if (!h || !in2) return fail(PQMC_EINVAL, "null argument");
  if (in2[0] < 0 || in2[1] < 0 || !(err_fast_max >= 0.0)) return fail(PQMC_EINVAL, "negative counter");
  unsigned long long s[2] = {(unsigned long long)in2[0], (unsigned long long)in2[1]};
  HIPCHK(hipMemcpyAsync(h->d_stats, s, sizeof(s), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->err_fast_max = err_fast_max;
  return PQMC_OK;
}

// Phase timers of the family pass (debug): enable != 0 allocates and
// zeroes them; out (8*P values, may be NULL) receives microseconds.  Rows:
// leader wave 0, the H helpers, leader waves 1, 2, 6 and 7.
extern "C" int pqmc_update_profile(pqmc_handle *h, int enable, double *out, int32_t *P) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  const int G = h->H + 5;
  if (P) *P = G;
  if (out && h->d_prof) {
    std::vector<unsigned long long> v(8 * G);
    HIPCHK(hipMemcpyAsync(v.data(), h->d_prof, sizeof(unsigned long long) * v.size(), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    int rate = 0;
    HIPCHK(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, 0));
    for (size_t i = 0; i < v.size(); ++i) out[i] = rate ? 1e3 * (double)v[i] / rate : (double)v[i];
  }
  if (enable) {
    if (!h->d_prof) HIPCHK(dalloc(&h->d_prof, 8 * G));
    HIPCHK(hipMemsetAsync(h->d_prof, 0, sizeof(unsigned long long) * 8 * G, h->st));
  } else if (h->d_prof) {
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipFree(h->d_prof));
    h->d_prof = nullptr;
  }
  return PQMC_OK;
}

extern "C" int pqmc_synchronize(pqmc_handle *h) {
  if (!h) return fail(PQMC_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  return PQMC_OK;
}

extern "C" int pqmc_timing(pqmc_handle *h, int enable, int kind, double *ms, int64_t *launches) {
  if (!h || kind < 0 || kind > 2) return fail(PQMC_EINVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(h->st));
  double tot = 0.0;
  for (auto &pr : h->spans[kind]) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, pr.first, pr.second));
    tot += t;
  }
  if (ms) *ms = tot;
  if (launches) *launches = (int64_t)h->spans[kind].size();
  h->spans[kind].clear();
  h->pools[kind].used = 0;
  h->timing[kind] = enable != 0;
  return PQMC_OK;
}

// ---------------------------------------------------------------------------
// ran2 (src/mod_ranstate.f90:171-191) and init_rng_by_seed (:267-283)

extern "C" void pqmc_ran2_seed(pqmc_ran2 *s, int32_t seed) {
  int32_t v[5] = {seed, seed, seed, seed, seed};
  for (int i = 0; i < 3; ++i)
    if (v[i] < 0) v[i] = ~v[i];
  for (int i = 3; i < 5; ++i)
    if (v[i] == 0) v[i] = 1;
  s->i = v[0]; s->j = v[1]; s->k = v[2]; s->m = v[3]; s->n = v[4];
}

extern "C" void pqmc_ran2_fill(pqmc_ran2 *s, int64_t n, float *out) {
  // amm = nearest(1.0_sp,-1.0_sp)/(-2^31)
  const float amm = -0x1.fffffep-32f;
  uint32_t I = (uint32_t)s->i, J = (uint32_t)s->j, K = (uint32_t)s->k, M = (uint32_t)s->m, N = (uint32_t)s->n;
  for (int64_t q = 0; q < n; ++q) {
    int32_t r = (int32_t)(I - K);
    if (r < 0) r = (int32_t)((uint32_t)r + 2147483579u);
    I = J;
    J = K;
    K = (uint32_t)r;
    N ^= N << 13;
    N ^= N >> 17;
    N ^= N << 5;
    const uint32_t lo = M & 65535u;
    M = ((3533u * (M >> 16) + lo) << 16) + 3533u * lo + 820265819u;
    const int32_t v = (int32_t)((N ^ K) + M);
    out[q] = amm * (float)((v < 0) ? v : ~v);
  }
  s->i = (int32_t)I; s->j = (int32_t)J; s->k = (int32_t)K; s->m = (int32_t)M; s->n = (int32_t)N;
}

// get_expKV_epsoc (src/mod_phonon_field.f90:572-604) for n (bond, slice)
// pairs, on the current device; host arrays in and out.
extern "C" int pqmc_expkv_epsoc(double dtau, int32_t nfam, const double *kcoe, const double *vcoe, const double *k12,
                                const double *v12, int64_t n, const int32_t *fam, const double *x,
                                const int32_t *cross, const double *bcph, int32_t inv, double *out) {
  if (n < 0 || nfam <= 0 || !(dtau > 0.0)) return fail(PQMC_EINVAL, "expkv_epsoc: n >= 0, nfam > 0 and dtau > 0");
  if (n == 0) return PQMC_OK;
  if (!kcoe || !vcoe || !k12 || !v12 || !fam || !x || !out || (!bcph && cross))
    return fail(PQMC_EINVAL, "expkv_epsoc: null argument");
  for (int64_t i = 0; i < n; ++i)
    if (fam[i] < 0 || fam[i] >= nfam) return fail(PQMC_EINVAL, "expkv_epsoc: family index out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PQMC_ENODEV, "no GPU");
  zd *d_c = nullptr, *d_ph = nullptr, *d_out = nullptr;
  int *d_fam = nullptr, *d_cross = nullptr;
  double *d_x = nullptr;
  hipStream_t st = nullptr;
  int rc = PQMC_OK;
  auto done = [&](int r) {
    if (st) hipStreamDestroy(st);
    hipFree(d_c);
    hipFree(d_ph);
    hipFree(d_out);
    hipFree(d_fam);
    hipFree(d_cross);
    hipFree(d_x);
    return r;
  };
#define EPCHK(expr)                                                                              \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return done(fail(PQMC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
  } while (0)
  EPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  EPCHK(dalloc(&d_c, 4 * (size_t)nfam));
  EPCHK(dalloc(&d_fam, (size_t)n));
  EPCHK(dalloc(&d_x, (size_t)n));
  EPCHK(dalloc(&d_out, 4 * (size_t)n));
  if (cross) {
    EPCHK(dalloc(&d_cross, (size_t)n));
    EPCHK(dalloc(&d_ph, (size_t)n));
  }
  const size_t zb = sizeof(zd) * nfam;
  EPCHK(hipMemcpyAsync(d_c, kcoe, zb, hipMemcpyHostToDevice, st));
  EPCHK(hipMemcpyAsync(d_c + nfam, vcoe, zb, hipMemcpyHostToDevice, st));
  EPCHK(hipMemcpyAsync(d_c + 2 * nfam, k12, zb, hipMemcpyHostToDevice, st));
  EPCHK(hipMemcpyAsync(d_c + 3 * nfam, v12, zb, hipMemcpyHostToDevice, st));
  EPCHK(hipMemcpyAsync(d_fam, fam, sizeof(int) * n, hipMemcpyHostToDevice, st));
  EPCHK(hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (cross) {
    EPCHK(hipMemcpyAsync(d_cross, cross, sizeof(int) * n, hipMemcpyHostToDevice, st));
    EPCHK(hipMemcpyAsync(d_ph, bcph, sizeof(zd) * n, hipMemcpyHostToDevice, st));
  }
  EPCHK(pqmc_launch_ekv_epsoc(dtau, d_c, d_c + nfam, d_c + 2 * nfam, d_c + 3 * nfam, nfam, (long)n, d_fam, d_x,
                              d_cross, d_ph, inv, d_out, st));
  EPCHK(hipMemcpyAsync(out, d_out, sizeof(zd) * 4 * n, hipMemcpyDeviceToHost, st));
  EPCHK(hipStreamSynchronize(st));
#undef EPCHK
  return done(rc);
}

// which = 0: C = alpha op(A) op(B) + beta C (pqmc_zgemm; A is lda x (opa ?
// M : K), B ldb x (opb ? K : N), C ldc x N); which = 1: C <- R^-1 C with R =
// A upper triangular (M x M, ld lda) and C ldc x N (pqmc_trsm_upper).  Host
// arrays of interleaved complex doubles, column-major; the library's own
// device buffers and the null stream.
extern "C" int pqmc_dense_check(int which, int opa, int opb, int M, int N, int K, const double *alpha_beta,
                                const double *A, int lda, const double *B, int ldb, double *C, int ldc) {
  if (which < 0 || which > 1 || M <= 0 || N <= 0 || (which == 0 && K <= 0) || !A || !C || lda < 1 || ldc < M ||
      (which == 0 && (!B || !alpha_beta || ldb < 1)))
    return fail(PQMC_EINVAL, "pqmc_dense_check: bad arguments");
  const size_t na = (size_t)lda * (which == 0 ? (opa ? M : K) : M), nb = which == 0 ? (size_t)ldb * (opb ? K : N) : 0,
               nc = (size_t)ldc * N, nw = (size_t)256 * ((M + 15) / 16);
  zd *dA = nullptr, *dB = nullptr, *dC = nullptr, *dW = nullptr;
  int rc = PQMC_OK;
  auto run = [&]() -> int {
    HIPCHK(hipMalloc(&dA, sizeof(zd) * na));
    HIPCHK(hipMalloc(&dC, sizeof(zd) * nc));
    HIPCHK(hipMemcpy(dA, A, sizeof(zd) * na, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dC, C, sizeof(zd) * nc, hipMemcpyHostToDevice));
    if (which == 0) {
      HIPCHK(hipMalloc(&dB, sizeof(zd) * nb));
      HIPCHK(hipMemcpy(dB, B, sizeof(zd) * nb, hipMemcpyHostToDevice));
      const zd alpha = make_double2(alpha_beta[0], alpha_beta[1]), beta = make_double2(alpha_beta[2], alpha_beta[3]);
      HIPCHK(pqmc_zgemm(opa, opb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, nullptr));
    } else {
      HIPCHK(hipMalloc(&dW, sizeof(zd) * nw));
      HIPCHK(pqmc_trsm_upper(dA, lda, M, dC, ldc, N, dW, nullptr));
    }
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(C, dC, sizeof(zd) * nc, hipMemcpyDeviceToHost));
    return PQMC_OK;
  };
  rc = run();
  hipFree(dA);
  hipFree(dB);
  hipFree(dC);
  hipFree(dW);
  return rc;
}

extern "C" int pqmc_mfma_selftest(double *max_abs_err) {
  double A[64], B[64], C[256], ref[256];
  for (int i = 0; i < 64; ++i) {
    A[i] = std::sin(0.37 * i + 0.1);
    B[i] = std::cos(0.23 * i - 0.4);
  }
  for (int j = 0; j < 16; ++j)
    for (int i = 0; i < 16; ++i) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A[i + 16 * k] * B[k + 4 * j];
      ref[i + 16 * j] = s;
    }
  double *dA, *dB, *dC;
  HIPCHK(hipMalloc(&dA, sizeof(A)));
  HIPCHK(hipMalloc(&dB, sizeof(B)));
  HIPCHK(hipMalloc(&dC, sizeof(C)));
  HIPCHK(hipMemcpy(dA, A, sizeof(A), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, B, sizeof(B), hipMemcpyHostToDevice));
  HIPCHK(pqmc_mfma_probe(dA, dB, dC, nullptr));
  HIPCHK(hipMemcpy(C, dC, sizeof(C), hipMemcpyDeviceToHost));
  hipFree(dA);
  hipFree(dB);
  hipFree(dC);
  double m = 0.0;
  for (int i = 0; i < 256; ++i) m = std::fmax(m, std::fabs(C[i] - ref[i]));
  if (max_abs_err) *max_abs_err = m;
  return PQMC_OK;
}
