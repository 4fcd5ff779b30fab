/*
 * pqmc_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A from-scratch CPU restatement, in C, of the projector-DQMC sweep of the
 * reference LEOMMM816/PQMC_allinone for the EPSOCZ model.  It is the checker
 * for the HIP library and the "port" CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every routine cites the reference file:line it restates.  Dense linear
 * algebra the reference delegates to LAPACK (ZGEQRF/ZUNGQR, ZGELQF/ZUNGLQ,
 * ZGETRF/ZGETRI) is restated here from the published LAPACK algorithms
 * (unblocked zgeqr2/zung2r/zgetf2 semantics, including the real-beta
 * Householder convention of zlarfg), because no LAPACK is linked.
 *
 * Build flags (oracle/Makefile): -O3 -fcx-fortran-rules -ffp-contract=off so
 * complex products use the plain formula (as gfortran does) and results do
 * not depend on the host's FMA support.
 */
#include "pqmc_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex zc;

#define SP_PI 3.14159274101257324219 /* REAL(SP) PI of src/mod_nrtype.f90:42 */

struct orc_model {
  orc_params p;
  int L1, L2, Ncell, Ns, ne, Nb, nfam, npla, Lt, ngroup, nblock;
  double dtau, M, D, biased, jump_distance, char_length, hop, lam;
  int ncopy, tr_weight;
  zc bc[2];
  int *sites; /* [nfam][npla][4] */
  int *bfi;   /* [nfam][npla] */
  int *cross; /* [nfam][npla] */
  zc *bcph;   /* [nfam][npla][16] col-major 4x4 */
  zc kcoe[4], vcoe[4];
  zc k12[4][4], v12[4][4]; /* Kmatrix(1:2,3:4) col-major: [a + 2b] */
  double *x;               /* boson_field(Nb, Lt) */
  zc *ekv, *ekvi;          /* expKV(4,4,npla,nfam,Lt) */
  zc *g;                   /* g_h(Ns,Ns) */
  zc *Q, *Dst, *R;         /* Q_string(Ns,ne,nblock), D_string(ne,nblock), R_string */
  zc *slQ, *slD, *slR;
  /* finite temperature (proj = .false., src/mod_evolution.f90:411-785):
   * block b holds a right string B(t,0) = Q diag(D) T (T unit upper
   * products) or a left string B(beta,t) = T diag(D) Q (Q rows orthonormal) */
  int ft;
  zc *FQ, *FD, *FT; /* Ns x Ns x nblock, Ns x nblock, Ns x Ns x nblock */
  zc ln_cw;
  double err_fast, err_fast_max;
  long pos_acc, neg_acc;
  long global_acc, global_rej;
  /* measurement registry (input/model/EPSOCZ.nml:262-283) */
  int total_width;
};

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

/* modi of src/mod_matrixlib.f90:66-77 */
static int modi(int n, int p) {
  int r = n % p;
  if (n < 0) return p + r;
  if (r == 0) return p;
  return r;
}

static inline zc cmul(zc a, zc b) { return a * b; }

/* det4x4 of src/mod_matrixlib.f90:563-590: cofactor expansion along row 1
 * (a is column-major 4x4, a(i,j) = a[(i-1) + 4(j-1)]) */
static zc det4(const zc *a) {
#define A(i, j) a[((i)-1) + 4 * ((j)-1)]
  zc res;
  res = A(1, 1) * (A(2, 2) * (A(3, 3) * A(4, 4) - A(3, 4) * A(4, 3)) -
                   A(2, 3) * (A(3, 2) * A(4, 4) - A(3, 4) * A(4, 2)) +
                   A(2, 4) * (A(3, 2) * A(4, 3) - A(3, 3) * A(4, 2)));
  res = res - A(1, 2) * (A(2, 1) * (A(3, 3) * A(4, 4) - A(3, 4) * A(4, 3)) -
                         A(2, 3) * (A(3, 1) * A(4, 4) - A(3, 4) * A(4, 1)) +
                         A(2, 4) * (A(3, 1) * A(4, 3) - A(3, 3) * A(4, 1)));
  res = res + A(1, 3) * (A(2, 1) * (A(3, 2) * A(4, 4) - A(3, 4) * A(4, 2)) -
                         A(2, 2) * (A(3, 1) * A(4, 4) - A(3, 4) * A(4, 1)) +
                         A(2, 4) * (A(3, 1) * A(4, 2) - A(3, 2) * A(4, 1)));
  res = res - A(1, 4) * (A(2, 1) * (A(3, 2) * A(4, 3) - A(3, 3) * A(4, 2)) -
                         A(2, 2) * (A(3, 1) * A(4, 3) - A(3, 3) * A(4, 1)) +
                         A(2, 3) * (A(3, 1) * A(4, 2) - A(3, 2) * A(4, 1)));
#undef A
  return res;
}

/* 4x4 product c = a*b (Fortran matmul order: sum over k ascending) */
static void mm4(zc *c, const zc *a, const zc *b) {
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      zc s = 0;
      for (int k = 0; k < 4; ++k) s += a[i + 4 * k] * b[k + 4 * j];
      c[i + 4 * j] = s;
    }
}

/* ------------------------------------------------------------------ */
/* LAPACK-equivalent dense kernels                                     */
/* ------------------------------------------------------------------ */

/* zgetf2 (partial pivoting on |Re|+|Im| as izamax); returns info */
static int lu_factor(int n, zc *a, int lda, int *ipiv) {
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = -1.0;
    for (int i = j; i < n; ++i) {
      double v = fabs(creal(a[i + (size_t)lda * j])) + fabs(cimag(a[i + (size_t)lda * j]));
      if (v > best) { best = v; p = i; }
    }
    ipiv[j] = p;
    if (a[p + (size_t)lda * j] != 0) {
      if (p != j)
        for (int c = 0; c < n; ++c) {
          zc t = a[j + (size_t)lda * c];
          a[j + (size_t)lda * c] = a[p + (size_t)lda * c];
          a[p + (size_t)lda * c] = t;
        }
      zc r = 1.0 / a[j + (size_t)lda * j];
      for (int i = j + 1; i < n; ++i) a[i + (size_t)lda * j] *= r;
    } else if (!info) {
      info = j + 1;
    }
    for (int c = j + 1; c < n; ++c) {
      zc ajc = a[j + (size_t)lda * c];
      if (ajc == 0) continue;
      for (int i = j + 1; i < n; ++i) a[i + (size_t)lda * c] -= a[i + (size_t)lda * j] * ajc;
    }
  }
  return info;
}

/* zinverse of src/mod_matrixlib.f90:410-431 (ZGETRF + ZGETRI): inverse via
 * the LU factors, one right-hand side e_j at a time. */
static int zinverse(int n, zc *a) {
  if (n == 1) { a[0] = 1.0 / a[0]; return 0; }
  int *ipiv = xcalloc(n, sizeof(int));
  zc *lu = xcalloc((size_t)n * n, sizeof(zc));
  memcpy(lu, a, sizeof(zc) * n * n);
  int info = lu_factor(n, lu, n, ipiv);
  zc *col = xcalloc(n, sizeof(zc));
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) col[i] = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i) { /* P b */
      int p = ipiv[i];
      if (p != i) { zc t = col[i]; col[i] = col[p]; col[p] = t; }
    }
    for (int i = 0; i < n; ++i) { /* L y = b (unit lower) */
      zc s = col[i];
      for (int k = 0; k < i; ++k) s -= lu[i + (size_t)n * k] * col[k];
      col[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) { /* U x = y */
      zc s = col[i];
      for (int k = i + 1; k < n; ++k) s -= lu[i + (size_t)n * k] * col[k];
      col[i] = s / lu[i + (size_t)n * i];
    }
    for (int i = 0; i < n; ++i) a[i + (size_t)n * j] = col[i];
  }
  free(col);
  free(lu);
  free(ipiv);
  return info;
}

/* zdet of src/mod_matrixlib.f90:515-537 (closed forms n<=4, zdet_lapack
 * :624-653 otherwise: product of U's diagonal, sign flip per interchange) */
static zc zdet(int n, const zc *a) {
  if (n == 1) return a[0];
  if (n == 2) return a[0] * a[3] - a[2] * a[1];
  if (n == 3) {
#define A(i, j) a[((i)-1) + 3 * ((j)-1)]
    return A(1, 1) * (A(2, 2) * A(3, 3) - A(2, 3) * A(3, 2)) -
           A(1, 2) * (A(2, 1) * A(3, 3) - A(2, 3) * A(3, 1)) +
           A(1, 3) * (A(2, 1) * A(3, 2) - A(2, 2) * A(3, 1));
#undef A
  }
  if (n == 4) return det4(a);
  zc *b = xcalloc((size_t)n * n, sizeof(zc));
  int *ipiv = xcalloc(n, sizeof(int));
  memcpy(b, a, sizeof(zc) * n * n);
  int info = lu_factor(n, b, n, ipiv);
  zc d = 0;
  if (!info) {
    int sgn = 1;
    d = 1.0;
    for (int i = 0; i < n; ++i) {
      if (ipiv[i] != i) sgn = -sgn;
      d = d * b[i + (size_t)n * i];
    }
    if (sgn < 0) d = -d;
  }
  free(ipiv);
  free(b);
  return d;
}

/* zgeqr2: Householder QR with the zlarfg convention (beta real,
 * beta = -sign(Re alpha) * |(alpha, x)|), H(j)^H applied from the left. */
static void zgeqr2(int m, int n, zc *a, int lda, zc *tau) {
  for (int j = 0; j < n; ++j) {
    zc *col = a + j + (size_t)lda * j;
    int len = m - j;
    zc alpha = col[0];
    double ss = 0.0;
    for (int i = 1; i < len; ++i) ss += creal(col[i]) * creal(col[i]) + cimag(col[i]) * cimag(col[i]);
    double xnorm = sqrt(ss);
    double ar = creal(alpha), ai = cimag(alpha);
    zc t = 0;
    if (!(xnorm == 0.0 && ai == 0.0)) {
      double beta = -copysign(sqrt(ar * ar + ai * ai + xnorm * xnorm), ar);
      t = CMPLX((beta - ar) / beta, -ai / beta);
      zc sc = 1.0 / (alpha - beta);
      for (int i = 1; i < len; ++i) col[i] *= sc;
      alpha = beta;
    }
    tau[j] = t;
    col[0] = 1.0;
    zc ct = conj(t);
    for (int c = j + 1; c < n; ++c) {
      zc *ac = a + j + (size_t)lda * c;
      zc w = 0;
      for (int i = 0; i < len; ++i) w += conj(col[i]) * ac[i];
      w *= ct;
      for (int i = 0; i < len; ++i) ac[i] -= col[i] * w;
    }
    col[0] = alpha;
  }
}

/* zung2r: explicit Q = H(1)...H(n) [I;0] */
static void zung2r(int m, int n, zc *a, int lda, const zc *tau) {
  for (int i = n - 1; i >= 0; --i) {
    if (i < n - 1) {
      a[i + (size_t)lda * i] = 1.0;
      for (int c = i + 1; c < n; ++c) {
        zc w = 0;
        for (int r = i; r < m; ++r) w += conj(a[r + (size_t)lda * i]) * a[r + (size_t)lda * c];
        w *= tau[i];
        for (int r = i; r < m; ++r) a[r + (size_t)lda * c] -= a[r + (size_t)lda * i] * w;
      }
    }
    for (int r = i + 1; r < m; ++r) a[r + (size_t)lda * i] *= -tau[i];
    a[i + (size_t)lda * i] = 1.0 - tau[i];
    for (int r = 0; r < i; ++r) a[r + (size_t)lda * i] = 0.0;
  }
}

/* zqdr of src/mod_matrixlib.f90:323-334 over zqr :452-467:
 * a(m,n) -> Q, r(n,n) unit upper (rows scaled by 1/d), d = diag(R) */
static void zqdr(int m, int n, zc *a, zc *r, zc *d) {
  zc *tau = xcalloc(n, sizeof(zc));
  zgeqr2(m, n, a, m, tau);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) r[i + (size_t)n * j] = (i <= j) ? a[i + (size_t)m * j] : 0.0;
  zung2r(m, n, a, m, tau);
  for (int i = 0; i < n; ++i) {
    d[i] = r[i + (size_t)n * i];
    for (int j = i; j < n; ++j) r[i + (size_t)n * j] /= d[i];
  }
  free(tau);
}

/* zldq of src/mod_matrixlib.f90:349-360 over zlq :488-503.  ZGELQF on a(m,n)
 * (m <= n) equals ZGEQRF on a^H (zgelq2 conjugates each row before zlarfg
 * and applies the reflector from the right), so: a^H = Q R, a <- Q^H,
 * l = R^H with columns scaled by 1/d, d = diag(l) (real). */
static void zldq(int m, int n, zc *a, zc *l, zc *d) {
  zc *ah = xcalloc((size_t)m * n, sizeof(zc));
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) ah[j + (size_t)n * i] = conj(a[i + (size_t)m * j]);
  zc *tau = xcalloc(m, sizeof(zc));
  zgeqr2(n, m, ah, n, tau);
  for (int j = 0; j < m; ++j) /* l(i,j) = conj(R(j,i)) for i >= j */
    for (int i = 0; i < m; ++i) l[i + (size_t)m * j] = (i >= j) ? conj(ah[j + (size_t)n * i]) : 0.0;
  zung2r(n, m, ah, n, tau);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) a[i + (size_t)m * j] = conj(ah[j + (size_t)n * i]);
  for (int i = 0; i < m; ++i) {
    d[i] = l[i + (size_t)m * i];
    for (int r = i; r < m; ++r) l[r + (size_t)m * i] /= d[i];
  }
  free(tau);
  free(ah);
}

/* C(m,n) = A(m,k) * B(k,n) (ZGEMM 'N','N', beta = 0) */
static void gemm_nn(int m, int n, int k, const zc *A, const zc *B, zc *C) {
  for (int j = 0; j < n; ++j) {
    zc *cj = C + (size_t)m * j;
    for (int i = 0; i < m; ++i) cj[i] = 0;
    for (int p = 0; p < k; ++p) {
      zc b = B[p + (size_t)k * j];
      const zc *ap = A + (size_t)m * p;
      for (int i = 0; i < m; ++i) cj[i] += ap[i] * b;
    }
  }
}

/* ------------------------------------------------------------------ */
/* model setup                                                         */
/* ------------------------------------------------------------------ */

/* EPSOCZ bond families, input/model/EPSOCZ.nml:120-238:
 * pla_tsl_dvec_uc = (2,0),(1,1) for all; offsets; pla_int_subsites;
 * Kmatrix(1:2,3:4) = I; Vmatrix(1:2,3:4) = [[0,-1],[1,0]] (x) or
 * [[0,i],[i,0]] (y); bf_index 1 (x) / 2 (y). */
static const int FAM_OFF[4][2] = {{0, 0}, {1, 0}, {0, 0}, {0, 1}};
static const int FAM_INT[4][4][3] = {
    {{0, 0, 1}, {0, 0, 2}, {1, 0, 1}, {1, 0, 2}},
    {{0, 0, 1}, {0, 0, 2}, {1, 0, 1}, {1, 0, 2}},
    {{0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2}},
    {{0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2}},
};
static const int FAM_BF[4] = {1, 1, 2, 2};

static void cell_dpos(const orc_model *m, int c, int *d) {
  d[0] = c % m->L1;
  d[1] = c / m->L1;
}

/* get_uc_index_from_dpos of src/mod_lattice.f90:256-280 (0-based result) */
static int uc_index(const orc_model *m, const int *dp) {
  int L[2] = {m->L1, m->L2};
  int idx = 0, mul = 1;
  for (int i = 0; i < 2; ++i) {
    int v = dp[i];
    if (v < 0) v += L[i];
    v %= L[i];
    idx += v * mul;
    mul *= L[i];
  }
  return idx;
}

/* set_typed_pf of src/mod_phonon_field.f90:249-421 for the four EPSOCZ
 * families (dim 4, two subsites per cell, site = 2*cell + subsite-1). */
static void build_families(orc_model *m) {
  int L[2] = {m->L1, m->L2};
  /* inverse of M = [[2,1],[0,1]] (columns (2,0),(1,1)) */
  const double inv[2][2] = {{0.5, -0.5}, {0.0, 1.0}};
  for (int f = 0; f < 4; ++f) {
    int cnt = 0;
    for (int c = 0; c < m->Ncell; ++c) {
      int dp[2];
      cell_dpos(m, c, dp);
      double cd[2];
      for (int j = 0; j < 2; ++j) {
        cd[j] = dp[j] - FAM_OFF[f][j];
        if (cd[j] < -0.0001) cd[j] += L[j];
        if (cd[j] >= L[j] - 0.0001) cd[j] -= L[j];
      }
      double xp0 = inv[0][0] * cd[0] + inv[0][1] * cd[1];
      double xp1 = inv[1][0] * cd[0] + inv[1][1] * cd[1];
      if (fabs(nearbyint(xp0) - xp0) > 1e-6 || fabs(nearbyint(xp1) - xp1) > 1e-6) continue;
      if (cnt >= m->npla) abort();
      m->sites[(f * m->npla + cnt) * 4 + 0] = 2 * c + (FAM_INT[f][0][2] - 1);
      ++cnt;
    }
    if (cnt != m->npla) abort();
    for (int b = 0; b < m->npla; ++b) {
      int *s = &m->sites[(f * m->npla + b) * 4];
      int pos[2];
      cell_dpos(m, s[0] / 2, pos);
      for (int i = 1; i < 4; ++i) {
        int cdp[2];
        for (int j = 0; j < 2; ++j) {
          cdp[j] = pos[j] + FAM_INT[f][i][j];
          if (cdp[j] > L[j] - 0.0001) cdp[j] -= L[j];
          if (cdp[j] < -0.0001) cdp[j] += L[j];
        }
        s[i] = 2 * uc_index(m, cdp) + (FAM_INT[f][i][2] - 1);
      }
      int cr = 0;
      zc *ph = &m->bcph[(f * m->npla + b) * 16];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          zc v = 1.0;
          int idp[2];
          cell_dpos(m, s[i] / 2, idp);
          for (int k = 0; k < 2; ++k) {
            int jd = idp[k] - FAM_INT[f][i][k] + FAM_INT[f][j][k];
            if (jd >= L[k] - 0.0001) { cr = 1; v = v * m->bc[k]; }
            if (jd < -0.0001) { cr = 1; v = v * conj(m->bc[k]); }
          }
          ph[i + 4 * j] = v;
        }
      m->cross[f * m->npla + b] = cr;
      m->bfi[f * m->npla + b] = (FAM_BF[f] - 1) + 2 * (s[0] / 2);
    }
    /* K_coe = 1*(-hop), V_coe = 1*ep_parameter (src/input.f90:207-216) */
    m->kcoe[f] = -m->hop;
    m->vcoe[f] = m->lam;
    m->k12[f][0] = 1.0; m->k12[f][1] = 0.0; m->k12[f][2] = 0.0; m->k12[f][3] = 1.0;
    if (f < 2) { /* V(1,3)=0, V(2,3)=1, V(1,4)=-1, V(2,4)=0 */
      m->v12[f][0] = 0.0; m->v12[f][1] = 1.0; m->v12[f][2] = -1.0; m->v12[f][3] = 0.0;
    } else {     /* V(1,3)=0, V(2,3)=i, V(1,4)=i, V(2,4)=0 */
      m->v12[f][0] = 0.0; m->v12[f][1] = I; m->v12[f][2] = I; m->v12[f][3] = 0.0;
    }
  }
}

orc_model *orc_create(const orc_params *p) {
  orc_model *m = xcalloc(1, sizeof(*m));
  m->p = *p;
  m->L1 = p->L1;
  m->L2 = p->L2;
  m->Ncell = p->L1 * p->L2;
  m->Ns = 2 * m->Ncell;
  m->ne = p->ne;
  m->Nb = 2 * m->Ncell;
  m->nfam = 4;
  m->npla = m->Ns / 4;
  m->Lt = p->Lt;
  m->ngroup = p->ngroup;
  m->nblock = m->Lt / m->ngroup + 2;
  m->dtau = p->dtau;
  m->hop = p->hop;
  m->lam = p->lam;
  m->biased = p->biased;
  m->ncopy = p->ncopy;
  m->tr_weight = p->tr_weight;
  /* init of src/input.f90:167-205 */
  m->D = 1.0;
  m->M = m->D * pow(p->omega, -2);
  if (m->M < 0.001) {
    m->M = 0.0;
    m->char_length = sqrt(1.0 / (m->D * p->lam));
  } else {
    m->char_length = 1.0 / (m->M * p->omega);
  }
  m->jump_distance = p->jump * (0.1 / p->dtau) * m->char_length;
  /* BC_phase = exp(i*phase*pi) (src/mod_lattice.f90:81-87) */
  m->bc[0] = cexp(I * p->bc_phase1 * acos(-1.0));
  m->bc[1] = cexp(I * p->bc_phase2 * acos(-1.0));
  m->sites = xcalloc((size_t)4 * m->npla * 4, sizeof(int));
  m->bfi = xcalloc((size_t)4 * m->npla, sizeof(int));
  m->cross = xcalloc((size_t)4 * m->npla, sizeof(int));
  m->bcph = xcalloc((size_t)4 * m->npla * 16, sizeof(zc));
  build_families(m);
  m->x = xcalloc((size_t)m->Nb * m->Lt, sizeof(double));
  m->ekv = xcalloc((size_t)16 * m->npla * 4 * m->Lt, sizeof(zc));
  m->ekvi = xcalloc((size_t)16 * m->npla * 4 * m->Lt, sizeof(zc));
  m->g = xcalloc((size_t)m->Ns * m->Ns, sizeof(zc));
  m->Q = xcalloc((size_t)m->Ns * m->ne * m->nblock, sizeof(zc));
  m->Dst = xcalloc((size_t)m->ne * m->nblock, sizeof(zc));
  m->R = xcalloc((size_t)m->ne * m->ne * m->nblock, sizeof(zc));
  m->slQ = xcalloc((size_t)m->Ns * m->ne, sizeof(zc));
  m->slD = xcalloc((size_t)m->ne, sizeof(zc));
  m->slR = xcalloc((size_t)m->ne * m->ne, sizeof(zc));
  m->err_fast_max = 0.000001; /* src/input.f90:53 */
  /* 3 scalars + 19 fields (each with a _k companion), EPSOCZ.nml:262-283 */
  m->total_width = 3 + 19 * 2 * m->Ncell;
  return m;
}

void orc_destroy(orc_model *m) {
  if (!m) return;
  free(m->sites); free(m->bfi); free(m->cross); free(m->bcph);
  free(m->x); free(m->ekv); free(m->ekvi); free(m->g);
  free(m->Q); free(m->Dst); free(m->R); free(m->slQ); free(m->slD); free(m->slR);
  free(m->FQ); free(m->FD); free(m->FT);
  free(m);
}

void orc_get_dims(const orc_model *m, orc_dims *d) {
  d->Ns = m->Ns; d->Ncell = m->Ncell; d->Nb = m->Nb; d->nfam = m->nfam;
  d->npla = m->npla; d->nblock = m->nblock; d->total_width = m->total_width;
  d->M = m->M; d->D = m->D; d->char_length = m->char_length; d->jump_distance = m->jump_distance;
}

void orc_get_model(const orc_model *m, int *sites, int *bfi, int *cross, double *bcph) {
  if (sites) memcpy(sites, m->sites, sizeof(int) * 4 * m->npla * 4);
  if (bfi) memcpy(bfi, m->bfi, sizeof(int) * 4 * m->npla);
  if (cross) memcpy(cross, m->cross, sizeof(int) * 4 * m->npla);
  if (bcph) memcpy(bcph, m->bcph, sizeof(zc) * 4 * m->npla * 16);
}

void orc_get_coeffs(const orc_model *m, double *kcoe, double *vcoe, double *k12, double *v12) {
  memcpy(kcoe, m->kcoe, sizeof(m->kcoe));
  memcpy(vcoe, m->vcoe, sizeof(m->vcoe));
  memcpy(k12, m->k12, sizeof(m->k12));
  memcpy(v12, m->v12, sizeof(m->v12));
}

void orc_set_fields(orc_model *m, const double *x) { memcpy(m->x, x, sizeof(double) * m->Nb * m->Lt); }
void orc_get_fields(const orc_model *m, double *x) { memcpy(x, m->x, sizeof(double) * m->Nb * m->Lt); }

/* ------------------------------------------------------------------ */
/* bond exponentials                                                   */
/* ------------------------------------------------------------------ */

/* get_expKV_epsocz of src/mod_phonon_field.f90:534-569:
 * E = cosh(n) I4 + [[0, +-h sinh(n)/n],[(.)^H, 0]],
 * h = -delt (K_coe K(1:2,3:4) + V_coe x V(1:2,3:4)) (.* BC phase),
 * n = delt sqrt(|K_coe|^2 + (|V_coe| x)^2); inv flips the sign of h. */
static void expkv(const orc_model *m, int f, int b, double xv, int inv, zc *e) {
  double ak = cabs(m->kcoe[f]), av = cabs(m->vcoe[f]);
  double norm = m->dtau * sqrt(ak * ak + (av * xv) * (av * xv));
  zc sub[4];
  zc vx = m->vcoe[f] * xv;
  for (int q = 0; q < 4; ++q) sub[q] = -m->dtau * (m->kcoe[f] * m->k12[f][q] + vx * m->v12[f][q]);
  if (m->cross[f * m->npla + b]) {
    const zc *ph = &m->bcph[(f * m->npla + b) * 16];
    /* BC_phases(1:2,3:4): sub(a,b) *= ph(a, 2+b) */
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) sub[a + 2 * c] = sub[a + 2 * c] * ph[a + 4 * (2 + c)];
  }
  for (int q = 0; q < 16; ++q) e[q] = 0;
  double ch = cosh(norm), sh = sinh(norm);
  for (int i = 0; i < 4; ++i) e[i + 4 * i] = ch;
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) {
      zc v = inv ? (-sub[a + 2 * c]) * sh / norm : sub[a + 2 * c] * sh / norm;
      e[a + 4 * (2 + c)] = v;
      e[(2 + c) + 4 * a] = conj(v);
    }
}

void orc_get_expkv(const orc_model *m, int pf_id, int pla, double xv, int inv, double *out) {
  expkv(m, pf_id - 1, pla, xv, inv, (zc *)out);
}

/* get_expKV_epsoc of src/mod_phonon_field.f90:572-604, the EPSOC model's
 * spin-up 2x2 bond: E = cosh(n) I2 + [[0, +-s sinh(n)/n],[conj(.), 0]],
 * s = -delt (K_coe K(1,2) + V_coe x V(1,2)) (* BC_phases(1,2) when the bond
 * crosses the boundary), n = delt sqrt(|K_coe|^2 + (|V_coe| x)^2); inv flips
 * the sign of s.  out: 2x2 column-major.  ph: null when the bond does not
 * cross. */
void orc_expkv_epsoc(double dtau, const double *kcoe, const double *vcoe, const double *k12, const double *v12,
                     const double *ph, double xv, int inv, double *out) {
  const zc kc = CMPLX(kcoe[0], kcoe[1]), vc = CMPLX(vcoe[0], vcoe[1]);
  const zc kk = CMPLX(k12[0], k12[1]), vv = CMPLX(v12[0], v12[1]);
  double ak = cabs(kc), av = cabs(vc);
  double norm = dtau * sqrt(ak * ak + (av * xv) * (av * xv));
  zc sub = -dtau * (kc * kk + (vc * xv) * vv);
  if (ph) sub = sub * CMPLX(ph[0], ph[1]);
  const double ch = cosh(norm), sh = sinh(norm);
  const zc e12 = inv ? (-sub) * sh / norm : sub * sh / norm;
  zc *e = (zc *)out;
  e[0] = ch;
  e[1] = conj(e12);
  e[2] = e12;
  e[3] = ch;
}

static zc *EKV(const orc_model *m, zc *tab, int f, int b, int t /*1-based*/) {
  return tab + 16 * ((size_t)b + (size_t)m->npla * (f + 4 * (size_t)(t - 1)));
}

/* init_expKV (src/mod_phonon_field.f90:423-449) + set_pf_expKV_mat (:605-628) */
void orc_init_expkv(orc_model *m) {
  for (int t = 1; t <= m->Lt; ++t)
    for (int f = 0; f < 4; ++f)
      for (int b = 0; b < m->npla; ++b) {
        double xv = m->x[m->bfi[f * m->npla + b] + (size_t)m->Nb * (t - 1)];
        expkv(m, f, b, xv, 0, EKV(m, m->ekv, f, b, t));
        expkv(m, f, b, xv, 1, EKV(m, m->ekvi, f, b, t));
      }
}

/* ------------------------------------------------------------------ */
/* propagation (src/mod_evolution.f90:236-328)                         */
/* ------------------------------------------------------------------ */

/* rows S of mat(ld, n): mat(S,:) = E * mat(S,:) */
static void rows_mix(zc *mat, int ld, int n, const int *S, const zc *E) {
  for (int j = 0; j < n; ++j) {
    zc *cj = mat + (size_t)ld * j;
    zc b[4] = {cj[S[0]], cj[S[1]], cj[S[2]], cj[S[3]]};
    for (int i = 0; i < 4; ++i) {
      zc s = 0;
      for (int k = 0; k < 4; ++k) s += E[i + 4 * k] * b[k];
      cj[S[i]] = s;
    }
  }
}

/* columns S of mat(n, ...): mat(:,S) = mat(:,S) * E */
static void cols_mix(zc *mat, int ld, int n, const int *S, const zc *E) {
  zc *c0 = mat + (size_t)ld * S[0], *c1 = mat + (size_t)ld * S[1];
  zc *c2 = mat + (size_t)ld * S[2], *c3 = mat + (size_t)ld * S[3];
  for (int i = 0; i < n; ++i) {
    zc b[4] = {c0[i], c1[i], c2[i], c3[i]};
    zc o[4];
    for (int a = 0; a < 4; ++a) {
      zc s = 0;
      for (int k = 0; k < 4; ++k) s += b[k] * E[k + 4 * a];
      o[a] = s;
    }
    c0[i] = o[0]; c1[i] = o[1]; c2[i] = o[2]; c3[i] = o[3];
  }
}

/* left_evolve: mat(ns,n) -> E_f mat [E_f^-1 if two_way] */
void orc_left_evolve(orc_model *m, int time, double *matd, int n, int pf_id, int two_way) {
  zc *mat = (zc *)matd;
  int f = pf_id - 1;
  for (int b = 0; b < m->npla; ++b)
    rows_mix(mat, m->Ns, n, &m->sites[(f * m->npla + b) * 4], EKV(m, m->ekv, f, b, time));
  if (two_way)
    for (int b = 0; b < m->npla; ++b)
      cols_mix(mat, m->Ns, m->Ns, &m->sites[(f * m->npla + b) * 4], EKV(m, m->ekvi, f, b, time));
}

/* right_evolve: mat(n,ns) -> mat E_f [E_f^-1 mat if two_way] */
void orc_right_evolve(orc_model *m, int time, double *matd, int n, int pf_id, int two_way) {
  zc *mat = (zc *)matd;
  int f = pf_id - 1;
  for (int b = 0; b < m->npla; ++b)
    cols_mix(mat, n, n, &m->sites[(f * m->npla + b) * 4], EKV(m, m->ekv, f, b, time));
  if (two_way)
    for (int b = 0; b < m->npla; ++b)
      rows_mix(mat, n, m->Ns, &m->sites[(f * m->npla + b) * 4], EKV(m, m->ekvi, f, b, time));
}

void orc_wrap_g(orc_model *m, int time, int pf_id, int forward) {
  if (forward)
    orc_left_evolve(m, time, (double *)m->g, m->Ns, pf_id, 1);
  else
    orc_right_evolve(m, time, (double *)m->g, m->Ns, pf_id, 1);
}

/* ------------------------------------------------------------------ */
/* G from scratch                                                      */
/* ------------------------------------------------------------------ */

static zc ln_branch(zc w) { /* cmplx(real, mod(aimag, 2*pi_SP)) */
  return CMPLX(creal(w), fmod(cimag(w), 2.0 * SP_PI));
}

/* cal_g_and_det_T0 of src/mod_evolution.f90:330-357 */
static void cal_g_and_det(orc_model *m, const zc *lmat /*ne x Ns*/, const zc *rmat /*Ns x ne*/) {
  int ne = m->ne, Ns = m->Ns;
  zc *pbp = xcalloc((size_t)ne * ne, sizeof(zc));
  gemm_nn(ne, ne, Ns, lmat, rmat, pbp);
  zc lw = 0;
  lw = lw + clog(zdet(ne, pbp));
  zc sd = 0;
  for (size_t i = 0; i < (size_t)ne * m->nblock; ++i) sd += clog(m->Dst[i]);
  lw = lw + sd;
  zinverse(ne, pbp);
  zc *buf = xcalloc((size_t)Ns * ne, sizeof(zc));
  gemm_nn(Ns, ne, ne, rmat, pbp, buf);
  gemm_nn(Ns, Ns, ne, buf, lmat, m->g);
  if (m->tr_weight) lw = lw + conj(lw);
  lw = lw * (double)m->ncopy;
  m->ln_cw = ln_branch(lw);
  free(buf);
  free(pbp);
}

static zc *QS(orc_model *m, int block /*1-based*/) { return m->Q + (size_t)m->Ns * m->ne * (block - 1); }
static zc *DS(orc_model *m, int block) { return m->Dst + (size_t)m->ne * (block - 1); }
static zc *RS(orc_model *m, int block) { return m->R + (size_t)m->ne * m->ne * (block - 1); }

static void conj_transpose(int r, int c, const zc *a, zc *out) { /* a(r,c) -> out(c,r) */
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) out[j + (size_t)c * i] = conj(a[i + (size_t)r * j]);
}

void orc_set_trial(orc_model *m, const double *slater) {
  memcpy(m->slQ, slater, sizeof(zc) * m->Ns * m->ne);
  zqdr(m->Ns, m->ne, m->slQ, m->slR, m->slD);
}

/* ------------------------------------------------------------------ */
/* finite temperature (src/mod_evolution.f90:411-785)                 */
/* ------------------------------------------------------------------ */
/* The reference's FT branch is marked "not used for now" and is not
 * self-consistent (init_g_FT's active call forms (1 + B B^H)^-1, the
 * commented call beside it (1 + B)^-1; the left string's unitary and
 * triangular factors are stored in swapped roles), so this restates the
 * quantities it is built to compute, with its structure: right strings
 * B(t,0) = Q D T from column-scaled QR (Q' = Q*D, then a UDV step), left
 * strings B(beta,t) = T D Q from row-scaled LQ, and G from the scale-split
 * formula of cal_g_and_det_FT (:494-648).  The reference's UDV is an SVD
 * (ZGESVD, src/mod_matrixlib.f90:704-743); a pre-pivoted QR (ft_prepivot)
 * gives the same G and ln det(1 + B) with graded scales, as the SVD does. */

static zc *FQS(orc_model *m, int b) { return m->FQ + (size_t)m->Ns * m->Ns * (b - 1); }
static zc *FDS(orc_model *m, int b) { return m->FD + (size_t)m->Ns * (b - 1); }
static zc *FTS(orc_model *m, int b) { return m->FT + (size_t)m->Ns * m->Ns * (b - 1); }

void orc_set_ft(orc_model *m, int on) {
  m->ft = on ? 1 : 0;
  if (m->ft && !m->FQ) {
    size_t n = (size_t)m->Ns * m->Ns * m->nblock;
    m->FQ = xcalloc(n, sizeof(zc));
    m->FT = xcalloc(n, sizeof(zc));
    m->FD = xcalloc((size_t)m->Ns * m->nblock, sizeof(zc));
  }
}

static void ft_identity(orc_model *m, int b) {
  int Ns = m->Ns;
  memset(FQS(m, b), 0, sizeof(zc) * Ns * Ns);
  memset(FTS(m, b), 0, sizeof(zc) * Ns * Ns);
  for (int i = 0; i < Ns; ++i) {
    FQS(m, b)[i + (size_t)Ns * i] = 1;
    FTS(m, b)[i + (size_t)Ns * i] = 1;
    FDS(m, b)[i] = 1;
  }
}

/* One string step, right (B Q D) or left (D Q B), factored so that no
 * intermediate squares a scale: scales of B(beta,0) reach exp(beta * band
 * edge) (exp(408) at beta = 100 here), and a Householder norm or a complex
 * division of such a number overflows.
 *
 * - Each scale splits exactly as D_j = Dh_j 2^e_j with |Dh_j| in [0.5, 1)
 *   (frexp); the factorization runs on X Dh (moderate entries), and the
 *   powers of two go back exactly afterwards.  Householder QR is invariant
 *   under power-of-two column scaling, so this is the QR of X D.
 * - Pre-pivoting (Bai, Lee, Li, Xu, "Stable solutions of linear systems
 *   involving long chain of matrix multiplications", Linear Algebra Appl.
 *   435 (2011) 659-673): the columns of X D are ordered by decreasing
 *   2-norm before the unpivoted QR, so that the new D comes out graded and
 *   T = D^-1 R stays bounded.  The reference's UDV is an SVD, whose D is
 *   sorted; an unpivoted QR leaves ratios of scales in T.
 *
 * Squared norms are compared as 4^e_j v_j with v_j = |Dh_j|^2 sum_i |X_ij|^2,
 * through ldexp (exact), ties by index.  perm[k] = the column taken to
 * position k, ep[k] = e_perm[k]. */
static void ft_prepivot(int n, const double *v, const int *ex, int *perm, int *ep) {
  for (int j = 0; j < n; ++j) {
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      double vi = ldexp(v[i], 2 * (ex[i] - ex[j]));
      rank += (vi > v[j]) || (vi == v[j] && i < j);
    }
    perm[rank] = j;
    ep[rank] = ex[j];
  }
}

/* v_j and e_j of column j (left: row j) of X against the scales D */
static void ft_keys(int n, const zc *x, const zc *D, int left, double *v, int *ex) {
  for (int j = 0; j < n; ++j) {
    double ss = 0.0;
    for (int i = 0; i < n; ++i) {
      zc a = left ? x[j + (size_t)n * i] : x[i + (size_t)n * j];
      ss += creal(a) * creal(a) + cimag(a) * cimag(a);
    }
    double m = frexp(cabs(D[j]), &ex[j]);
    v[j] = m * m * ss;
  }
}

/* right string of block b from block b-1 and the slices t0..t1:
 * B Q_{b-1} D_{b-1} P = Q_b D_b T'; T_b = T' P^T T_{b-1} */
static void ft_right(orc_model *m, int b, int t0, int t1) {
  int Ns = m->Ns;
  zc *x = xcalloc((size_t)Ns * Ns, sizeof(zc)), *r = xcalloc((size_t)Ns * Ns, sizeof(zc));
  zc *xp = xcalloc((size_t)Ns * Ns, sizeof(zc)), *tg = xcalloc((size_t)Ns * Ns, sizeof(zc));
  double *v = xcalloc(Ns, sizeof(double));
  int *ex = xcalloc(Ns, sizeof(int)), *perm = xcalloc(Ns, sizeof(int)), *ep = xcalloc(Ns, sizeof(int));
  const zc *D = FDS(m, b - 1);
  memcpy(x, FQS(m, b - 1), sizeof(zc) * Ns * Ns);
  for (int p = t0; p <= t1; ++p)
    for (int f = 1; f <= 4; ++f) orc_left_evolve(m, p, (double *)x, Ns, f, 0);
  ft_keys(Ns, x, D, 0, v, ex);
  ft_prepivot(Ns, v, ex, perm, ep);
  for (int k = 0; k < Ns; ++k) { /* (X D P) 2^-ep: column k = X(:, perm[k]) Dh_perm[k] */
    zc dh = CMPLX(ldexp(creal(D[perm[k]]), -ep[k]), ldexp(cimag(D[perm[k]]), -ep[k]));
    for (int i = 0; i < Ns; ++i) xp[i + (size_t)Ns * k] = x[i + (size_t)Ns * perm[k]] * dh;
  }
  zqdr(Ns, Ns, xp, r, FDS(m, b));
  for (int k = 0; k < Ns; ++k) {
    zc *d = &FDS(m, b)[k];
    *d = CMPLX(ldexp(creal(*d), ep[k]), ldexp(cimag(*d), ep[k]));
    for (int j = k + 1; j < Ns; ++j) {
      zc *t = &r[k + (size_t)Ns * j];
      *t = CMPLX(ldexp(creal(*t), ep[j] - ep[k]), ldexp(cimag(*t), ep[j] - ep[k]));
    }
  }
  memcpy(FQS(m, b), xp, sizeof(zc) * Ns * Ns);
  const zc *tp = FTS(m, b - 1);
  for (int j = 0; j < Ns; ++j) /* rows of T_{b-1} in pivot order: P^T T_{b-1} */
    for (int k = 0; k < Ns; ++k) tg[k + (size_t)Ns * j] = tp[perm[k] + (size_t)Ns * j];
  gemm_nn(Ns, Ns, Ns, r, tg, FTS(m, b));
  free(ep);
  free(perm);
  free(ex);
  free(v);
  free(tg);
  free(xp);
  free(r);
  free(x);
}

/* left string of block b from block b+1 and the slices t1 down to t0, rows
 * pre-pivoted: P^T D_{b+1} Q_{b+1} B = L' D_b Q_b; T_b = T_{b+1} P L' */
static void ft_left(orc_model *m, int b, int t0, int t1) {
  int Ns = m->Ns;
  zc *y = xcalloc((size_t)Ns * Ns, sizeof(zc)), *l = xcalloc((size_t)Ns * Ns, sizeof(zc));
  zc *yp = xcalloc((size_t)Ns * Ns, sizeof(zc)), *tg = xcalloc((size_t)Ns * Ns, sizeof(zc));
  double *v = xcalloc(Ns, sizeof(double));
  int *ex = xcalloc(Ns, sizeof(int)), *perm = xcalloc(Ns, sizeof(int)), *ep = xcalloc(Ns, sizeof(int));
  const zc *D = FDS(m, b + 1);
  memcpy(y, FQS(m, b + 1), sizeof(zc) * Ns * Ns);
  for (int p = t1; p >= t0; --p)
    for (int f = 4; f >= 1; --f) orc_right_evolve(m, p, (double *)y, Ns, f, 0);
  ft_keys(Ns, y, D, 1, v, ex);
  ft_prepivot(Ns, v, ex, perm, ep);
  for (int k = 0; k < Ns; ++k) { /* (P^T D Q B) 2^-ep: row k = Dh_perm[k] Y(perm[k], :) */
    zc dh = CMPLX(ldexp(creal(D[perm[k]]), -ep[k]), ldexp(cimag(D[perm[k]]), -ep[k]));
    for (int j = 0; j < Ns; ++j) yp[k + (size_t)Ns * j] = dh * y[perm[k] + (size_t)Ns * j];
  }
  zldq(Ns, Ns, yp, l, FDS(m, b));
  for (int k = 0; k < Ns; ++k) {
    zc *d = &FDS(m, b)[k];
    *d = CMPLX(ldexp(creal(*d), ep[k]), ldexp(cimag(*d), ep[k]));
    for (int i = k + 1; i < Ns; ++i) {
      zc *t = &l[i + (size_t)Ns * k];
      *t = CMPLX(ldexp(creal(*t), ep[i] - ep[k]), ldexp(cimag(*t), ep[i] - ep[k]));
    }
  }
  memcpy(FQS(m, b), yp, sizeof(zc) * Ns * Ns);
  const zc *tn = FTS(m, b + 1);
  for (int k = 0; k < Ns; ++k) memcpy(tg + (size_t)Ns * k, tn + (size_t)Ns * perm[k], sizeof(zc) * Ns);
  gemm_nn(Ns, Ns, Ns, tg, l, FTS(m, b)); /* T_b = (T_{b+1} P) L' */
  free(ep);
  free(perm);
  free(ex);
  free(v);
  free(tg);
  free(yp);
  free(l);
  free(y);
}

/* G = (1 + Q_r D_r T_r T_l D_l Q_l)^-1 with D split at |d| = 1
 * (cal_g_and_det_FT, src/mod_evolution.f90:494-648); g_h = 1 - G;
 * ln_cw = ln det(1 + B(beta,0)) */
static void ft_cal_g(orc_model *m, int br, int bl) {
  int Ns = m->Ns;
  const zc *Qr = FQS(m, br), *Dr = FDS(m, br), *Tr = FTS(m, br);
  const zc *Ql = FQS(m, bl), *Dl = FDS(m, bl), *Tl = FTS(m, bl);
  zc *dbr = xcalloc(Ns, sizeof(zc)), *dsr = xcalloc(Ns, sizeof(zc));
  zc *dbl = xcalloc(Ns, sizeof(zc)), *dsl = xcalloc(Ns, sizeof(zc));
  for (int i = 0; i < Ns; ++i) {
    if (cabs(Dr[i]) > 1.0) { dbr[i] = Dr[i]; dsr[i] = 1; } else { dbr[i] = 1; dsr[i] = Dr[i]; }
    if (cabs(Dl[i]) > 1.0) { dbl[i] = Dl[i]; dsl[i] = 1; } else { dbl[i] = 1; dsl[i] = Dl[i]; }
  }
  zc *qq = xcalloc((size_t)Ns * Ns, sizeof(zc)), *tt = xcalloc((size_t)Ns * Ns, sizeof(zc));
  zc *mm = xcalloc((size_t)Ns * Ns, sizeof(zc));
  gemm_nn(Ns, Ns, Ns, Ql, Qr, qq); /* (Q_r^H Q_l^H) = (Q_l Q_r)^H */
  gemm_nn(Ns, Ns, Ns, Tr, Tl, tt);
  for (int j = 0; j < Ns; ++j)
    for (int i = 0; i < Ns; ++i)
      mm[i + (size_t)Ns * j] = conj(qq[j + (size_t)Ns * i]) / dbr[i] / dbl[j] + dsr[i] * tt[i + (size_t)Ns * j] * dsl[j];
  zc lw = clog(zdet(Ns, mm));
  zc lq = clog(zdet(Ns, Qr)) + clog(zdet(Ns, Ql));
  for (int i = 0; i < Ns; ++i) lw += clog(dbr[i]) + clog(dbl[i]);
  lw += lq;
  zinverse(Ns, mm);
  /* G = Q_l^H Db_l^-1 M^-1 Db_r^-1 Q_r^H */
  for (int j = 0; j < Ns; ++j)
    for (int i = 0; i < Ns; ++i) mm[i + (size_t)Ns * j] = mm[i + (size_t)Ns * j] / dbl[i] / dbr[j];
  zc *qlh = xcalloc((size_t)Ns * Ns, sizeof(zc)), *qrh = xcalloc((size_t)Ns * Ns, sizeof(zc));
  conj_transpose(Ns, Ns, Ql, qlh);
  conj_transpose(Ns, Ns, Qr, qrh);
  gemm_nn(Ns, Ns, Ns, qlh, mm, tt);
  gemm_nn(Ns, Ns, Ns, tt, qrh, m->g);
  for (size_t e = 0; e < (size_t)Ns * Ns; ++e) m->g[e] = -m->g[e];
  for (int i = 0; i < Ns; ++i) m->g[i + (size_t)Ns * i] += 1;
  if (m->tr_weight) lw = lw + conj(lw);
  m->ln_cw = ln_branch(lw * (double)m->ncopy);
  free(qlh); free(qrh); free(qq); free(tt); free(mm);
  free(dbr); free(dsr); free(dbl); free(dsl);
}

/* init_g_FT: right strings of blocks 2..nblock-1 from the identity at block
 * 1, the identity as the left string at block nblock, G at ntime+1 */
static void ft_init(orc_model *m, int start_time) {
  int ng = m->ngroup, nb = m->nblock;
  (void)start_time;
  ft_identity(m, 1);
  ft_identity(m, nb);
  for (int b = 2; b <= nb - 1; ++b) ft_right(m, b, (b - 2) * ng + 1, (b - 1) * ng);
  ft_cal_g(m, nb - 1, nb);
}

/* get_g_scratch_FT (src/mod_evolution.f90:651-740) */
static void ft_scratch(orc_model *m, int time, int forward) {
  int ng = m->ngroup;
  if (forward) {
    int block = time / ng + 1;
    ft_right(m, block, time - ng, time - 1);
    ft_cal_g(m, block, block + 1);
  } else {
    int block = time / ng + 2;
    if (ng == 1) block -= 1;
    ft_left(m, block, time, time + ng - 1);
    ft_cal_g(m, block - 1, block);
  }
}

/* init_g_T0 of src/mod_evolution.f90:46-111 */
void orc_init_g(orc_model *m, int start_time) {
  if (m->ft) {
    m->ln_cw = 0;
    ft_init(m, start_time);
    return;
  }
  int Ns = m->Ns, ne = m->ne, ng = m->ngroup, nb = m->nblock;
  m->ln_cw = 0;
  memcpy(QS(m, 1), m->slQ, sizeof(zc) * Ns * ne);
  memcpy(QS(m, nb), m->slQ, sizeof(zc) * Ns * ne);
  memcpy(DS(m, 1), m->slD, sizeof(zc) * ne);
  memcpy(DS(m, nb), m->slD, sizeof(zc) * ne);
  memcpy(RS(m, 1), m->slR, sizeof(zc) * ne * ne);
  memcpy(RS(m, nb), m->slR, sizeof(zc) * ne * ne);
  int start_block = start_time / ng + 1;
  int stb = (start_block - 1) * ng + 1;
  if (stb > start_time) abort();
  zc *rmat = xcalloc((size_t)Ns * ne, sizeof(zc));
  zc *lmat = xcalloc((size_t)Ns * ne, sizeof(zc));
  memcpy(rmat, QS(m, start_block), sizeof(zc) * Ns * ne);
  for (int p = stb; p <= m->Lt; ++p) {
    for (int f = 1; f <= 4; ++f) orc_left_evolve(m, p, (double *)rmat, ne, f, 0);
    if (p % ng != 0) continue;
    int block = (p - 1) / ng + 2;
    zqdr(Ns, ne, rmat, RS(m, block), DS(m, block));
    memcpy(QS(m, block), rmat, sizeof(zc) * Ns * ne);
  }
  conj_transpose(Ns, ne, QS(m, nb), lmat);
  cal_g_and_det(m, lmat, rmat);
  free(lmat);
  free(rmat);
}

/* get_g_scratch_T0 of src/mod_evolution.f90:113-195 */
void orc_get_g_scratch(orc_model *m, int time, int forward, double *err_out) {
  int Ns = m->Ns, ne = m->ne, ng = m->ngroup;
  zc ln_fast = m->ln_cw;
  m->ln_cw = 0;
  if (m->ft) {
    zc *gf = xcalloc((size_t)Ns * Ns, sizeof(zc));
    memcpy(gf, m->g, sizeof(zc) * Ns * Ns);
    ft_scratch(m, time, forward);
    double err = 0;
    for (size_t i = 0; i < (size_t)Ns * Ns; ++i) {
      double v = cabs(gf[i] - m->g[i]);
      if (v > err) err = v;
    }
    m->err_fast = err;
    if (err > m->err_fast_max) m->err_fast_max = err;
    double e2 = fabs(creal(m->ln_cw - ln_fast));
    if (e2 > m->err_fast_max) m->err_fast_max = e2;
    if (err_out) *err_out = err;
    free(gf);
    return;
  }
  zc *rmat = xcalloc((size_t)Ns * ne, sizeof(zc));
  zc *lmat = xcalloc((size_t)Ns * ne, sizeof(zc));
  zc *gfast = xcalloc((size_t)Ns * Ns, sizeof(zc));
  memcpy(gfast, m->g, sizeof(zc) * Ns * Ns);
  int block;
  if (forward) {
    block = time / ng + 1;
    memcpy(rmat, QS(m, block - 1), sizeof(zc) * Ns * ne);
    for (int p = time - ng; p <= time - 1; ++p)
      for (int f = 1; f <= 4; ++f) orc_left_evolve(m, p, (double *)rmat, ne, f, 0);
    zqdr(Ns, ne, rmat, RS(m, block), DS(m, block));
    memcpy(QS(m, block), rmat, sizeof(zc) * Ns * ne);
    conj_transpose(Ns, ne, QS(m, block + 1), lmat);
  } else {
    block = time / ng + 2;
    if (ng == 1) block = block - 1;
    conj_transpose(Ns, ne, QS(m, block + 1), lmat);
    for (int p = time + ng - 1; p >= time; --p)
      for (int f = 4; f >= 1; --f) orc_right_evolve(m, p, (double *)lmat, ne, f, 0);
    zldq(ne, Ns, lmat, RS(m, block), DS(m, block));
    conj_transpose(ne, Ns, lmat, QS(m, block));
    memcpy(rmat, QS(m, block - 1), sizeof(zc) * Ns * ne);
  }
  cal_g_and_det(m, lmat, rmat);
  double err = 0;
  for (size_t i = 0; i < (size_t)Ns * Ns; ++i) {
    double v = cabs(gfast[i] - m->g[i]);
    if (v > err) err = v;
  }
  m->err_fast = err;
  if (err > m->err_fast_max) m->err_fast_max = err;
  double e2 = fabs(creal(m->ln_cw - ln_fast));
  if (e2 > m->err_fast_max) m->err_fast_max = e2;
  if (err_out) *err_out = err;
  free(gfast);
  free(lmat);
  free(rmat);
}

void orc_get_g(const orc_model *m, double *g) { memcpy(g, m->g, sizeof(zc) * m->Ns * m->Ns); }
void orc_set_g(orc_model *m, const double *g) { memcpy(m->g, g, sizeof(zc) * m->Ns * m->Ns); }
void orc_get_qstring(const orc_model *m, int block, double *q) {
  memcpy(q, m->Q + (size_t)m->Ns * m->ne * (block - 1), sizeof(zc) * m->Ns * m->ne);
}
void orc_get_dstring(const orc_model *m, double *d) { memcpy(d, m->Dst, sizeof(zc) * m->ne * m->nblock); }
void orc_get_lncw(const orc_model *m, double *ri) { ri[0] = creal(m->ln_cw); ri[1] = cimag(m->ln_cw); }
void orc_get_counters(const orc_model *m, long *pa, long *na, double *efm) {
  *pa = m->pos_acc; *na = m->neg_acc; *efm = m->err_fast_max;
}

/* ------------------------------------------------------------------ */
/* local update (src/mod_update.f90:13-245, 475-500)                   */
/* ------------------------------------------------------------------ */

/* cal_energy_difference (:203-228) -> -delt*(E_2-E_1) */
static double delta_action(const orc_model *m, int field, int time, double jump) {
  const double *x = m->x;
  size_t Nb = m->Nb;
  double be = x[field + Nb * (modi(time - 1, m->Lt) - 1)];
  double af = x[field + Nb * (modi(time + 1, m->Lt) - 1)];
  double now = x[field + Nb * (time - 1)];
  double nw = now + jump;
  double dt2 = m->dtau * m->dtau;
  double e1 = 0.5 * m->M * ((be - now) * (be - now) + (af - now) * (af - now)) / dt2 +
              0.5 * m->D * (now * now) - m->biased * now;
  double e2 = 0.5 * m->M * ((be - nw) * (be - nw) + (af - nw) * (af - nw)) / dt2 +
              0.5 * m->D * (nw * nw) - m->biased * nw;
  return -m->dtau * (e2 - e1);
}

/* cal_prob (:475-500), Metropolis branch */
static double cal_prob(zc w) {
  w = CMPLX(creal(w), fmod(cimag(w), 2.0 * SP_PI));
  double rt = exp(fmin(10.0, creal(w)));
  return fmin(1.0, rt);
}

void orc_update_phonon_field(orc_model *m, int pf_id, int time, const float *u, uint8_t *acc) {
  int f = pf_id - 1, Ns = m->Ns;
  zc *g = m->g;
  zc *T = xcalloc((size_t)Ns * 4, sizeof(zc));
  zc *P = xcalloc((size_t)Ns * 4, sizeof(zc));
  for (int b = 0; b < m->npla; ++b) {
    const int *S = &m->sites[(f * m->npla + b) * 4];
    int field = m->bfi[f * m->npla + b];
    double *pbf = &m->x[field + (size_t)m->Nb * (time - 1)];
    double old_bf = *pbf;
    zc gs[16];
    for (int j = 0; j < 4; ++j)
      for (int i = 0; i < 4; ++i) gs[i + 4 * j] = g[S[i] + (size_t)Ns * S[j]];
    /* ran_sysm = ran2*2-1 in REAL(SP) (src/mod_ranstate.f90:233-236) */
    float rs = u[2 * b] * 2.0f - 1.0f;
    double jump = m->jump_distance * (double)rs;
    double dS = delta_action(m, field, time, jump);
    *pbf = *pbf + jump;
    /* calculate_Delta (:230-245) */
    zc enew[16], delta[16], rm[16];
    expkv(m, f, b, *pbf, 0, enew);
    mm4(delta, EKV(m, m->ekvi, f, b, time), enew);
    for (int i = 0; i < 4; ++i) delta[i + 4 * i] -= 1.0;
    mm4(rm, gs, delta);
    for (int i = 0; i < 4; ++i) rm[i + 4 * i] += 1.0;
    zc relec = 1.0;
    relec = relec * det4(rm);
    if (m->tr_weight) relec = relec * conj(relec);
    if (m->ncopy != 1) relec = cpow(relec, m->ncopy);
    zc w = dS + clog(relec);
    double prob = cal_prob(w);
    int accept = !(prob < (double)u[2 * b + 1]);
    if (!accept) {
      *pbf = *pbf - jump;
      *pbf = old_bf;
      m->neg_acc++;
      if (acc) acc[b] = 0;
      continue;
    }
    expkv(m, f, b, *pbf, 0, EKV(m, m->ekv, f, b, time));
    expkv(m, f, b, *pbf, 1, EKV(m, m->ekvi, f, b, time));
    m->ln_cw = m->ln_cw + log(creal(relec));
    m->pos_acc++;
    if (acc) acc[b] = 1;
    /* final update (:146-187): Delta from the old inverse and new expKV */
    zc eold_inv[16], d2[16], r2[16], x4[16];
    expkv(m, f, b, old_bf, 1, eold_inv);
    mm4(d2, eold_inv, EKV(m, m->ekv, f, b, time));
    for (int i = 0; i < 4; ++i) d2[i + 4 * i] -= 1.0;
    mm4(r2, gs, d2);
    for (int i = 0; i < 4; ++i) r2[i + 4 * i] += 1.0;
    zinverse(4, r2);
    mm4(x4, d2, r2);
    /* gh_part_nd = g_h(:,S); g_part_dn(i,:) = -g_h(S(i),:) + e_S(i) */
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < Ns; ++i) {
        zc s = 0;
        for (int q = 0; q < 4; ++q) s += g[i + (size_t)Ns * S[q]] * x4[q + 4 * k];
        T[i + (size_t)Ns * k] = s;
      }
    for (int j = 0; j < Ns; ++j)
      for (int k = 0; k < 4; ++k) P[k + 4 * (size_t)j] = -g[S[k] + (size_t)Ns * j];
    for (int k = 0; k < 4; ++k) P[k + 4 * (size_t)S[k]] += 1.0;
    for (int j = 0; j < Ns; ++j) {
      zc *gj = g + (size_t)Ns * j;
      zc p0 = P[0 + 4 * (size_t)j], p1 = P[1 + 4 * (size_t)j], p2 = P[2 + 4 * (size_t)j],
         p3 = P[3 + 4 * (size_t)j];
      const zc *t0 = T, *t1 = T + Ns, *t2 = T + 2 * (size_t)Ns, *t3 = T + 3 * (size_t)Ns;
      for (int i = 0; i < Ns; ++i) {
        zc v = gj[i];
        v += t0[i] * p0;
        v += t1[i] * p1;
        v += t2[i] * p2;
        v += t3[i] * p3;
        gj[i] = v;
      }
    }
  }
  free(P);
  free(T);
}

/* ------------------------------------------------------------------ */
/* measurement (src/mod_meas.f90:165-183, 186-208, 210-566)            */
/* ------------------------------------------------------------------ */

enum { /* registration order, EPSOCZ.nml:262-283 */
       O_BF_KE, O_BF_PE, O_BF_X,
       F_BFxBFx, F_BFxBFy, F_BFyBFy, F_BFxyBFxy, F_JxxJxx, F_JxyJxy, F_JyxJyx, F_JyyJyy,
       F_JxxJyx, F_JxyJyy, F_JxxJyy, F_JxyJyx, F_cJxy, F_cJyx, F_den, F_SC, F_Sz, F_Sx, F_Sy,
       N_FIELD_END };

static size_t field_off(const orc_model *m, int fobs) { /* fobs >= F_BFxBFx */
  return 3 + (size_t)(fobs - F_BFxBFx) * 2 * m->Ncell;
}

typedef struct { const zc *gh; const zc *gg; int Ns; } wick_ctx;

/* cal_wick_contraction (:165-183): mat(i,j) = gh(l2,l1) gh(r2,r1) + gh(r2,l1) g(l2,r1) */
static void wick(const wick_ctx *w, int n, const int (*ls)[2], const int (*rs)[2], zc *mat) {
  int Ns = w->Ns;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      zc v = 0;
      v = v + w->gh[ls[i][1] + (size_t)Ns * ls[i][0]] * w->gh[rs[j][1] + (size_t)Ns * rs[j][0]];
      v = v + w->gh[rs[j][1] + (size_t)Ns * ls[i][0]] * w->gg[ls[i][1] + (size_t)Ns * rs[j][0]];
      mat[i + n * j] = v;
    }
}

/* sum(vec_l * matmul(mat, vec_r)) */
static zc contract(int n, const zc *mat, const zc *vl, const zc *vr) {
  zc t[8];
  for (int i = 0; i < n; ++i) {
    zc s = 0;
    for (int j = 0; j < n; ++j) s += mat[i + n * j] * vr[j];
    t[i] = s;
  }
  zc s = 0;
  for (int i = 0; i < n; ++i) s += vl[i] * t[i];
  return s;
}

/* spin_current_matrix (:64-87): sites(1)=up, sites(2)=dn of cells a,b,c,d */
static void spin_current(const wick_ctx *w, int a, int b, int c, int d, zc *mat) {
  int ls[4][2] = {{2 * a, 2 * b + 1}, {2 * a + 1, 2 * b}, {2 * b, 2 * a + 1}, {2 * b + 1, 2 * a}};
  int rs[4][2] = {{2 * c, 2 * d + 1}, {2 * c + 1, 2 * d}, {2 * d, 2 * c + 1}, {2 * d + 1, 2 * c}};
  wick(w, 4, (const int(*)[2])ls, (const int(*)[2])rs, mat);
}

/* spin_momentum_matrix (:89-108) */
static void spin_momentum(const wick_ctx *w, int a, int c, zc *mat) {
  int ls[4][2] = {{2 * a, 2 * a}, {2 * a, 2 * a + 1}, {2 * a + 1, 2 * a}, {2 * a + 1, 2 * a + 1}};
  int rs[4][2] = {{2 * c, 2 * c}, {2 * c, 2 * c + 1}, {2 * c + 1, 2 * c}, {2 * c + 1, 2 * c + 1}};
  wick(w, 4, (const int(*)[2])ls, (const int(*)[2])rs, mat);
}

/* physical_spin_current_matrix (:129-163) */
static void phys_current(const wick_ctx *w, int a, int b, int c, int d, zc *mat) {
  int ls[8][2] = {{2 * a, 2 * b + 1}, {2 * a + 1, 2 * b}, {2 * b, 2 * a + 1}, {2 * b + 1, 2 * a},
                  {2 * a, 2 * b},     {2 * a + 1, 2 * b + 1}, {2 * b, 2 * a}, {2 * b + 1, 2 * a + 1}};
  int rs[8][2] = {{2 * c, 2 * d + 1}, {2 * c + 1, 2 * d}, {2 * d, 2 * c + 1}, {2 * d + 1, 2 * c},
                  {2 * c, 2 * d},     {2 * c + 1, 2 * d + 1}, {2 * d, 2 * c}, {2 * d + 1, 2 * c + 1}};
  wick(w, 8, (const int(*)[2])ls, (const int(*)[2])rs, mat);
}

/* add_boundary_phase (:199-208) */
static void bphase(const orc_model *m, zc *v, int crosses, int dir) {
  if (!crosses) return;
  zc p = m->bc[dir];
  v[0] *= p; v[1] *= p; v[2] *= conj(p); v[3] *= conj(p);
}

/* ms_take_measurement (:210-544) followed by ms_k_space_obs (:546-566) */
void orc_measure(orc_model *m, int time, double *outd) {
  zc *out = (zc *)outd;
  int Ns = m->Ns, Nc = m->Ncell, L1 = m->L1, L2 = m->L2;
  memset(out, 0, sizeof(zc) * m->total_width);
  zc *gg = xcalloc((size_t)Ns * Ns, sizeof(zc));
  for (int j = 0; j < Ns; ++j) {
    for (int i = 0; i < Ns; ++i) gg[i + (size_t)Ns * j] = -m->g[i + (size_t)Ns * j];
    gg[j + (size_t)Ns * j] = gg[j + (size_t)Ns * j] + 1.0;
  }
  wick_ctx w = {m->g, gg, Ns};
  const double *xt = m->x + (size_t)m->Nb * (time - 1);
  const double *xt1 = m->x + (size_t)m->Nb * (time); /* time+1 (BF_PE) */
  double hop = m->hop, lam = m->lam;
  const zc vjx[4] = {-hop * (-I), -hop * (-I), -hop * I, -hop * I};
  const zc vjy[4] = {-hop * -1.0, -hop * 1.0, -hop * 1.0, -hop * -1.0};
  for (int c1 = 0; c1 < Nc; ++c1) {
    int d1[2] = {c1 % L1, c1 / L1};
    int t[2];
    t[0] = d1[0] + 1; t[1] = d1[1];
    int c1x = uc_index(m, t);
    t[0] = d1[0]; t[1] = d1[1] + 1;
    int c1y = uc_index(m, t);
    int bf1x = 2 * c1, bf1y = 2 * c1 + 1;
    zc o;
    o = 0;
    o = o + 0.5 * m->D * (xt[bf1x] * xt[bf1x] + xt[bf1y] * xt[bf1y]) - m->biased * (xt[bf1x] + xt[bf1y]);
    o = o / (double)Nc;
    out[O_BF_KE] += o;
    o = 0;
    {
      double a = xt[bf1x] - xt1[bf1x], b = xt[bf1y] - xt1[bf1y];
      o = o - 0.5 * m->M * (a * a + b * b) / (m->dtau * m->dtau);
    }
    o = o + 1.0 / (2.0 * m->dtau) * 2;
    o = o / (double)Nc;
    out[O_BF_PE] += o;
    o = 0;
    o = o + (xt[bf1x] + xt[bf1y]) / (double)Nc;
    out[O_BF_X] += o;
    int cb1[2] = {d1[0] == L1 - 1, d1[1] == L2 - 1};
    for (int c2 = 0; c2 < Nc; ++c2) {
      int d2[2] = {c2 % L1, c2 / L1};
      t[0] = d2[0] + 1; t[1] = d2[1];
      int c2x = uc_index(m, t);
      t[0] = d2[0]; t[1] = d2[1] + 1;
      int c2y = uc_index(m, t);
      int cb2[2] = {d2[0] == L1 - 1, d2[1] == L2 - 1};
      int bf2x = 2 * c2, bf2y = 2 * c2 + 1;
      /* get_relative_index (src/mod_lattice.f90:282-306) */
      int rel[2] = {(d2[0] - d1[0] + L1) % L1, (d2[1] - d1[1] + L2) % L2};
      int ind = uc_index(m, rel);
      zc mat[64], vl[8], vr[8];
      out[field_off(m, F_BFxBFx) + ind] += xt[bf1x] * xt[bf2x] / (double)Nc;
      out[field_off(m, F_BFxBFy) + ind] += xt[bf1x] * xt[bf2y] / (double)Nc;
      out[field_off(m, F_BFyBFy) + ind] += xt[bf1y] * xt[bf2y] / (double)Nc;
      out[field_off(m, F_BFxyBFxy) + ind] += xt[bf1x] * xt[bf1y] * xt[bf2x] * xt[bf2y] / (double)Nc;
      /* x-bond currents */
      spin_current(&w, c1, c1x, c2, c2x, mat);
      memcpy(vl, vjx, sizeof(vjx)); memcpy(vr, vjx, sizeof(vjx));
      bphase(m, vl, cb1[0], 0); bphase(m, vr, cb2[0], 0);
      out[field_off(m, F_JxxJxx) + ind] += contract(4, mat, vl, vr) / (double)Nc;
      memcpy(vl, vjy, sizeof(vjy)); memcpy(vr, vjy, sizeof(vjy));
      bphase(m, vl, cb1[0], 0); bphase(m, vr, cb2[0], 0);
      out[field_off(m, F_JxyJxy) + ind] += contract(4, mat, vl, vr) / (double)Nc;
      /* y-bond currents */
      spin_current(&w, c1, c1y, c2, c2y, mat);
      memcpy(vl, vjx, sizeof(vjx)); memcpy(vr, vjx, sizeof(vjx));
      bphase(m, vl, cb1[1], 1); bphase(m, vr, cb2[1], 1);
      out[field_off(m, F_JyxJyx) + ind] += contract(4, mat, vl, vr) / (double)Nc;
      memcpy(vl, vjy, sizeof(vjy)); memcpy(vr, vjy, sizeof(vjy));
      bphase(m, vl, cb1[1], 1); bphase(m, vr, cb2[1], 1);
      /* Jyy_Jyy contracts with vec_l on both sides (:353) */
      out[field_off(m, F_JyyJyy) + ind] += contract(4, mat, vl, vl) / (double)Nc;
      /* x-bond vs y-bond */
      spin_current(&w, c1, c1x, c2, c2y, mat);
      memcpy(vl, vjx, sizeof(vjx)); memcpy(vr, vjx, sizeof(vjx));
      bphase(m, vl, cb1[0], 0); bphase(m, vr, cb2[1], 1);
      out[field_off(m, F_JxxJyx) + ind] += contract(4, mat, vl, vr) / (double)Nc;
      memcpy(vl, vjy, sizeof(vjy)); memcpy(vr, vjy, sizeof(vjy));
      bphase(m, vl, cb1[0], 0); bphase(m, vr, cb2[1], 1);
      out[field_off(m, F_JxyJyy) + ind] += contract(4, mat, vl, vr) / (double)Nc;
      /* Jxx_Jyy and Jxy_Jyx are registered but never recorded (the code
       * records only their Hartree-only variants, which are unregistered) */
      /* spin momentum */
      spin_momentum(&w, c1, c2, mat);
      {
        zc vd[4] = {0.5, 0, 0, 0.5};
        out[field_off(m, F_den) + ind] += contract(4, mat, vd, vd) / (double)Nc;
      }
      {
        int s1u = 2 * c1, s1d = 2 * c1 + 1, s2u = 2 * c2, s2d = 2 * c2 + 1;
        const zc *gh = m->g;
        out[field_off(m, F_SC) + ind] += (gh[s2u + (size_t)Ns * s1u] * gh[s2d + (size_t)Ns * s1d] -
                                          gh[s2d + (size_t)Ns * s1u] * gh[s2u + (size_t)Ns * s1d]) /
                                         (double)Nc;
      }
      {
        zc vz[4] = {0.5, 0, 0, -0.5};
        out[field_off(m, F_Sz) + ind] += contract(4, mat, vz, vz) / (double)Nc;
        zc vx[4] = {0, 0.5, 0.5, 0};
        out[field_off(m, F_Sx) + ind] += contract(4, mat, vx, vx) / (double)Nc;
        zc vy[4] = {0, I * -0.5, I * 0.5, 0};
        out[field_off(m, F_Sy) + ind] += contract(4, mat, vy, vy) / (double)Nc;
      }
      /* physical currents (8x8) */
      phys_current(&w, c1, c1y, c2, c2y, mat);
      for (int q = 0; q < 4; ++q) { vl[q] = hop * ((q < 2) ? -I : I); }
      for (int q = 4; q < 8; ++q) vl[q] = lam * xt[bf1y] * 1.0;
      for (int q = 0; q < 4; ++q) vr[q] = vl[q];
      for (int q = 4; q < 8; ++q) vr[q] = lam * xt[bf2y] * 1.0;
      bphase(m, vl, cb1[1], 1); bphase(m, vr, cb2[1], 1);
      bphase(m, vl + 4, cb1[1], 1); bphase(m, vr + 4, cb2[1], 1);
      out[field_off(m, F_cJyx) + ind] += contract(8, mat, vl, vr) / (double)Nc;
      phys_current(&w, c1, c1x, c2, c2x, mat);
      {
        const double sy[4] = {-1.0, 1.0, 1.0, -1.0};
        for (int q = 0; q < 4; ++q) vl[q] = hop * sy[q];
      }
      for (int q = 4; q < 8; ++q) vl[q] = lam * xt[bf1x] * 1.0;
      for (int q = 0; q < 4; ++q) vr[q] = vl[q];
      for (int q = 4; q < 8; ++q) vr[q] = lam * xt[bf2x] * 1.0;
      bphase(m, vl, cb1[0], 0); bphase(m, vr, cb2[0], 0);
      bphase(m, vl + 4, cb1[0], 0); bphase(m, vr + 4, cb2[0], 0);
      out[field_off(m, F_cJxy) + ind] += contract(8, mat, vl, vr) / (double)Nc;
    }
  }
  /* begin_measure: divide by 2*meas_number+1 = 1, then the k-space companions
   * k_data(k) = sum_i phase(k,i) r(i) / N_cell, phase = exp(i FBZ(k).rpos(i))
   * (src/mod_lattice.f90:164-205 for a unit square lattice, even or odd L) */
  double kx0 = 0, ky0 = 0;
  double two_pi = 2.0 * acos(-1.0);
  kx0 += (L1 % 2) * two_pi / L1 / 2.0; kx0 -= 0.5 * two_pi;
  ky0 += (L2 % 2) * two_pi / L2 / 2.0; ky0 -= 0.5 * two_pi;
  zc *ph = xcalloc((size_t)Nc * Nc, sizeof(zc));
  for (int k = 0; k < Nc; ++k) {
    double fk0 = two_pi / L1 * (k % L1) + kx0;
    double fk1 = two_pi / L2 * (k / L1) + ky0;
    for (int i = 0; i < Nc; ++i) {
      double r0 = i % L1, r1 = i / L1;
      ph[k + (size_t)Nc * i] = cexp(I * (fk0 * r0 + fk1 * r1));
    }
  }
  for (int fo = F_BFxBFx; fo < N_FIELD_END; ++fo) {
    zc *r = out + field_off(m, fo);
    zc *kd = r + Nc;
    for (int k = 0; k < Nc; ++k) {
      zc s = 0;
      for (int i = 0; i < Nc; ++i) s = s + ph[k + (size_t)Nc * i] * r[i];
      kd[k] += s / (double)Nc;
    }
  }
  free(ph);
  free(gg);
}

/* ------------------------------------------------------------------ */
/* sweep (src/Main_PQMC.f90:81-156, one `loop` iteration)              */
/* ------------------------------------------------------------------ */

void orc_sweep(orc_model *m, int forward, const float *u, int ntc, uint8_t *acc, double *meas_out,
               int *n_meas) {
  int Lt = m->Lt, npla = m->npla;
  if (ntc <= 0 || ntc > Lt) ntc = Lt;
  size_t off = 0;
  if (n_meas) *n_meas = 0;
  for (int tc = 1; tc <= ntc; ++tc) {
    int time;
    if (!forward) {
      time = Lt - tc + 1;
      for (int f = 4; f >= 1; --f) {
        orc_right_evolve(m, time, (double *)m->g, m->Ns, f, 1);
        orc_update_phonon_field(m, f, time, u + off, acc ? acc + off / 2 : NULL);
        off += 2 * (size_t)npla;
      }
    } else {
      time = tc;
      for (int f = 1; f <= 4; ++f) {
        orc_update_phonon_field(m, f, time, u + off, acc ? acc + off / 2 : NULL);
        orc_left_evolve(m, time, (double *)m->g, m->Ns, f, 1);
        off += 2 * (size_t)npla;
      }
    }
    if (tc % m->ngroup == 0) orc_get_g_scratch(m, forward ? time + 1 : time, forward, NULL);
    if (meas_out && time == Lt / 2) {
      orc_measure(m, time, meas_out);
      if (n_meas) (*n_meas)++;
    }
  }
}

/* ------------------------------------------------------------------ */
/* ran2 (src/mod_ranstate.f90:171-191, 267-283)                         */
/* ------------------------------------------------------------------ */

void orc_ran2_seed(orc_ran2 *s, int32_t seed) {
  int32_t v[5] = {seed, seed, seed, seed, seed};
  for (int i = 0; i < 3; ++i)
    if (v[i] < 0) v[i] = ~v[i];
  for (int i = 3; i < 5; ++i)
    if (v[i] == 0) v[i] = 1;
  s->i = v[0]; s->j = v[1]; s->k = v[2]; s->m = v[3]; s->n = v[4];
}

void orc_ran2_fill(orc_ran2 *s, int64_t n, float *out) {
  /* amm = nearest(1.0_sp,-1.0_sp)/(-2^31), exact in REAL(SP) */
  const float amm = -0x1.fffffep-32f;
  for (int64_t q = 0; q < n; ++q) {
    int32_t r = (int32_t)((uint32_t)s->i - (uint32_t)s->k);
    if (r < 0) r = (int32_t)((uint32_t)r + 2147483579u);
    s->i = s->j;
    s->j = s->k;
    s->k = r;
    uint32_t nn = (uint32_t)s->n;
    nn ^= nn << 13;
    nn ^= nn >> 17;
    nn ^= nn << 5;
    s->n = (int32_t)nn;
    uint32_t mm = (uint32_t)s->m;
    uint32_t lo = mm & 65535u;
    mm = ((3533u * (mm >> 16) + lo) << 16) + 3533u * lo + 820265819u;
    s->m = (int32_t)mm;
    int32_t v = (int32_t)((nn ^ (uint32_t)s->k) + mm);
    int32_t sel = (v < 0) ? v : ~v;
    out[q] = amm * (float)sel;
  }
}

/* ------------------------------------------------------------------ */
/* global moves (src/mod_update.f90:251-467, src/mod_evolution.f90:359-408) */
/* ------------------------------------------------------------------ */

/* zlq of src/mod_matrixlib.f90:488-503 (ZGELQF + ZUNGLQ): a(d x n) = L Q,
 * a <- Q, l (d x d) the lower factor.  ZGELQF reflects the conjugated rows
 * with zlarfg, which is zgeqr2 on a^H: L = R^H and Q = Q'^H.  ZGELQ2
 * conjugates each row back after zlarfg, the diagonal included, so the
 * diagonal is conj((beta, +0)) = (beta, -0): a negative entry has
 * log = log|beta| - i pi in the reference. */
static void zlq(int d, int n, zc *a, zc *l) {
  zc *at = xcalloc((size_t)n * d, sizeof(zc));
  zc *tau = xcalloc((size_t)d, sizeof(zc));
  conj_transpose(d, n, a, at);
  zgeqr2(n, d, at, n, tau);
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i)
      l[i + (size_t)d * j] = (i >= j) ? conj(at[j + (size_t)n * i]) : 0;
  zung2r(n, d, at, n, tau);
  conj_transpose(n, d, at, a);
  free(tau);
  free(at);
}

/* cal_new_weight_T0 (src/mod_evolution.f90:359-408): the full-determinant
 * weight of the current fields, from the trial strings at both ends. */
/* cal_new_weight_FT (src/mod_evolution.f90:743-785): ln det(1 + B(beta,0))
 * of the current expKV, from a right string built from the identity over all
 * slices and the identity left string of block nblock.  The reference scales
 * the rows of Q by D where its comment says Q' = Q*D (:771-773); this
 * restates Q*D, the same step as get_g_scratch.  The strings and G are left
 * as they were. */
static void ft_new_weight(orc_model *m, double *ri) {
  size_t nq = (size_t)m->Ns * m->Ns * m->nblock, nd = (size_t)m->Ns * m->nblock, ng = (size_t)m->Ns * m->Ns;
  zc *q = xcalloc(nq, sizeof(zc)), *t = xcalloc(nq, sizeof(zc)), *d = xcalloc(nd, sizeof(zc));
  zc *g = xcalloc(ng, sizeof(zc));
  memcpy(q, m->FQ, nq * sizeof(zc));
  memcpy(t, m->FT, nq * sizeof(zc));
  memcpy(d, m->FD, nd * sizeof(zc));
  memcpy(g, m->g, ng * sizeof(zc));
  ft_init(m, 1);
  ri[0] = creal(m->ln_cw);
  ri[1] = cimag(m->ln_cw);
  memcpy(m->FQ, q, nq * sizeof(zc));
  memcpy(m->FT, t, nq * sizeof(zc));
  memcpy(m->FD, d, nd * sizeof(zc));
  memcpy(m->g, g, ng * sizeof(zc));
  free(q);
  free(t);
  free(d);
  free(g);
}

void orc_cal_new_weight(orc_model *m, double *ri) {
  if (m->ft) {
    ft_new_weight(m, ri);
    return;
  }
  int ne = m->ne, Ns = m->Ns, ng = m->ngroup;
  zc *lmat = xcalloc((size_t)ne * Ns, sizeof(zc));
  zc *tmat = xcalloc((size_t)ne * ne, sizeof(zc));
  conj_transpose(Ns, ne, m->slQ, lmat); /* Q_string(:, :, nblock) */
  zc lndet = 0;
  for (int i = 0; i < ne; ++i) lndet = lndet + clog(m->slD[i]);
  int flag = 0;
  for (int p = m->Lt; p >= 1; --p) {
    for (int f = 4; f >= 1; --f) orc_right_evolve(m, p, (double *)lmat, ne, f, 0);
    flag = flag + 1;
    if (flag != ng) continue;
    flag = 0;
    zlq(ne, Ns, lmat, tmat);
    for (int i = 0; i < ne; ++i) lndet = lndet + clog(tmat[i + (size_t)ne * i]);
  }
  for (int i = 0; i < ne; ++i) lndet = lndet + clog(m->slD[i]); /* D_string(:, 1) */
  gemm_nn(ne, ne, Ns, lmat, m->slQ, tmat);                      /* rmat = Q_string(:, :, 1) */
  lndet = lndet + clog(zdet(ne, tmat));
  if (m->tr_weight) lndet = lndet + conj(lndet);
  m->ln_cw = ln_branch(0 + lndet * (double)m->ncopy);
  ri[0] = creal(m->ln_cw);
  ri[1] = cimag(m->ln_cw);
  free(tmat);
  free(lmat);
}

/* ------------------------------------------------------------------ */
/* field gradient of ln_cw (no counterpart in the reference)           */
/* ------------------------------------------------------------------ */

/* d/dx of expkv(.., inv = 0), same layout.  With n = delt sqrt(|K|^2 +
 * (|V| x)^2), s = sinh(n)/n and n dn/dx = delt^2 |V|^2 x =: nd:
 * d cosh(n)/dx = s nd and ds/dx = r nd, r = (n cosh n - sinh n)/n^3, taken
 * from its series 1/3 + n^2/30 below n = 1e-3; dh/dx = -delt V_coe V12. */
static void dexpkv(const orc_model *m, int f, int b, double xv, zc *de) {
  double ak = cabs(m->kcoe[f]), av = cabs(m->vcoe[f]);
  double n = m->dtau * sqrt(ak * ak + (av * xv) * (av * xv));
  double nd = m->dtau * m->dtau * av * av * xv;
  double s = n > 0 ? sinh(n) / n : 1.0;
  double r = n < 1e-3 ? 1.0 / 3.0 + n * n / 30.0 : (n * cosh(n) - sinh(n)) / (n * n * n);
  zc sub[4], dsub[4];
  zc vx = m->vcoe[f] * xv;
  for (int q = 0; q < 4; ++q) {
    sub[q] = -m->dtau * (m->kcoe[f] * m->k12[f][q] + vx * m->v12[f][q]);
    dsub[q] = -m->dtau * m->vcoe[f] * m->v12[f][q];
  }
  if (m->cross[f * m->npla + b]) {
    const zc *ph = &m->bcph[(f * m->npla + b) * 16];
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c) {
        sub[a + 2 * c] = sub[a + 2 * c] * ph[a + 4 * (2 + c)];
        dsub[a + 2 * c] = dsub[a + 2 * c] * ph[a + 4 * (2 + c)];
      }
  }
  for (int q = 0; q < 16; ++q) de[q] = 0;
  for (int i = 0; i < 4; ++i) de[i + 4 * i] = s * nd;
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) {
      zc v = dsub[a + 2 * c] * s + sub[a + 2 * c] * (r * nd);
      de[a + 4 * (2 + c)] = v;
      de[(2 + c) + 4 * a] = conj(v);
    }
}

void orc_get_dexpkv(const orc_model *m, int pf_id, int pla, double xv, double *out) {
  dexpkv(m, pf_id - 1, pla, xv, (zc *)out);
}

/* d ln_cw / dx(i, t) for every field, complex, grad[i + Nb (t-1)].  A
 * backward pass over the slices without updates: where E_f(t) is the
 * outermost factor of G's right string (just before the backward sweep's
 * two-way wrap E^-1 G E), d ln det(L E R') / dx = Tr[G dE E^-1], and for a
 * bond only its four sites contribute: sum_ij (de e^-1)(i,j) G(S_j, S_i).
 * The same holds for g_h = 1 - (1 + B)^-1 and ln det(1 + B) at finite
 * temperature.  ln_cw's tr_weight doubling and ncopy power are applied.
 * Starts where a backward sweep starts (after init_g(1) or a forward sweep)
 * and leaves the chain where a backward sweep that rejects every proposal
 * leaves it. */
void orc_gradient(orc_model *m, double *grad_out) {
  int Lt = m->Lt, npla = m->npla, Ns = m->Ns;
  size_t Nb = m->Nb;
  zc *grad = (zc *)grad_out;
  for (size_t i = 0; i < Nb * Lt; ++i) grad[i] = 0;
  zc de[16], md[16];
  for (int tc = 1; tc <= Lt; ++tc) {
    int time = Lt - tc + 1;
    for (int f = 4; f >= 1; --f) {
      for (int b = 0; b < npla; ++b) {
        int fb = (f - 1) * npla + b;
        size_t idx = (size_t)m->bfi[fb] + Nb * (time - 1);
        dexpkv(m, f - 1, b, m->x[idx], de);
        const zc *ei = EKV(m, m->ekvi, f - 1, b, time);
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            zc s = 0;
            for (int k = 0; k < 4; ++k) s = s + de[i + 4 * k] * ei[k + 4 * j];
            md[i + 4 * j] = s;
          }
        const int *S = &m->sites[fb * 4];
        zc tr = 0;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) tr = tr + md[i + 4 * j] * m->g[S[j] + (size_t)Ns * S[i]];
        if (m->tr_weight) tr = tr + conj(tr);
        grad[idx] = grad[idx] + tr * (double)m->ncopy;
      }
      orc_right_evolve(m, time, (double *)m->g, Ns, f, 1);
    }
    if (tc % m->ngroup == 0) orc_get_g_scratch(m, time, 0, NULL);
  }
}

static float g_rands(orc_ran2 *s) {
  float r;
  orc_ran2_fill(s, 1, &r);
  return r;
}
/* irands (src/mod_ranstate.f90:221-226): int(r * n) in REAL(SP) */
static int g_irands(orc_ran2 *s, int n) { return (int)(g_rands(s) * (float)n); }
static float g_ran_sysm(orc_ran2 *s) { return g_rands(s) * 2.0f - 1.0f; }

/* generate_newfield_space (src/mod_update.f90:441-467) with
 * K_vec = 2*PI*[0.5d0, 0.5d0] (PI in REAL(SP)) and phi = 0 */
static void newfield_space(orc_model *m, double gud, const double *incell, orc_ran2 *s, double *bj) {
  const double distance = 2 * gud / m->Ncell;
  const double amplitude = distance * (double)g_rands(s);
  const double K = (double)(2.0f * (float)SP_PI) * 0.5;
  for (int c = 0; c < m->Ncell; ++c) {
    int dp[2];
    cell_dpos(m, c, dp); /* rpos = dpos for lat_vec = identity */
    const double arg = (K * (double)dp[0] + K * (double)dp[1]) + 0.0;
    for (int i = 0; i < 2; ++i) bj[2 * c + i] = incell[i] * amplitude * cos(arg);
  }
}

/* updata_expKV_global(t0, t1): expKV of slices t0..t1 from the fields */
static void expkv_slices(orc_model *m, int t0, int t1) {
  for (int t = t0; t <= t1; ++t)
    for (int f = 0; f < 4; ++f)
      for (int b = 0; b < m->npla; ++b) {
        double xv = m->x[m->bfi[f * m->npla + b] + (size_t)m->Nb * (t - 1)];
        expkv(m, f, b, xv, 0, EKV(m, m->ekv, f, b, t));
        expkv(m, f, b, xv, 1, EKV(m, m->ekvi, f, b, t));
      }
}

/* generate_newfield_global (src/mod_update.f90:292-439) for the kinds the
 * sweep loop calls (src/Main_PQMC.f90:88-94).  Returns -delt*dE. */
static double newfield_global(orc_model *m, int kind, const orc_global *gp, orc_ran2 *s) {
  const int Nb = m->Nb, Lt = m->Lt;
  const double gud = gp->distance * m->char_length;
  double *old = xcalloc((size_t)Nb * Lt, sizeof(double));
  memcpy(old, m->x, sizeof(double) * Nb * Lt);
  double de = 0.0;
  if (kind == ORC_GU_KSPACE || kind == ORC_GU_KSPACE_TIME) {
    double incell[2], *bj = xcalloc((size_t)Nb, sizeof(double));
    for (int i = 0; i < 2; ++i) incell[i] = g_irands(s, 2) * 2.0 - 1.0;
    newfield_space(m, gud, incell, s, bj);
    int t0 = 1, t1 = Lt;
    if (kind == ORC_GU_KSPACE_TIME) {
      t0 = g_irands(s, Lt / 2) + 1;
      t1 = t0 + (int)(1.0 * m->M / (m->dtau * m->dtau));
      if (t1 < Lt) t1 = Lt;
      /* the reference indexes past ntime when start + M/dtau^2 > ntime;
       * the window is cut at ntime here */
      if (t1 > Lt) t1 = Lt;
    }
    for (int t = t0; t <= t1; ++t)
      for (int i = 0; i < Nb; ++i) m->x[i + (size_t)Nb * (t - 1)] = m->x[i + (size_t)Nb * (t - 1)] + bj[i];
    for (int t = t0; t <= t1; ++t)
      for (int i = 0; i < Nb; ++i) {
        const double xn = m->x[i + (size_t)Nb * (t - 1)], xo = old[i + (size_t)Nb * (t - 1)];
        de = de + 0.5 * m->D * (xn * xn - xo * xo);
        de = de + (-m->biased * (xn - xo));
      }
    if (kind == ORC_GU_KSPACE_TIME) {
      const double kin = 0.5 * (m->M / (m->dtau * m->dtau));
      for (int i = 0; i < Nb; ++i) {
#define XN(t) m->x[i + (size_t)Nb * ((t) - 1)]
#define XO(t) old[i + (size_t)Nb * ((t) - 1)]
        const int tb = modi(t0 - 1, Lt), ta = modi(t1 + 1, Lt);
        de = de + kin * ((XN(t0) - XN(tb)) * (XN(t0) - XN(tb)) - (XO(t0) - XO(tb)) * (XO(t0) - XO(tb)));
        de = de + kin * ((XN(t1) - XN(ta)) * (XN(t1) - XN(ta)) - (XO(t1) - XO(ta)) * (XO(t1) - XO(ta)));
#undef XN
#undef XO
      }
    }
    free(bj);
    expkv_slices(m, t0, t1);
  } else if (kind == ORC_GU_ROTATE) {
    const double ang = (double)g_ran_sysm(s) * 0.1 * (double)(float)SP_PI;
    double r11 = cos(ang), r12 = -sin(ang), r21 = sin(ang), r22 = cos(ang);
    if (g_ran_sysm(s) < 0.0f) {
      r12 = -r12;
      r22 = -r22;
    }
    for (int c = 0; c < m->Ncell; ++c)
      for (int t = 0; t < Lt; ++t) {
        double *x = &m->x[(size_t)Nb * t + 2 * c];
        const double a = x[0], b = x[1];
        x[0] = r11 * a + r12 * b;
        x[1] = r21 * a + r22 * b;
      }
    expkv_slices(m, 1, Lt);
  } else {
    abort();
  }
  free(old);
  return -de * m->dtau;
}

int orc_update_global(orc_model *m, int kind, const orc_global *gp, orc_ran2 *s) {
  const size_t nx = (size_t)m->Nb * m->Lt, ne = (size_t)16 * m->npla * 4 * m->Lt;
  double *x_old = xcalloc(nx, sizeof(double));
  zc *e_old = xcalloc(ne, sizeof(zc)), *ei_old = xcalloc(ne, sizeof(zc));
  memcpy(x_old, m->x, sizeof(double) * nx);
  memcpy(e_old, m->ekv, sizeof(zc) * ne);
  memcpy(ei_old, m->ekvi, sizeof(zc) * ne);
  const zc w_old = m->ln_cw;
  const double ph = newfield_global(m, kind, gp, s);
  double ri[2];
  orc_cal_new_weight(m, ri);
  const double prob = cal_prob(m->ln_cw + ph - w_old);
  const int accept = (double)g_rands(s) < prob;
  if (accept) {
    m->global_acc++;
  } else {
    memcpy(m->x, x_old, sizeof(double) * nx);
    memcpy(m->ekv, e_old, sizeof(zc) * ne);
    memcpy(m->ekvi, ei_old, sizeof(zc) * ne);
    m->ln_cw = w_old;
    m->global_rej++;
  }
  free(ei_old);
  free(e_old);
  free(x_old);
  return accept;
}

void orc_propose_global(orc_model *m, int kind, const orc_global *gp, orc_ran2 *s, double *x_out, double *ph_out) {
  const size_t nx = (size_t)m->Nb * m->Lt, ne = (size_t)16 * m->npla * 4 * m->Lt;
  double *x_old = xcalloc(nx, sizeof(double));
  zc *e_old = xcalloc(ne, sizeof(zc)), *ei_old = xcalloc(ne, sizeof(zc));
  memcpy(x_old, m->x, sizeof(double) * nx);
  memcpy(e_old, m->ekv, sizeof(zc) * ne);
  memcpy(ei_old, m->ekvi, sizeof(zc) * ne);
  *ph_out = newfield_global(m, kind, gp, s);
  memcpy(x_out, m->x, sizeof(double) * nx);
  memcpy(m->x, x_old, sizeof(double) * nx);
  memcpy(m->ekv, e_old, sizeof(zc) * ne);
  memcpy(m->ekvi, ei_old, sizeof(zc) * ne);
  free(ei_old);
  free(e_old);
  free(x_old);
}

void orc_get_global_counters(const orc_model *m, long *acc, long *rej) {
  *acc = m->global_acc;
  *rej = m->global_rej;
}

/* ------------------------------------------------------------------ */
/* exposed dense kernels                                               */
/* ------------------------------------------------------------------ */

void orc_zqdr(int m, int n, double *a, double *r, double *d) { zqdr(m, n, (zc *)a, (zc *)r, (zc *)d); }
void orc_zldq(int m, int n, double *a, double *l, double *d) { zldq(m, n, (zc *)a, (zc *)l, (zc *)d); }
void orc_zinverse(int n, double *a) { zinverse(n, (zc *)a); }
void orc_zdet(int n, const double *a, double *det) {
  zc d = zdet(n, (const zc *)a);
  det[0] = creal(d);
  det[1] = cimag(d);
}
