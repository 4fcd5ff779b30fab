// One family pass at one time slice: the local phonon-field updates of
// update_phonon_field (src/mod_update.f90:13-201), fused with the two-way
// wrap of the same family (right_evolve before it on a backward sweep,
// left_evolve after it on a forward sweep, src/mod_evolution.f90:236-328).
//
// The reference visits the family's bonds in order and, after each accepted
// bond i, applies the rank-4 update
//     G <- G + u_i v_i,   u_i = G[:, S_i] X_i,   v_i = (I - G)[S_i, :],
//     X_i = Delta_i (I + G[S_i, S_i] Delta_i)^-1.
// The bonds of one family act on disjoint sites, so the updates of the a
// accepted bonds of any stretch of bonds fold into one update of rank 4a:
//     G_after[:, q] = G[:, q] + G[:, S_A] B w[q],   w[q] = (I - G)[S_A, q],
// where G is taken before the stretch and the 4a x 4a matrix B = A T^-1
// collects the sequence:
//     A[:, m] = (sel_m + sum_{m'<m} A[:, m'] v_{m'}[S_m]) X_m   (block upper),
//     T[m][m'] = u_{m'}[S_m] for m' < m, identity blocks on the diagonal.
// T^-1 is block lower with identity diagonal blocks, so A and T^-1 share one
// 4a x 4a array: A in the upper part with the diagonal blocks, T^-1 in the
// strictly lower part.  The bonds are cut into decision blocks of nb
// consecutive bonds; one leader workgroup decides them and H helper
// workgroups own the columns of G:
//
//  * Leader, wave 0: decides the bonds of block k from an LDS copy
//    W = G[S_k, S_k] that it keeps current with the rank-4 updates (only the
//    rows and columns of the block's later bonds change).
//  * Leader, wave 1: extends A and T^-1 by each accept while wave 0 goes on,
//    and publishes them with the accepted indices.
//  * Leader, waves 2-5: fetch G[S_{k+1}, S_k], G[S_k, S_{k+1}] and
//    G[S_{k+1}, S_{k+1}] after block k-1 and, accept by accept, the factors
//    P = G[S_{k+1}, S_A] A and Q = T^-1 G[S_A, S_{k+1}]; after block k they
//    fold it into the next window,  G[S_{k+1}, S_{k+1}] -= P Q.  The helpers
//    therefore lag by one block.  Wave 7 prepares the next block's proposals.
//  * Helper h keeps the columns of bph bonds in LDS for the whole pass and
//    applies block k to them in one step, G[:, q] += G[:, S_A] (B w[q]) with
//    B = A T^-1 formed locally, and the pivot columns G[:, S_A] read from a
//    slot that the owners of block k's columns filled after step k-1.
//
// Accept test: the reference rejects when prob < u with prob =
// min(1, exp(min(10, w))), w = dS + ncopy*(1+TR)*log|det(I + G_SS Delta)|
// (cal_prob, :475-500).  As log u < 0 this is exactly w < log u, i.e.
// |det|^2 < exp(2 (log u - dS) / (ncopy*(1+TR))); the right-hand side does
// not depend on G and is prepared before the bond's turn.
#include <mutex>

#include "pqmc_common.h"

#include "k_pass.h"

struct ps_prop {
  zd D[16];  // Delta = E(old)^-1 E(new) - I, column-major
  ekv_t en;
  double thr, xnew;
  int fld, pad[3];
};

// Data exchanged between the workgroups of a pass (records, slot columns,
// updated exponentials) is written only with device-scope (sc1) stores,
// which go through to memory, and a publisher waits for its own stores to
// complete (vmcnt 0) before it raises a flag.  Agent-scope fences, which
// write back and invalidate the whole L2 of the XCD, are avoided on purpose.
// Readers use plain, L2-cached loads for records and slot columns: every
// such address is written exactly once per pass (one slot per step) and is
// read only after its flag, so no cache can hold an older copy of it from
// this pass, and the caches are invalidated at kernel start.  The H helpers
// that read one block's columns then share one copy per XCD L2 instead of
// all fetching the same lines from memory.  Flags and the exponentials,
// which are read before and after they change, use coherent loads.
static __device__ __forceinline__ zd ld_coh(const zd *p) {
  const double *d = reinterpret_cast<const double *>(p);
  return zmk(__hip_atomic_load(d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
             __hip_atomic_load(d + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
static __device__ __forceinline__ void st_coh(zd *p, zd v) {
  double *d = reinterpret_cast<double *>(p);
  __hip_atomic_store(d, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(d + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ int ld_coh_i(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void st_coh_i(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void ld_coh_ekv(const ekv_t *p, ekv_t &e) {
  const zd *z = reinterpret_cast<const zd *>(p);
  const zd c = ld_coh(z);
  e.c = c.x;
  e.pad = c.y;
#pragma unroll
  for (int q = 0; q < 4; ++q) e.m[q] = ld_coh(z + 1 + q);
}
static __device__ __forceinline__ void st_coh_ekv(ekv_t *p, const ekv_t &e) {
  zd *z = reinterpret_cast<zd *>(p);
  st_coh(z, zmk(e.c, e.pad));
#pragma unroll
  for (int q = 0; q < 4; ++q) st_coh(z + 1 + q, e.m[q]);
}
// this thread's stores are complete (and, being device-scope, visible)
static __device__ __forceinline__ void ps_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
}
static __device__ __forceinline__ void ps_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

static __device__ __forceinline__ unsigned long long ps_ld(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void ps_st(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Single-lane wait until *p >= v (and *q >= w when q is set); returns the
// last value read from p, or 0 when the wait failed.  Bounded by wall-clock
// time and by the abort word.
static __device__ unsigned long long ps_wait(const ps_args &A, unsigned long long *p, unsigned long long v,
                                             unsigned long long *q, unsigned long long w) {
  // one round trip per poll: the flag words are loaded together; the abort
  // word and the clock are looked at every 32 polls
  const unsigned long long t0 = wall_clock64();
  unsigned long long *q2 = q ? q : p;
  const unsigned long long w2 = q ? w : v;
  for (unsigned it = 1;; ++it) {
    const unsigned long long a = ps_ld(p), b = ps_ld(q2);
    if (a >= v && b >= w2) {
      ps_acquire();
      return a;
    }
    if ((it & 31) == 0) {
      if (ps_ld(&A.sync[2]) != 0) return 0;
      if (wall_clock64() - t0 > PS_WAIT_TICKS) {
        __hip_atomic_store(&A.sync[2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(&A.stats[2], 1ull);
        return 0;
      }
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// helpers that have finished the backward wrap at the start of the pass
static __device__ __forceinline__ unsigned long long *ps_wrapped(const ps_args &A) { return &A.sync[3]; }

// count of the owners of block kstep+1+mb that have published after step kstep
static __device__ __forceinline__ unsigned long long *ps_pub(const ps_args &A, int kstep, int mb) {
  return &A.sync[8 + 2 * (kstep + 1) + mb];
}
static __device__ __forceinline__ unsigned long long *ps_flag(const ps_args &A, int k) {
  return &A.sync[10 + 2 * A.nblk + k];
}
// helpers that own at least one bond of block blk
static __device__ __forceinline__ unsigned long long ps_owners(const ps_args &A, int blk) {
  const int b0 = blk * A.nb, b1 = min((blk + 1) * A.nb, A.md.npla) - 1;
  return (unsigned long long)(b1 / A.bph - b0 / A.bph + 1);
}
// column cl (of 2 PS_BS rows: block kstep+1, then block kstep+2) of block kstep+1+mb in the tile slot
static __device__ __forceinline__ zd *ps_tcol(const ps_args &A, int kstep, int mb, int cl) {
  return A.tslot + ((size_t)(((kstep + 1) * 2 + mb) * PS_BS + cl)) * (2 * PS_BS);
}
// column cl of block kstep+1+mb in the slot written after step kstep
static __device__ __forceinline__ zd *ps_scol(const ps_args &A, int kstep, int mb, int cl) {
  return A.slot + ((size_t)(((kstep + 1) * 2 + mb) * PS_BS + cl)) * A.lds;
}
static __device__ __forceinline__ int ps_nsites(const ps_args &A, int k) {
  return 4 * min(A.nb, A.md.npla - k * A.nb);
}

// acc + a * b with four FMAs (throughput form for long accumulations)
static __device__ __forceinline__ zd zfma(zd acc, zd a, zd b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}

// v where ok, else zero (a select on the components, not on addresses)
static __device__ __forceinline__ zd zsel(bool ok, zd v) { return zmk(ok ? v.x : 0.0, ok ? v.y : 0.0); }

static __device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum over the lanes l, l+4, l+8, l+12 of each 16-lane row
static __device__ __forceinline__ double row4_sum(double v) {
  v += dpp_move<0x124, 0xF>(v);  // row_ror:4
  v += dpp_move<0x128, 0xF>(v);  // row_ror:8
  return v;
}

// max over the 16 lanes of each 16-lane row
static __device__ __forceinline__ double row16_max(double v) {
  v = fmax(v, dpp_move<0xB1, 0xF>(v));   // quad_perm [1,0,3,2]
  v = fmax(v, dpp_move<0x4E, 0xF>(v));   // quad_perm [2,3,0,1]
  v = fmax(v, dpp_move<0x141, 0xF>(v));  // row_half_mirror
  v = fmax(v, dpp_move<0x140, 0xF>(v));  // row_mirror
  return v;
}

// value of lane l (uniform l) in every lane
static __device__ __forceinline__ double rdlane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
static __device__ __forceinline__ zd rdlane_z(zd v, int l) { return zmk(rdlane_d(v.x, l), rdlane_d(v.y, l)); }
static __device__ __forceinline__ zd shfl_z(zd v, int l) { return zmk(__shfl(v.x, l, 64), __shfl(v.y, l, 64)); }

// The adjugate gives M^-1 = C^T / det accurately only while the cofactor
// expansion does not cancel: its error grows like max|M|^4 / |det| (in units
// of the rounding), and after an ill-conditioned accept (a small ratio that
// the bosonic action still lets through) G briefly holds entries of 1e5 and
// more, where it would lose most digits.  Beyond this bound the inverse is
// taken by Gauss-Jordan elimination with partial pivoting instead, as the
// reference inverts R with LU (inverse, src/mod_matrixlib.f90:410-431).
#define PS_ADJ_LIMIT 1.0e4

// modi of src/mod_matrixlib.f90:66-77
static __device__ __forceinline__ int modi_d(int n, int p) {
  int r = n % p;
  if (n < 0) return p + r;
  if (r == 0) return p;
  return r;
}

// 3x3 determinant of M (column-major 4x4 in LDS) without row ri and column cj.
static __device__ __forceinline__ zd minor3(const zd *M, int ri, int cj) {
  const int r0 = ri == 0 ? 1 : 0, r1 = ri <= 1 ? 2 : 1, r2 = ri <= 2 ? 3 : 2;
  const int c0 = cj == 0 ? 1 : 0, c1 = cj <= 1 ? 2 : 1, c2 = cj <= 2 ? 3 : 2;
#define E_(r, c) M[(r) + 4 * (c)]
  zd t = zmul(E_(r0, c0), zsub(zmul(E_(r1, c1), E_(r2, c2)), zmul(E_(r1, c2), E_(r2, c1))));
  t = zsub(t, zmul(E_(r0, c1), zsub(zmul(E_(r1, c0), E_(r2, c2)), zmul(E_(r1, c2), E_(r2, c0)))));
  t = zadd(t, zmul(E_(r0, c2), zsub(zmul(E_(r1, c0), E_(r2, c1)), zmul(E_(r1, c1), E_(r2, c0)))));
#undef E_
  return t;
}

// Optional phase timers (wall clock), kept in registers and added to
// A.prof[blockIdx.x * 8 + phase] once at the end.
#define PROF_T0()                                                \
  unsigned long long _pt = A.prof ? wall_clock64() : 0ull;       \
  unsigned long long _pa[8] = {0, 0, 0, 0, 0, 0, 0, 0}
#define PROF_ACC(k)                                              \
  do {                                                           \
    if (A.prof) {                                                \
      const unsigned long long _now = wall_clock64();            \
      _pa[k] += _now - _pt;                                      \
      _pt = _now;                                                \
    }                                                            \
  } while (0)
#define PROF_FLUSH_ROW(row)                                                               \
  do {                                                                                    \
    if (A.prof && (threadIdx.x & 63) == 0)                                                \
      for (int _k = 0; _k < 8; ++_k) A.prof[(row) * 8 + _k] += _pa[_k];                   \
  } while (0)
// rows: 0 leader wave 0, 1..H helpers, H+1 leader wave 1, H+2 leader wave 2
#define PROF_FLUSH()                                                                      \
  do {                                                                                    \
    if (threadIdx.x == 0)                                                                 \
      PROF_FLUSH_ROW(blockIdx.x);                                                         \
    else if (blockIdx.x == 0 && threadIdx.x == 64)                                        \
      PROF_FLUSH_ROW(A.H + 1);                                                            \
    else if (blockIdx.x == 0 && threadIdx.x == 128)                                       \
      PROF_FLUSH_ROW(A.H + 2);                                                            \
  } while (0)

// Proposal of bond p (family order): jump, bosonic action change
// (cal_energy_difference, :203-228), threshold, Delta (calculate_Delta,
// :230-245) and the new compact exponential.
static __device__ void ps_propose(const ps_args &A, int p, ps_prop &out) {
  const model_t &md = A.md;
  const size_t tbase = (size_t)md.Nb * (A.t - 1);
  const int fld = A.fbfi[p];
  const double xold = A.x[fld + tbase];
  // ran_sysm = ran2*2-1 in REAL(SP) (src/mod_ranstate.f90:233-236); jump in DP
  const float rs = A.u[2 * p] * 2.0f - 1.0f;
  const double jump = md.jump * (double)rs;
  const double be = A.x[fld + (size_t)md.Nb * (modi_d(A.t - 1, md.Lt) - 1)];
  const double af = A.x[fld + (size_t)md.Nb * (modi_d(A.t + 1, md.Lt) - 1)];
  const double nw = xold + jump;
  const double dt2 = md.dtau * md.dtau;
  const double e1 = 0.5 * md.M * ((be - xold) * (be - xold) + (af - xold) * (af - xold)) / dt2 +
                    0.5 * md.D * (xold * xold) - md.biased * xold;
  const double e2 = 0.5 * md.M * ((be - nw) * (be - nw) + (af - nw) * (af - nw)) / dt2 +
                    0.5 * md.D * (nw * nw) - md.biased * nw;
  const double dS = -md.dtau * (e2 - e1);
  const double kpow = (double)md.ncopy * (md.tr_weight ? 2.0 : 1.0);
  out.thr = exp(2.0 * (log((double)A.u[2 * p + 1]) - dS) / kpow);
  out.xnew = nw;
  out.fld = fld;
  ekv_t en;
  expkv_compact(md, A.f, nw, A.fcross[p] ? &A.fph[4 * p] : nullptr, en);
  out.en = en;
  const ekv_t eo = A.ekv[((size_t)(A.t - 1) * 4 + A.f) * md.npla + p];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj)
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
      zd s = zmk(0.0, 0.0);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) s = zadd(s, zmul(ekv_entry(eo, true, ii, kk), ekv_entry(en, false, kk, jj)));
      if (ii == jj) s.x -= 1.0;
      out.D[ii + 4 * jj] = s;
    }
}

// The three tiles that fold block k into the window of block k+1, as they
// were after step k-1: T10 = G[S_{k+1}, S_k] stored transposed
// (T10t[c * PS_BS + r]), T01 = G[S_k, S_{k+1}] and T11 = G[S_{k+1}, S_{k+1}]
// row-major; entries outside the blocks are zero.  256 threads (t0 = 0..255)
// issue all their loads before the first LDS store.
static __device__ void ps_load_fold_tiles(const ps_args &A, int k, zd *T10t, zd *T01, zd *T11, int t0) {
  // thread t0 takes row r = t0 % 32 and columns c0 + 8j (j = 0..3) of each tile,
  // from the compact tile slot written after step k-1 (rows: block k, then k+1)
  const int n0 = ps_nsites(A, k), n1 = ps_nsites(A, k + 1);
  const int r = t0 & 31, c0 = t0 >> 5;
  zd v10[4], v01[4], v11[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = c0 + 8 * j;
    v10[j] = ps_tcol(A, k - 1, 0, c)[PS_BS + r];  // G[S_{k+1}(r), S_k(c)]
    v01[j] = ps_tcol(A, k - 1, 1, c)[r];          // G[S_k(r), S_{k+1}(c)]
    v11[j] = ps_tcol(A, k - 1, 1, c)[PS_BS + r];  // G[S_{k+1}(r), S_{k+1}(c)]
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = c0 + 8 * j;
    T10t[c * PS_BS + r] = zsel(r < n1 && c < n0, v10[j]);
    T01[r * PS_BS + c] = zsel(r < n0 && c < n1, v01[j]);
    T11[r * PS_BS + c] = zsel(r < n1 && c < n1, v11[j]);
  }
}

// --------------------------------------------------------------------------
// Leader: decisions (wave 0), B (wave 1), next block (waves 2-7), fold.
// --------------------------------------------------------------------------
struct ps_leader_lds {
  zd buf[4][PS_BS * PS_BS];     // W, T10, T01, T11 (W and T11 rotate)
  zd sU[PS_NB][PS_BS * 4];      // u_m at the block's rows (row*4 + comp)
  zd sV[PS_NB][4 * PS_BS];      // v_m at the block's cols (comp*BS + col)
  zd sX[PS_NB][16];
  zd sAT[PS_BS * PS_BS];        // A (upper part with diagonal blocks) and T^-1 (strictly lower), stride PS_BS
  zd sPt[PS_BS * PS_BS];        // P transposed: sPt[g * PS_BS + r] = P[r][g]
  zd sQ[PS_BS * PS_BS];         // Q (row-major, stride PS_BS)
  zd sTt[PS_BS * 4];
  zd sM[4][16], sCof[16];
  zd dets[PS_NB];
  ps_prop prop[2][PS_NB];
  int li[PS_NB];
  int npub, done;    // wave 0 -> wave 1: accepts handed over, block decided
  int nfold, bdone;  // wave 1 -> waves 2-5: accepts folded into A and T^-1, block folded
  int tiles;         // waves 2-5 that have stored their part of the fold tiles
  int wdone;         // wave 6 -> wave 0: accepts whose update of W off its diagonal 4x4 blocks is done
  int nacc, abort;
};

// One wave waits until the LDS count *cnt exceeds have, or until *fin is
// set; returns the count then seen (have or less means no more will come),
// or -1 after the clock bound.  The producer sets *fin after its last count.
static __device__ int ps_lds_progress(const ps_args &A, ps_leader_lds &L, int *cnt, int *fin, int have) {
  const unsigned long long t0 = wall_clock64();
  for (unsigned it = 1;; ++it) {
    const int np = __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (np > have) return np;
    if (__hip_atomic_load(fin, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP))
      return __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    // the producers never wait inside a block; the bound only guards against a logic error
    if ((it & 63) == 0 && wall_clock64() - t0 > PS_WAIT_TICKS) {
      if ((threadIdx.x & 63) == 0) {
        __hip_atomic_store(&A.sync[2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(&A.stats[2], 1ull);
        L.abort = 1;
      }
      return -1;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Barrier among the n waves that call it, through an LDS counter that is
// reset between uses (lane 0 of each wave counts and waits).
static __device__ void ps_lds_bar(const ps_args &A, ps_leader_lds &L, int *ctr, int n) {
  if ((threadIdx.x & 63) == 0) {
    __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const unsigned long long t0 = wall_clock64();
    for (unsigned it = 1; __hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < n; ++it) {
      if ((it & 63) == 0 && wall_clock64() - t0 > PS_WAIT_TICKS) {
        __hip_atomic_store(&A.sync[2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(&A.stats[2], 1ull);
        L.abort = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  wave_sync();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Wave 1: extend A and T^-1 by accept m of the current block.
static __device__ void ps_fold_accept(ps_leader_lds &L, int m, int lane) {
  const int n = 4 * (m + 1);
  const int jm = L.li[m];
  zd *__restrict__ sAT = L.sAT;
  zd *__restrict__ sTt = L.sTt;
  // T(g, kk) = sel + sum_{m'<m} A[g][4m'+d] v_{m'}[d][S_m(kk)]  (A[g][4m'+d] = 0 unless g/4 <= m')
  // T^-1[4m + r][g] = - sum_{m'<m} u_{m'}[S_m(r)][d] T^-1[4m'+d][g]  for g < 4m, where
  // T^-1[4m'+d][g] is the identity on the diagonal block m' = g/4 and stored below it
  for (int e = lane; e < 4 * n + 4 * 4 * m; e += 64) {
    if (e < 4 * n) {
      const int g = e >> 2, kk = e & 3;
      zd t = zmk(g == 4 * m + kk ? 1.0 : 0.0, 0.0);
      for (int mp = g >> 2; mp < m; ++mp) {
        zd av[4], vv[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          av[d] = sAT[g * PS_BS + 4 * mp + d];
          vv[d] = L.sV[mp][d * PS_BS + 4 * jm + kk];
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) t = zadd(t, zmul(av[d], vv[d]));
      }
      sTt[g * 4 + kk] = t;
    } else {
      const int e2 = e - 4 * n, nl = 4 * m;
      const int r = e2 / nl, g = e2 % nl;
      const int gb = g >> 2;
      // diagonal block m' = g/4: -u_{m'}[S_m(r)][g % 4]
      zd t = zneg(L.sU[gb][(4 * jm + r) * 4 + (g & 3)]);
      for (int mp = gb + 1; mp < m; ++mp) {
        zd uv[4], tv[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          uv[d] = L.sU[mp][(4 * jm + r) * 4 + d];
          tv[d] = sAT[(4 * mp + d) * PS_BS + g];
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) t = zsub(t, zmul(uv[d], tv[d]));
      }
      sAT[(4 * m + r) * PS_BS + g] = t;
    }
  }
  wave_sync();
  // A[:, block m] = T X_m
  for (int e = lane; e < 4 * n; e += 64) {
    const int g = e >> 2, c = e & 3;
    zd tv[4], xv[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      tv[kk] = sTt[g * 4 + kk];
      xv[kk] = L.sX[m][kk + 4 * c];
    }
    zd s = zmk(0.0, 0.0);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) s = zadd(s, zmul(tv[kk], xv[kk]));
    sAT[g * PS_BS + 4 * m + c] = s;
  }
  wave_sync();
}

// Waves 2-5 (t2 = 0..255): the contributions of accept m to
//     P[:, block m] = G[S_{k+1}, S_A] A[:, block m]      (t2 < 128: row t2/4, column t2%4)
//     Q[block m, :] = T^-1[block m, :] G[S_A, S_{k+1}]   (t2 >= 128: row (t2-128)/32, column t2%32)
// with T10 stored transposed.  Entries are independent, so no barrier is needed.
static __device__ void ps_pq_accept(ps_leader_lds &L, int m, int t2, const zd *T10t, const zd *T01) {
  if (t2 < 128) {
    const int r = t2 >> 2, c = t2 & 3;
    zd s0 = zmk(0.0, 0.0), s1 = s0;
    for (int mb = 0; mb <= m; ++mb) {
      const int cb = 4 * L.li[mb];
      zd tv[4], av[4];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        tv[d] = T10t[(cb + d) * PS_BS + r];
        av[d] = L.sAT[(4 * mb + d) * PS_BS + 4 * m + c];
      }
      s0 = zadd(s0, zadd(zmul(tv[0], av[0]), zmul(tv[1], av[1])));
      s1 = zadd(s1, zadd(zmul(tv[2], av[2]), zmul(tv[3], av[3])));
    }
    L.sPt[(4 * m + c) * PS_BS + r] = zadd(s0, s1);
  } else {
    const int e = t2 - 128, rr = e >> 5, c = e & 31;
    zd s0 = T01[(4 * L.li[m] + rr) * PS_BS + c], s1 = zmk(0.0, 0.0);
    for (int mb = 0; mb < m; ++mb) {
      const int cb = 4 * L.li[mb];
      zd tv[4], ov[4];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        tv[d] = L.sAT[(4 * m + rr) * PS_BS + 4 * mb + d];
        ov[d] = T01[(cb + d) * PS_BS + c];
      }
      s0 = zadd(s0, zadd(zmul(tv[0], ov[0]), zmul(tv[1], ov[1])));
      s1 = zadd(s1, zadd(zmul(tv[2], ov[2]), zmul(tv[3], ov[3])));
    }
    L.sQ[(4 * m + rr) * PS_BS + c] = zadd(s0, s1);
  }
}

// Wave 6: W += u_m v_m on the later rows and columns of accept m, except the
// diagonal 4 x 4 blocks of the later bonds (the deciding wave updates those),
// on FP64 MFMA: per 16 x 16 tile of W that meets them, D += U V with U the
// u_m rows (zero above row o) and V the v_m columns (zero left of column o),
// complex as four real products of K = 4.  Lane l supplies U[l % 16][l / 16]
// and V[l / 16][l % 16] and holds rows l / 16 + 4 i of column l % 16.  Every
// entry of W has one writer, in accept order, so the result does not depend
// on how the two waves interleave.
static __device__ void ps_w_offdiag(ps_leader_lds &L, zd *W, int m, int nsk) {
  const int lane = threadIdx.x & 63;
  const int o = 4 * (L.li[m] + 1);
  if (o >= nsk) return;
  const int li = lane & 15, lk = lane >> 4;
  const int t0 = o >> 4, t1 = (nsk - 1) >> 4;
  for (int rt = t0; rt <= t1; ++rt) {
    const int ra = 16 * rt + li;
    const zd ua = zsel(ra >= o && ra < nsk, L.sU[m][ra * 4 + lk]);
    for (int ct = t0; ct <= t1; ++ct) {
      const int cb = 16 * ct + li;
      const zd vb = zsel(cb >= o && cb < nsk, L.sV[m][lk * PS_BS + cb]);
      double4_t dre, dim;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const zd t = W[(16 * rt + lk + 4 * i) * PS_BS + cb];
        dre[i] = t.x;
        dim[i] = t.y;
      }
      dre = mfma_f64(ua.x, vb.x, dre);
      dim = mfma_f64(ua.x, vb.y, dim);
      dre = mfma_f64(-ua.y, vb.y, dre);
      dim = mfma_f64(ua.y, vb.x, dim);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * rt + lk + 4 * i;
        if (row >= o && row < nsk && cb >= o && cb < nsk && (row >> 2) != (cb >> 2))
          W[row * PS_BS + cb] = zmk(dre[i], dim[i]);
      }
    }
  }
  wave_sync();
}

static __device__ void ps_leader(const ps_args &A, ps_leader_lds &L) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const model_t &md = A.md;
  const int npla = md.npla, nb = A.nb, nblk = A.nblk;
  const size_t tbase = (size_t)md.Nb * (A.t - 1);
  PROF_T0();
  int iW = 0, i10 = 1, i01 = 2, i11 = 3;
  bool ok = true;
  // a backward pass wraps every column with the exponentials before the
  // update; none may change until all helpers have read them
  bool ekv_free = A.wrap != 2;
  // proposals of block 0; the window after the helpers' load
  if (tid < min(nb, npla)) ps_propose(A, tid, L.prop[0][tid]);
  if (tid == 0) {
    L.abort = 0;
    L.npub = 0;
    L.done = 0;
    L.nfold = 0;
    L.bdone = 0;
    L.tiles = 0;
    L.wdone = 0;
    L.nacc = ps_wait(A, ps_pub(A, -1, 0), ps_owners(A, 0), nullptr, 0) ? 1 : 0;
  }
  __syncthreads();
  ok = L.nacc != 0;
  if (ok) {
    const int n0 = ps_nsites(A, 0);
    for (int e = tid; e < PS_BS * PS_BS; e += PS_THREADS) {
      const int c = e >> 5, r = e & 31;
      L.buf[iW][r * PS_BS + c] = zsel(r < n0 && c < n0, ps_tcol(A, -1, 0, c)[r]);
    }
  }
  __syncthreads();
  PROF_ACC(0);
  int total_acc = 0, total_rej = 0;
  for (int k = 0; k < nblk && ok; ++k) {
    const int p0 = k * nb;
    const int nsk = ps_nsites(A, k);
    const int nbk = nsk / 4;
    const bool has_next = k + 1 < nblk;
    if (wave == 0) {
      // ---- decisions of block k ----
      zd *W = L.buf[iW];
      const ps_prop *pp = L.prop[k & 1];
      int m = 0;
      for (int jb = 0; jb < nbk;) {
        // One round evaluates the next four bonds from the same window, one
        // per 16-lane row (bond jb + gq).  Bonds before the first accepted one
        // are rejected exactly as in the sequential scan, since a rejection
        // leaves G unchanged; the round then continues after that bond.
        // M = I + G_SS Delta (src/mod_update.f90:97-100), lane a + 4c of a row holds M(a, c)
        const int gq = lane >> 4, ll = lane & 15, ra = ll & 3, rc = ll >> 2;
        const int nv = min(4, nbk - jb);
        const int jq = jb + min(gq, nv - 1);
        {
          zd wv[4], dv[4];
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            wv[kk] = W[(4 * jq + ra) * PS_BS + 4 * jq + kk];
            dv[kk] = pp[jq].D[kk + 4 * rc];
          }
          zd mv = zmk(ra == rc ? 1.0 : 0.0, 0.0);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) mv = zadd(mv, zmul(wv[kk], dv[kk]));
          L.sM[gq][ll] = mv;
        }
        wave_sync();
        // cofactor C(ra, rc); det = sum_c M(0, c) C(0, c) over lanes 0, 4, 8, 12 of the row
        zd cof = minor3(L.sM[gq], ra, rc);
        if ((ra + rc) & 1) cof = zneg(cof);
        const zd term = ra == 0 ? zmul(L.sM[gq][ll], cof) : zmk(0.0, 0.0);
        const double tr = row4_sum(term.x), ti = row4_sum(term.y);
        const bool accq = gq < nv && !(tr * tr + ti * ti < pp[jq].thr);
        const unsigned long long bal = __ballot(ll == 0 && accq);
        const int f = bal ? (int)(__builtin_ctzll(bal) >> 4) : nv;
        if (ll == 0 && gq < f) A.acc[p0 + jb + gq] = 0;
        total_rej += f;
        PROF_ACC(4);
        if (f == nv) {
          jb += nv;
          continue;
        }
        const int j = jb + f;
        jb = j + 1;
        const double dre = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(tr), 16 * f),
                                            __builtin_amdgcn_readlane(__double2loint(tr), 16 * f));
        const double dim = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(ti), 16 * f),
                                            __builtin_amdgcn_readlane(__double2loint(ti), 16 * f));
        const double n2 = dre * dre + dim * dim;
        if (lane == 0) A.acc[p0 + j] = 1;
        if (gq == f) L.sCof[ll] = cof;
        // X = Delta M^-1 with M^-1(kk, c) = C(c, kk) / det
        wave_sync();
        {
          const zd me = L.sM[gq][ll];
          const double nrm = rdlane_d(row16_max(fabs(me.x) + fabs(me.y)), 16 * f);
          if (nrm * nrm * nrm * nrm > PS_ADJ_LIMIT * sqrt(n2)) {
            // Gauss-Jordan on [M | I] with partial pivoting (largest |re| + |im|,
            // the first of equals, as izamax in zgetrf); lane l < 32 holds entry
            // (l & 3, l >> 2) of the 4 x 8 array, lanes 32..63 idle along
            const int gr = lane & 3, gc = lane >> 2;
            zd a = zmk(0.0, 0.0);
            if (lane < 32) a = gc < 4 ? L.sM[f][gr + 4 * gc] : zmk(gr == gc - 4 ? 1.0 : 0.0, 0.0);
            for (int k = 0; k < 4; ++k) {
              int pr = k;
              double best = -1.0;
              for (int i = k; i < 4; ++i) {
                const zd v = rdlane_z(a, i + 4 * k);
                const double sv = fabs(v.x) + fabs(v.y);
                if (sv > best) {
                  best = sv;
                  pr = i;
                }
              }
              a = shfl_z(a, (gr == k ? pr : (gr == pr ? k : gr)) + 4 * gc);
              const zd piv = rdlane_z(a, 5 * k);
              const double pn = piv.x * piv.x + piv.y * piv.y;
              if (gr == k) a = zmul(a, zmk(piv.x / pn, -piv.y / pn));
              const zd rk = shfl_z(a, k + 4 * gc);  // row k, scaled
              const zd fk = shfl_z(a, gr + 4 * k);  // this row at column k
              if (gr != k) a = zsub(a, zmul(fk, rk));
            }
            // det M^-1(kk, c) in the place of the cofactor C(c, kk)
            if (lane < 32 && gc >= 4) L.sCof[(gc - 4) + 4 * gr] = zmul(zmk(dre, dim), a);
            wave_sync();
          }
        }
        if (lane < 16) {
          zd dv[4], cv[4];
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            dv[kk] = pp[j].D[ra + 4 * kk];
            cv[kk] = L.sCof[rc + 4 * kk];
          }
          const double in2 = 1.0 / n2;
          const zd idet = zmk(dre * in2, -dim * in2);
          zd xv = zmk(0.0, 0.0);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) xv = zadd(xv, zmul(dv[kk], cv[kk]));
          L.sX[m][lane] = zmul(xv, idet);
        }
        if (lane == 0) {
          const ekv_t en = pp[j].en;
          const double xn = pp[j].xnew;
          const int fld = pp[j].fld;
          L.li[m] = j;
          L.dets[m] = zmk(dre, dim);
          A.x[fld + tbase] = xn;
          if (!ekv_free) ekv_free = ps_wait(A, ps_wrapped(A), (unsigned long long)(gridDim.x - 1), nullptr, 0) != 0;
          st_coh_ekv(&A.ekv[((size_t)(A.t - 1) * 4 + A.f) * npla + p0 + j], en);
        }
        wave_sync();
        // u_m and v_m read W off its diagonal blocks: wave 6 has applied the
        // earlier accepts there
        if (m > 0 && L.wdone < m) ps_lds_progress(A, L, &L.wdone, &L.abort, m - 1);
        PROF_ACC(5);
        // u_m, v_m at the later sites of the block
        const int o = 4 * (j + 1), Ln = nsk - o;
        for (int e = lane; e < Ln * 4; e += 64) {
          const int r = o + (e >> 2), c = e & 3;
          zd s = zmk(0.0, 0.0);
#pragma unroll
          for (int a = 0; a < 4; ++a) s = zadd(s, zmul(W[r * PS_BS + 4 * j + a], L.sX[m][a + 4 * c]));
          L.sU[m][r * 4 + c] = s;
          const int ca = e / Ln, cc = o + e % Ln;
          L.sV[m][ca * PS_BS + cc] = zneg(W[(4 * j + ca) * PS_BS + cc]);
        }
        wave_sync();
        // hand accept m to wave 1
        if (lane == 0) __hip_atomic_store(&L.npub, m + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        PROF_ACC(6);
        // W += u_m v_m on the diagonal 4 x 4 blocks of the later bonds, the
        // only part of W the decision rounds read; wave 6 updates the rest
        // (ps_w_offdiag) while the next round runs
        for (int e = lane; e < Ln * 4; e += 64) {
          const int row = o + (e >> 2), col = 4 * (row >> 2) + (e & 3);
          zd uv[4], vv[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            uv[c] = L.sU[m][row * 4 + c];
            vv[c] = L.sV[m][c * PS_BS + col];
          }
          zd acc = W[row * PS_BS + col];
#pragma unroll
          for (int c = 0; c < 4; ++c) acc = zadd(acc, zmul(uv[c], vv[c]));
          W[row * PS_BS + col] = acc;
        }
        wave_sync();
        PROF_ACC(7);
        ++m;
        ++total_acc;
      }
      if (lane == 0) {
        L.nacc = m;
        __hip_atomic_store(&L.done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      ps_release();  // acc flags, fields and exponentials are complete before block k is published
      PROF_ACC(0);
    } else if (wave == 1) {
      // ---- A and T^-1 of block k, accept by accept ----
      int md_ = 0;
      for (;;) {
        PROF_ACC(0);
        const int np = ps_lds_progress(A, L, &L.npub, &L.done, md_);
        PROF_ACC(0);
        if (np <= md_) break;
        for (; md_ < np; ++md_) {
          ps_fold_accept(L, md_, lane);
          if (lane == 0) __hip_atomic_store(&L.nfold, md_ + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
          // the record's entries that accept md_ fixed: T^-1 rows 4 md_..4 md_+3
          // left of the diagonal block, A columns 4 md_..4 md_+3 down to row
          // 4 md_+3; later accepts leave them unchanged
          ps_rec *R = A.rec + k;
          const int m4 = 4 * md_;
          for (int e = lane; e < 4 * m4 + 4 * (m4 + 4); e += 64) {
            const int r = e < 4 * m4 ? m4 + e / m4 : (e - 4 * m4) >> 2;
            const int c = e < 4 * m4 ? e % m4 : m4 + ((e - 4 * m4) & 3);
            st_coh(&R->AT[r * PS_BS + c], L.sAT[r * PS_BS + c]);
          }
        }
        PROF_ACC(1);
      }
      if (lane == 0) __hip_atomic_store(&L.bdone, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      // the record's accepted indices now (its A and T^-1 went out accept by
      // accept); its flag after the barrier
      const int a = L.nacc;
      ps_rec *R = A.rec + k;
      if (lane < a) st_coh_i(&R->li[lane], L.li[lane]);
      PROF_ACC(4);
    } else if (has_next && wave <= 5) {
      // ---- waves 2..5: the fold tiles after step k-1, then P and Q ----
      bool wok = true;
      if (lane == 0)
        wok = ps_wait(A, ps_pub(A, k - 1, 0), ps_owners(A, k), ps_pub(A, k - 1, 1), ps_owners(A, k + 1)) != 0;
      wok = __shfl(wok ? 1 : 0, 0, 64) != 0;
      PROF_ACC(1);
      if (wok) {
        ps_load_fold_tiles(A, k, L.buf[i10], L.buf[i01], L.buf[i11], tid - 128);
        ps_lds_bar(A, L, &L.tiles, 4);  // P and Q read all of the tiles
        PROF_ACC(7);
        int md_ = 0;
        for (;;) {
          const int np = ps_lds_progress(A, L, &L.nfold, &L.bdone, md_);
          if (np <= md_) break;
          for (; md_ < np; ++md_) ps_pq_accept(L, md_, tid - 128, L.buf[i10], L.buf[i01]);
        }
      } else if (lane == 0) {
        L.abort = 1;
      }
      PROF_ACC(5);
    } else if (wave == 6) {
      // ---- W off its diagonal blocks, accept by accept ----
      int md_ = 0;
      for (;;) {
        const int np = ps_lds_progress(A, L, &L.npub, &L.done, md_);
        if (np <= md_) break;
        for (; md_ < np; ++md_) {
          ps_w_offdiag(L, L.buf[iW], md_, nsk);
          if (lane == 0) __hip_atomic_store(&L.wdone, md_ + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    } else if (has_next && wave == 7) {
      // ---- wave 7: the next block's proposals ----
      const int p1 = (k + 1) * nb, nb1 = ps_nsites(A, k + 1) / 4;
      if (lane < nb1) ps_propose(A, p1 + lane, L.prop[(k + 1) & 1][lane]);
    }
    if (tid == 0 && ps_ld(&A.sync[2]) != 0) L.abort = 1;
    __syncthreads();
    if (L.abort) { ok = false; break; }
    PROF_ACC(2);
    const int a = L.nacc;
    if (wave == 1) {
      // ---- publish block k: the record stores are complete, then the flag ----
      ps_release();
      if (lane == 0) ps_st(ps_flag(A, k), (unsigned long long)(a + 1));
      // ln_cw += log(real(R_elec)) per accept (src/mod_update.f90:142)
      double lv = 0.0;
      if (lane < a) {
        zd rel = L.dets[lane];
        if (md.tr_weight) rel = zmul(rel, zconj(rel));
        zd rp = rel;
        for (int c = 1; c < md.ncopy; ++c) rp = zmul(rp, rel);
        lv = log(rp.x);
      }
      lv = wave_sum(lv);
      if (lane == 0 && a > 0) atomicAdd(A.lncw_acc, lv);
      PROF_ACC(4);
    }
    if (!has_next) break;
    zd *T11 = L.buf[i11];
    if (a > 0 && wave >= 2 && wave <= 5) {
      // ---- fold block k into the next window: T11 -= P Q, K = 4a ----
      // on FP64 MFMA, one 16 x 16 tile per wave; lane l supplies A[l%16][l/16]
      // and B[l/16][l%16], and holds rows l/16 + 4i of column l%16 of the result.
      const int tt = wave - 2;
      const int rt = tt >> 1, ct = tt & 1;
      const int lr = lane & 15, lk = lane >> 4;
      double4_t dre0, dim0, dre1 = {0.0, 0.0, 0.0, 0.0}, dim1 = dre1;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const zd t = T11[(16 * rt + lk + 4 * i) * PS_BS + 16 * ct + lr];
        dre0[i] = t.x;
        dim0[i] = t.y;
      }
      for (int ks = 0; ks < a; ks += 2) {
        const int k2 = min(ks + 1, a - 1);
        const zd a0 = L.sPt[(4 * ks + lk) * PS_BS + 16 * rt + lr];
        const zd b0 = L.sQ[(4 * ks + lk) * PS_BS + 16 * ct + lr];
        const zd a1 = zsel(ks + 1 < a, L.sPt[(4 * k2 + lk) * PS_BS + 16 * rt + lr]);
        const zd b1 = L.sQ[(4 * k2 + lk) * PS_BS + 16 * ct + lr];
        dre0 = mfma_f64(-a0.x, b0.x, dre0);
        dim0 = mfma_f64(-a0.x, b0.y, dim0);
        dre1 = mfma_f64(-a1.x, b1.x, dre1);
        dim1 = mfma_f64(-a1.x, b1.y, dim1);
        dre0 = mfma_f64(a0.y, b0.y, dre0);
        dim0 = mfma_f64(-a0.y, b0.x, dim0);
        dre1 = mfma_f64(a1.y, b1.y, dre1);
        dim1 = mfma_f64(-a1.y, b1.x, dim1);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        T11[(16 * rt + lk + 4 * i) * PS_BS + 16 * ct + lr] = zmk(dre0[i] + dre1[i], dim0[i] + dim1[i]);
    }
    if (wave == 0)
      PROF_ACC(1);
    else
      PROF_ACC(6);
    if (tid == 0) {
      L.npub = 0;
      L.done = 0;
      L.nfold = 0;
      L.bdone = 0;
      L.tiles = 0;
      L.wdone = 0;
    }
    __syncthreads();
    // the window of block k+1 is T11; the old window's buffer becomes T11
    const int tmp = iW;
    iW = i11;
    i11 = tmp;
    PROF_ACC(3);
  }
  if (tid == 0) {
    atomicAdd(&A.stats[0], (unsigned long long)total_acc);
    atomicAdd(&A.stats[1], (unsigned long long)total_rej);
  }
  PROF_FLUSH();
}

// --------------------------------------------------------------------------
// Helpers: own columns, wraps, block steps.
// --------------------------------------------------------------------------
struct ps_helper_hdr {
  zd sAT[PS_BS * PS_BS];       // A and T^-1 of the block (packed as in the record, stride PS_BS)
  zd sB[PS_BS * PS_BS];        // v = T^-1 w per own column (stride PS_BS)
  int li[PS_NB];
  int a, ok, sink;
};
// after the header: sw, sy [4 bph][PS_BS], sC [4 bph][Ns], then the family's
// sites [4 npla] (int)

// Own columns: row r of local column q at sC[q * Ns + r].
static __device__ void ps_wrap_cols(const ps_args &A, zd *sC, int b0, int b1, bool forward) {
  const int Ns = A.md.Ns, npla = A.md.npla, tid = threadIdx.x;
  const ekv_t *ek = A.ekv + ((size_t)(A.t - 1) * 4 + A.f) * npla;
  const int nbo = b1 - b0;
  // column mixing of the own bonds: backward G E, forward G E^-1
  for (int e = tid; e < nbo * Ns; e += PS_THREADS) {
    const int bl = e / Ns, r = e % Ns;
    ekv_t E;
    ld_coh_ekv(ek + b0 + bl, E);
    zd v[4], y[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = sC[(4 * bl + c) * Ns + r];
    ekv_apply_right(E, forward, v, y);
#pragma unroll
    for (int c = 0; c < 4; ++c) sC[(4 * bl + c) * Ns + r] = y[c];
  }
  __syncthreads();
  // row mixing over every bond of the family: backward E^-1 G, forward E G
  const int ncol = 4 * nbo;
  for (int e = tid; e < npla * ncol; e += PS_THREADS) {
    const int b = e / ncol, q = e % ncol;
    ekv_t E;
    ld_coh_ekv(ek + b, E);
    const int *S = A.fsites + 4 * b;
    zd v[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) v[a] = sC[q * Ns + S[a]];
    ekv_apply(E, !forward, v, y);
#pragma unroll
    for (int a = 0; a < 4; ++a) sC[q * Ns + S[a]] = y[a];
  }
  __syncthreads();
}

// After step kstep the slot holds the columns of blocks kstep+1, kstep+2 and
// the tile slot the same columns at the rows of those two blocks.
static __device__ void ps_publish_cols(const ps_args &A, const zd *sC, const int *sS, int b0, int b1, int kstep) {
  const int Ns = A.md.Ns, tid = threadIdx.x;
  for (int b = b0; b < b1; ++b) {
    const int blk = b / A.nb;
    const int mb = blk - (kstep + 1);
    if (mb < 0 || mb > 1) continue;
    const int cl0 = 4 * (b - blk * A.nb);
    const zd *col = sC + 4 * (b - b0) * Ns;
    for (int e = tid; e < 4 * Ns; e += PS_THREADS) {
      const int c = e / Ns, r = e % Ns;
      st_coh(ps_scol(A, kstep, mb, cl0 + c) + r, col[c * Ns + r]);
    }
    for (int e = tid; e < 4 * 2 * PS_BS; e += PS_THREADS) {
      const int c = e >> 6, i = e & 63, j = kstep + 1 + (i >> 5), r = i & 31;
      if (j < A.nblk && r < ps_nsites(A, j)) st_coh(ps_tcol(A, kstep, mb, cl0 + c) + i, col[c * Ns + sS[4 * A.nb * j + r]]);
    }
  }
}

// One thread: count this helper among the publishers of the blocks
// kstep+1, kstep+2 that it owns bonds of (after its stores are complete).
static __device__ void ps_announce(const ps_args &A, int b0, int b1, int kstep) {
  bool own[2] = {false, false};
  for (int b = b0; b < b1; ++b) {
    const int mb = b / A.nb - (kstep + 1);
    if (mb == 0 || mb == 1) own[mb] = true;
  }
  for (int mb = 0; mb < 2; ++mb)
    if (own[mb]) atomicAdd(ps_pub(A, kstep, mb), 1ull);
}

// G[r, q] += sum_g G_k[r, S_A(g)] y[q][g] for the own columns, four at a
// time.  All pivot loads of a row are issued before the first use (entries
// past 4a read one fixed word, a single request per wave, and are not used);
// y is staged in registers four pivots at a time.
static __device__ void ps_pivot_update(const ps_args &A, const ps_helper_hdr &Hh, int kstep, int n4, int ncol,
                                       const zd *sy, zd *sC) {
  const int Ns = A.md.Ns;
  constexpr int HB = PS_BS / 2;
  int col[PS_NB];  // first slot column of each accepted bond
#pragma unroll
  for (int m = 0; m < PS_NB; ++m) col[m] = 4 * Hh.li[m];
  for (int r = threadIdx.x; r < Ns; r += PS_THREADS) {
    for (int q0 = 0; q0 < ncol; q0 += 4) {
      zd acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = zmk(0.0, 0.0);
#pragma unroll
      for (int hb = 0; hb < 2; ++hb) {
        const int g0 = hb * HB;
        if (g0 >= n4) break;
        zd P[HB];
#pragma unroll
        for (int i = 0; i < HB; ++i) {
          const int g = g0 + i;
          const zd *p = g < n4 ? ps_scol(A, kstep, 0, col[g >> 2] + (g & 3)) + r : A.slot;
          P[i] = *p;
        }
#pragma unroll
        for (int gg = 0; gg < HB; gg += 4) {
          if (g0 + gg < n4) {
            zd Y[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int q = 0; q < 4; ++q) Y[i][q] = sy[(q0 + q) * PS_BS + g0 + gg + i];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int q = 0; q < 4; ++q) acc[q] = zfma(acc[q], P[gg + i], Y[i][q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) sC[(q0 + q) * Ns + r] = zadd(sC[(q0 + q) * Ns + r], acc[q]);
    }
  }
}

// The same update on FP64 MFMA when a helper owns 16 columns (four bonds):
// per 16-row tile, C(16 x 16) += P(16 x n4) Y(n4 x 16) with P the pivot
// columns G_k[:, S_A] and Y[g][q] = y[q][g], complex as four real products.
// Lane l supplies P[t0 + l%16][g0 + l/16] and Y[g0 + l/16][l%16] and holds C
// rows l/16 + 4i of column l%16.  All pivot loads of a tile are issued
// before the first product.
static __device__ void ps_pivot_update_mfma(const ps_args &A, const ps_helper_hdr &Hh, int kstep, int n4,
                                            const zd *sy, zd *sC) {
  const int Ns = A.md.Ns, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  constexpr int NWV = PS_THREADS / 64, NCH = PS_BS / 4, NT = 4;  // NT tiles per wave cover Ns <= 512
  zd p[NT][NCH];
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int r = 16 * (wv + NWV * j) + li;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      const int g = 4 * ch + lk;
      p[j][ch] = (4 * ch < n4 && r < Ns) ? ps_scol(A, kstep, 0, 4 * Hh.li[g >> 2] + (g & 3))[r] : zmk(0.0, 0.0);
    }
  }
  zd y[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) y[ch] = 4 * ch < n4 ? sy[li * PS_BS + 4 * ch + lk] : zmk(0.0, 0.0);
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int t0 = 16 * (wv + NWV * j);
    if (t0 >= Ns) break;  // uniform
    double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      if (4 * ch < n4) {  // uniform
        cr = mfma_f64(p[j][ch].x, y[ch].x, cr);
        cr = mfma_f64(-p[j][ch].y, y[ch].y, cr);
        ci = mfma_f64(p[j][ch].x, y[ch].y, ci);
        ci = mfma_f64(p[j][ch].y, y[ch].x, ci);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = t0 + lk + 4 * i;
      if (row < Ns) sC[li * Ns + row] = zadd(sC[li * Ns + row], zmk(cr[i], ci[i]));
    }
  }
}

template <int BPH>
static __device__ void ps_helper(const ps_args &A, ps_helper_hdr &Hh, zd *sC, zd *sw, zd *sy, int *sS) {
  constexpr int NC = 4 * BPH;
  const int tid = threadIdx.x;
  const model_t &md = A.md;
  const int Ns = md.Ns, npla = md.npla;
  const int h = blockIdx.x - 1;
  const int b0 = h * BPH, b1 = min(npla, b0 + BPH);
  const int ncol = 4 * (b1 - b0);
  PROF_T0();
  // ---- load own columns (natural column = site), optional backward wrap ----
  for (int e = tid; e < ncol * Ns; e += PS_THREADS) {
    const int q = e / Ns, r = e % Ns;
    sC[e] = A.G[r + (size_t)Ns * A.fsites[4 * b0 + q]];
  }
  for (int e = ncol * Ns + tid; e < NC * Ns; e += PS_THREADS) sC[e] = zmk(0.0, 0.0);
  for (int e = tid; e < 4 * npla; e += PS_THREADS) sS[e] = A.fsites[e];
  __syncthreads();
  if (A.wrap == 2) ps_wrap_cols(A, sC, b0, b1, false);
  ps_publish_cols(A, sC, sS, b0, b1, -1);
  ps_release();
  __syncthreads();
  if (tid == 0) {
    ps_announce(A, b0, b1, -1);
    // this workgroup has read the exponentials its backward wrap needs
    atomicAdd(ps_wrapped(A), 1ull);
  }
  PROF_ACC(0);
  for (int k = 0; k < A.nblk; ++k) {
    // Block k's columns are in the slot once their owners have published
    // after step k-1; the leader's decision of block k comes later.
    // Meanwhile the helpers of an XCD (workgroups are spread over the 8 XCDs
    // round-robin) touch those columns in slices, so that the pivot loads of
    // step k hit in L2.
    if (tid == 0) Hh.ok = ps_wait(A, ps_pub(A, k - 1, 0), ps_owners(A, k), nullptr, 0) != 0;
    __syncthreads();
    if (!Hh.ok) break;
    {
      // one load per 128-byte line (8 entries of a column): line ln goes to
      // the helper whose slice is ln % 12
      const int nsk = ps_nsites(A, k), slice = (blockIdx.x >> 3) % 12, nl = (Ns + 7) >> 3;
      double t = 0.0;
      for (int e = tid; e < nsk * nl; e += PS_THREADS) {
        const int c = e / nl, ln = e % nl;
        if (ln % 12 == slice) t += ps_scol(A, k - 1, 0, c)[8 * ln].x;
      }
      if (t == 1.2345e300) Hh.sink = 1;  // keeps the loads; never true for Green's function entries
    }
    if (tid == 0) {
      const unsigned long long fl = ps_wait(A, ps_flag(A, k), 1ull, nullptr, 0);
      Hh.ok = fl != 0;
      Hh.a = fl != 0 ? (int)fl - 1 : 0;
    }
    __syncthreads();
    if (!Hh.ok) break;
    PROF_ACC(1);
    const int a = Hh.a, n4 = 4 * a;
    if (a > 0) {
      const ps_rec *R = A.rec + k;
      if (tid < a) Hh.li[tid] = R->li[tid];
      for (int e = tid; e < n4 * n4; e += PS_THREADS) {
        const int r = e / n4, c = e % n4;
        Hh.sAT[r * PS_BS + c] = R->AT[r * PS_BS + c];
      }
      __syncthreads();
      PROF_ACC(2);
      // w[q][g] = e_q(S_g) - G_k[S_g, q], then y[q] = B w[q] = A (T^-1 w[q]) as
      // two block-triangular products (no n4 x n4 product B = A T^-1):
      //   v[q][g] = w[q][g] + sum_{h/4 < g/4} T^-1[g][h] w[q][h]
      //   y[q][g] = sum_{h/4 >= g/4} A[g][h] v[q][h]
      const int *Sk = sS + 4 * A.nb * k;
      for (int e = tid; e < ncol * n4; e += PS_THREADS) {
        const int q = e / n4, g = e % n4;
        const int st = Sk[4 * Hh.li[g >> 2] + (g & 3)];
        sw[q * PS_BS + g] = zsub(zmk(sS[4 * b0 + q] == st ? 1.0 : 0.0, 0.0), sC[q * Ns + st]);
      }
      __syncthreads();
      zd *sv = Hh.sB;  // v[q][g] at sv[q * PS_BS + g]
      for (int e = tid; e < ncol * n4; e += PS_THREADS) {
        const int q = e / n4, g = e % n4;
        const int hend = 4 * (g >> 2);
        zd v0 = sw[q * PS_BS + g], v1 = zmk(0.0, 0.0);
        for (int hh = 0; hh < hend; hh += 4) {
          zd tv[4], wv[4];
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            tv[d] = Hh.sAT[g * PS_BS + hh + d];
            wv[d] = sw[q * PS_BS + hh + d];
          }
          v0 = zfma(zfma(v0, tv[0], wv[0]), tv[1], wv[1]);
          v1 = zfma(zfma(v1, tv[2], wv[2]), tv[3], wv[3]);
        }
        sv[q * PS_BS + g] = zadd(v0, v1);
      }
      __syncthreads();
      for (int e = tid; e < ncol * n4; e += PS_THREADS) {
        const int q = e / n4, g = e % n4;
        zd y0 = zmk(0.0, 0.0), y1 = y0;
        for (int hh = 4 * (g >> 2); hh < n4; hh += 4) {
          zd av[4], vv[4];
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            av[d] = Hh.sAT[g * PS_BS + hh + d];
            vv[d] = sv[q * PS_BS + hh + d];
          }
          y0 = zfma(zfma(y0, av[0], vv[0]), av[1], vv[1]);
          y1 = zfma(zfma(y1, av[2], vv[2]), av[3], vv[3]);
        }
        sy[q * PS_BS + g] = zadd(y0, y1);
      }
      __syncthreads();
      PROF_ACC(3);
      // G[:, q] += G_k[:, S_A] y[q]; pivot columns from the slot of step k-1
      if constexpr (BPH == 4)
        ps_pivot_update_mfma(A, Hh, k - 1, n4, sy, sC);
      else
        ps_pivot_update(A, Hh, k - 1, n4, NC, sy, sC);
      __syncthreads();
    }
    PROF_ACC(5);
    ps_publish_cols(A, sC, sS, b0, b1, k);
    ps_release();
    __syncthreads();
    PROF_ACC(6);
    if (tid == 0) ps_announce(A, b0, b1, k);
    PROF_ACC(7);
  }
  if (A.wrap == 1) ps_wrap_cols(A, sC, b0, b1, true);
  for (int e = tid; e < ncol * Ns; e += PS_THREADS) {
    const int q = e / Ns, r = e % Ns;
    A.G[r + (size_t)Ns * A.fsites[4 * b0 + q]] = sC[e];
  }
  PROF_ACC(4);
  PROF_FLUSH();
}

__global__ __launch_bounds__(PS_THREADS) void k_family_pass(ps_args A) {
  extern __shared__ __align__(16) char ps_lds[];
  __shared__ int s_skip;
  // after any expired wait the chain is broken: later passes do nothing until
  // the host has seen the error
  if (threadIdx.x == 0) s_skip = __hip_atomic_load(&A.stats[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
  __syncthreads();
  if (s_skip) {
  } else if (blockIdx.x == 0) {
    ps_leader(A, *reinterpret_cast<ps_leader_lds *>(ps_lds));
  } else {
    ps_helper_hdr &Hh = *reinterpret_cast<ps_helper_hdr *>(ps_lds);
    zd *sw = reinterpret_cast<zd *>(ps_lds + sizeof(ps_helper_hdr));
    zd *sy = sw + 4 * A.bph * PS_BS;
    zd *sC = sy + 4 * A.bph * PS_BS;
    int *sS = reinterpret_cast<int *>(sC + 4 * A.bph * A.md.Ns);
    if (A.bph == 1)
      ps_helper<1>(A, Hh, sC, sw, sy, sS);
    else if (A.bph == 2)
      ps_helper<2>(A, Hh, sC, sw, sy, sS);
    else
      ps_helper<4>(A, Hh, sC, sw, sy, sS);
  }
  // the last workgroup to finish clears the sync words for the next launch
  __syncthreads();
  if (threadIdx.x == 0) {
    ps_release();
    const unsigned long long done = atomicAdd(&A.sync[1], 1ull);
    if (done == gridDim.x - 1) {
      ps_st(&A.sync[2], 0ull);
      ps_st(ps_wrapped(A), 0ull);
      for (int k = -1; k < A.nblk; ++k) {
        ps_st(ps_pub(A, k, 0), 0ull);
        ps_st(ps_pub(A, k, 1), 0ull);
      }
      for (int k = 0; k < A.nblk; ++k) ps_st(ps_flag(A, k), 0ull);
      ps_st(&A.sync[1], 0ull);
    }
  }
}

extern "C" size_t pqmc_pass_lds(int Ns, int bph) {
  const size_t lead = sizeof(ps_leader_lds);
  const size_t help =
      sizeof(ps_helper_hdr) + sizeof(zd) * (size_t)4 * bph * (2 * PS_BS + Ns) + sizeof(int) * (size_t)2 * Ns;
  return lead > help ? lead : help;
}
extern "C" size_t pqmc_pass_rec_bytes(void) { return sizeof(ps_rec); }

extern "C" hipError_t pqmc_launch_pass(const ps_args *a, hipStream_t s) {
  if (a->bph != 1 && a->bph != 2 && a->bph != 4) return hipErrorInvalidValue;
  if (a->bph == 4 && a->md.Ns > 512) return hipErrorInvalidValue;  // ps_pivot_update_mfma covers 512 rows
  const size_t lds = pqmc_pass_lds(a->md.Ns, a->bph);
  static std::mutex attr_mu;  // launches may come from several host threads
  static size_t attr = 0;
  {
    std::lock_guard<std::mutex> lk(attr_mu);
    if (lds > attr) {
      hipError_t e = hipFuncSetAttribute((const void *)k_family_pass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = lds;
    }
  }
  hipLaunchKernelGGL(k_family_pass, dim3(a->H + 1), dim3(PS_THREADS), lds, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t pqmc_pass_max_groups(int Ns, int bph, int *out) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_family_pass, PS_THREADS,
                                                              pqmc_pass_lds(Ns, bph));
  if (e != hipSuccess) return e;
  int dev = 0, ncu = 0;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *out = per_cu * ncu;
  return hipSuccess;
}
