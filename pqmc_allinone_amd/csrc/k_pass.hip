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
// accepted bonds of any stretch of bonds fold into one update of the
// stretch's start G:
//     G_after[:, q] = G[:, q] + sum_m G[:, S_m] y_m(q),
//     v_m(q) = w_m(q) - sum_{m'<m} T[m][m'] v_m'(q),   w_m(q) = (I - G)[S_m, q],
//     y_m(q) = X_m (v_m(q) + sum_{m'>m} V[m][m'] y_m'(q))   (m = a-1 .. 0),
// with the 4 x 4 blocks T[m][m'] = u_m'[S_m] and V[m'][m] = v_m'[S_m]
// (forward and backward block substitutions).  T holds X as the reference's
// u does (G[:, S] X is formed first, src/mod_update.f90:177); in y the sum
// is formed before X_m multiplies it, which keeps the drift after an
// ill-conditioned accept (large X_m) at the level of the sequential
// updates.  The bonds are cut into
// decision blocks of nb consecutive bonds; one leader workgroup decides them
// and H helper workgroups own the columns of G:
//
//  * Leader, wave 0: decides the bonds of block k from an LDS copy
//    W = G[S_k, S_k] that it keeps current with the rank-4 updates (only the
//    rows and columns of the block's later bonds change); wave 6 takes the
//    part of each update off W's diagonal 4 x 4 blocks.
//  * Leader, wave 1: copies each accept's X_m, T and V blocks (from wave
//    0's u_m, v_m at the block's sites) to the block's record and publishes
//    it as soon as the block is decided.
//  * Leader, waves 2-5: fetch G[S_{k+1}, S_k], G[S_k, S_{k+1}] and
//    G[S_{k+1}, S_{k+1}] as they were after block k-1 into MFMA
//    accumulators and apply block k's accepts to them one by one, so that
//    the window of block k+1 is ready when block k is decided.  The helpers
//    therefore lag by one block.  Wave 7 prepares the next block's proposals,
//    and in the last block those of the next pass's block 0 (ps_args.nxt).
//  * Pass start: the leader gathers block 0's window from G itself (with the
//    backward wrap restricted to the block) and takes block 0's proposals
//    from the previous pass (ps_args.pre) when it formed them.
//  * Helper h keeps the columns of bph bonds in LDS for the whole pass and
//    applies block k to them in one step from the record, with the pivot
//    columns G[:, S_A] from a slot that the owners of block k's columns
//    filled after step k-1 (with one bond per helper, loaded into registers
//    while the leader decides the block).

// Accept test: the reference rejects when prob < u with prob =
// min(1, exp(min(10, w))), w = dS + ncopy*(1+TR)*log|det(I + G_SS Delta)|
// (cal_prob, :475-500).  As log u < 0 this is exactly w < log u, i.e.
// |det|^2 < exp(2 (log u - dS) / (ncopy*(1+TR))); the right-hand side does
// not depend on G and is prepared before the bond's turn.
#include <mutex>

#include "pqmc_common.h"

#include "k_pass.h"

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
static __device__ __forceinline__ unsigned long long ps_wait(const ps_args &A, unsigned long long *p, unsigned long long v,
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
// count of the owners of block kstep+1 whose slot columns are out after step kstep
static __device__ __forceinline__ unsigned long long *ps_spub(const ps_args &A, int kstep) {
  return &A.sync[10 + 3 * A.nblk + kstep + 1];
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

// sum over the four lanes of each quad, in every lane of the quad
static __device__ __forceinline__ double quad_sum(double v) {
  v += dpp_move<0xB1, 0xF>(v);  // quad_perm [1,0,3,2]
  v += dpp_move<0x4E, 0xF>(v);  // quad_perm [2,3,0,1]
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
// rows: 0 leader wave 0, 1..H helpers, H+1 leader wave 1, H+2 leader wave 2,
// H+3 leader wave 6, H+4 leader wave 7
#define PROF_FLUSH()                                                                      \
  do {                                                                                    \
    if (threadIdx.x == 0)                                                                 \
      PROF_FLUSH_ROW(blockIdx.x);                                                         \
    else if (blockIdx.x == 0 && threadIdx.x == 64)                                        \
      PROF_FLUSH_ROW(A.H + 1);                                                            \
    else if (blockIdx.x == 0 && threadIdx.x == 128)                                       \
      PROF_FLUSH_ROW(A.H + 2);                                                            \
    else if (blockIdx.x == 0 && threadIdx.x == 384)                                       \
      PROF_FLUSH_ROW(A.H + 3);                                                            \
    else if (blockIdx.x == 0 && threadIdx.x == 448)                                       \
      PROF_FLUSH_ROW(A.H + 4);                                                            \
  } while (0)

// Proposal of bond p (family order): jump, bosonic action change
// (cal_energy_difference, :203-228), threshold, Delta (calculate_Delta,
// :230-245) and the new compact exponential.
// Family f at slice t with its draws u and tables (A.f, A.t, A.u, ... for
// this pass).
static __device__ __forceinline__ void ps_propose(const ps_args &A, int f, int t, const float *u, const int *fbfi,
                                                  const int *fcross, const zd *fph, int p, ps_prop &out) {
  const model_t &md = A.md;
  const size_t tbase = (size_t)md.Nb * (t - 1);
  const int fld = fbfi[p];
  const double xold = A.x[fld + tbase];
  // ran_sysm = ran2*2-1 in REAL(SP) (src/mod_ranstate.f90:233-236); jump in DP
  const float rs = u[2 * p] * 2.0f - 1.0f;
  const double jump = md.jump * (double)rs;
  const double be = A.x[fld + (size_t)md.Nb * (modi_d(t - 1, md.Lt) - 1)];
  const double af = A.x[fld + (size_t)md.Nb * (modi_d(t + 1, md.Lt) - 1)];
  const double nw = xold + jump;
  const double dt2 = md.dtau * md.dtau;
  const double e1 = 0.5 * md.M * ((be - xold) * (be - xold) + (af - xold) * (af - xold)) / dt2 +
                    0.5 * md.D * (xold * xold) - md.biased * xold;
  const double e2 = 0.5 * md.M * ((be - nw) * (be - nw) + (af - nw) * (af - nw)) / dt2 +
                    0.5 * md.D * (nw * nw) - md.biased * nw;
  const double dS = -md.dtau * (e2 - e1);
  const double kpow = (double)md.ncopy * (md.tr_weight ? 2.0 : 1.0);
  out.thr = exp(2.0 * (log((double)u[2 * p + 1]) - dS) / kpow);
  out.xnew = nw;
  out.fld = fld;
  ekv_t en;
  expkv_compact(md, f, nw, fcross[p] ? &fph[4 * p] : nullptr, en);
  out.en = en;
  const ekv_t eo = A.ekv[((size_t)(t - 1) * 4 + f) * md.npla + p];
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

// --------------------------------------------------------------------------
// Leader: decisions (wave 0), record (wave 1), next window (waves 2-5),
// W off its diagonal blocks (wave 6), next proposals (wave 7).
// --------------------------------------------------------------------------
struct ps_leader_lds {
  zd buf[2][PS_BS * PS_BS];     // the window W of block k and that of block k+1 (they rotate)
  zd sU[PS_NB][PS_BS * 4];      // u_m at the block's rows (row*4 + comp)
  zd sV[PS_NB][4 * PS_BS];      // v_m at the block's cols (comp*BS + col)
  zd sX[PS_NB][16];
  zd nC[2][PS_BS * 4];          // T10[R, S_j] of the accept being folded (row*4 + a), by accept parity
  zd nR[2][4 * PS_BS];          // T01[S_j, C] (a*BS + col), by accept parity
  zd sM[4][16], sCof[16];
  zd dets[PS_NB];
  ps_prop prop[2][PS_NB];
  int li[PS_NB];
  int npub, done;    // wave 0 -> wave 1: accepts handed over, block decided
  int fbar;          // waves 2-5: barrier count of the accept-by-accept fold
  int wdone;         // wave 6 -> wave 0: accepts whose update of W off its diagonal 4x4 blocks is done
  int nacc, abort;
};

// One wave waits until the LDS count *cnt exceeds have, or until *fin is
// set; returns the count then seen (have or less means no more will come),
// or -1 after the clock bound.  The producer sets *fin after its last count.
static __device__ __forceinline__ int ps_lds_progress(const ps_args &A, ps_leader_lds &L, int *cnt, int *fin, int have) {
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

// Barrier among the n waves that call it, through a monotone LDS counter:
// the i-th use by each wave waits for the count n * i.
static __device__ __forceinline__ void ps_lds_bar_to(const ps_args &A, ps_leader_lds &L, int *ctr, int target) {
  if ((threadIdx.x & 63) == 0) {
    __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const unsigned long long t0 = wall_clock64();
    for (unsigned it = 1; __hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < target; ++it) {
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

// Waves 2-5 fold the accepts of block k, one by one, into the tiles of block
// k+1, as the reference's rank-4 update restricted to them,
// G <- G + u_m v_m with u_m = G[:, S_j] X_m and v_m = (I - G)[S_j, :]
// (accept m at bond j, o = 4 (j + 1), R and C the sites of block k+1):
//   u_m[R] = T10[R, S_j] X_m,  v_m[C] = -T01[S_j, C],
//   T11 += u_m[R] v_m[C],  T10[R, c] += u_m[R] v_m[c],  T01[r, C] += u_m[r] v_m[C]
// for the block's later columns c and rows r (>= o), whose u_m[r] and v_m[c]
// wave 0 has formed from W.  Only the later bonds' columns of T10 and rows of
// T01 are read again.  Wave w (0..3) keeps the 16 x 16 tile (w / 2, w % 2) of
// T11, T10 and T01 in MFMA accumulators for the whole block (lane l: rows
// l/16 + 4i, column l%16); per accept the owners of S_j's column of T10 and
// row of T01 put them in LDS, and each product is three real FP64 MFMAs of
// K = 4 (zmfma).  Entries outside the blocks are zero and stay zero.
struct ps_tiles {
  double4_t r11, i11, r10, i10, r01, i01;
};

static __device__ __forceinline__ void ps_tiles_load(const ps_args &A, int k, ps_tiles &T) {
  const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) - 2, rt = w >> 1, ct = w & 1;
  const int lr = lane & 15, lk = lane >> 4, col = 16 * ct + lr;
  const int n0 = ps_nsites(A, k), n1 = ps_nsites(A, k + 1);
  const zd *c0 = ps_tcol(A, k - 1, 0, col), *c1 = ps_tcol(A, k - 1, 1, col);
  zd v11[4], v10[4], v01[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = 16 * rt + lk + 4 * i;
    v11[i] = c1[PS_BS + row];  // G[S_{k+1}(row), S_{k+1}(col)]
    v10[i] = c0[PS_BS + row];  // G[S_{k+1}(row), S_k(col)]
    v01[i] = c1[row];          // G[S_k(row), S_{k+1}(col)]
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = 16 * rt + lk + 4 * i;
    const zd a = zsel(row < n1 && col < n1, v11[i]), b = zsel(row < n1 && col < n0, v10[i]),
             c = zsel(row < n0 && col < n1, v01[i]);
    T.r11[i] = a.x;
    T.i11[i] = a.y;
    T.r10[i] = b.x;
    T.i10[i] = b.y;
    T.r01[i] = c.x;
    T.i01[i] = c.y;
  }
}

// D += a b for complex operands on FP64 MFMA by three real products from
// zero, P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi): re += P1 - P2,
// im += P3 - P1 - P2 (one MFMA deep instead of two chained)
static __device__ __forceinline__ void zmfma(double4_t &dr, double4_t &di, zd a, zd b) {
  const double4_t z = {0.0, 0.0, 0.0, 0.0};
  const double4_t p1 = mfma_f64(a.x, b.x, z), p2 = mfma_f64(a.y, b.y, z), p3 = mfma_f64(a.x + a.y, b.x + b.y, z);
  dr += p1 - p2;
  di += (p3 - p1) - p2;
}

// Step 1 of accept m: the owners put T10[R, S_j] and T01[S_j, C] in LDS.
static __device__ __forceinline__ void ps_tiles_share(ps_leader_lds &L, int m, const ps_tiles &T) {
  const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) - 2, rt = w >> 1, ct = w & 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int j4 = 4 * L.li[m], par = m & 1;
  const int a = 16 * ct + lr - j4;
  if (a >= 0 && a < 4)
#pragma unroll
    for (int i = 0; i < 4; ++i) L.nC[par][(16 * rt + lk + 4 * i) * 4 + a] = zmk(T.r10[i], T.i10[i]);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int b = 16 * rt + lk + 4 * i - j4;
    if (b >= 0 && b < 4) L.nR[par][b * PS_BS + 16 * ct + lr] = zmk(T.r01[i], T.i01[i]);
  }
}

// Step 2 of accept m: the three rank-4 products on the wave's tiles.
static __device__ __forceinline__ void ps_tiles_fold(ps_leader_lds &L, int m, int nsk, ps_tiles &T) {
  const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) - 2, rt = w >> 1, ct = w & 1;
  const int lr = lane & 15, lk = lane >> 4;
  const int o = 4 * (L.li[m] + 1), par = m & 1;
  const int row = 16 * rt + lr, col = 16 * ct + lr;
  // A operands: u_m[R] at (row, lk) and, for T01, wave 0's u_m at the block's rows
  zd u = zmk(0.0, 0.0);
#pragma unroll
  for (int a = 0; a < 4; ++a) u = zadd(u, zmul(L.nC[par][row * 4 + a], L.sX[m][a + 4 * lk]));
  const zd ub = zsel(row >= o && row < nsk, L.sU[m][row * 4 + lk]);
  // B operands: v_m[C] at (lk, col) and, for T10, wave 0's v_m at the block's columns
  const zd v = zneg(L.nR[par][lk * PS_BS + col]);
  const zd vb = zsel(col >= o && col < nsk, L.sV[m][lk * PS_BS + col]);
  zmfma(T.r11, T.i11, u, v);
  zmfma(T.r10, T.i10, u, vb);
  zmfma(T.r01, T.i01, ub, v);
}

// The wave's tile of T11 (the next window) to LDS, row-major.
static __device__ __forceinline__ void ps_tiles_store(zd *W, const ps_tiles &T) {
  const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) - 2, rt = w >> 1, ct = w & 1;
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) W[(16 * rt + lk + 4 * i) * PS_BS + 16 * ct + lr] = zmk(T.r11[i], T.i11[i]);
}

// Wave 6: W += u_m v_m on the later rows and columns of accept m, except the
// diagonal 4 x 4 blocks of the later bonds (the deciding wave updates those),
// on FP64 MFMA: per 16 x 16 tile of W that meets them, D += U V with U the
// u_m rows (zero above row o) and V the v_m columns (zero left of column o),
// complex as three real products of K = 4 (zmfma).  Lane l supplies U[l % 16][l / 16]
// and V[l / 16][l % 16] and holds rows l / 16 + 4 i of column l % 16.  Every
// entry of W has one writer, in accept order, so the result does not depend
// on how the two waves interleave.
static __device__ __forceinline__ void ps_w_offdiag(ps_leader_lds &L, zd *W, int m, int nsk) {
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
      zmfma(dre, dim, ua, vb);
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

static __device__ __forceinline__ void ps_leader(const ps_args &A, ps_leader_lds &L) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const model_t &md = A.md;
  const int npla = md.npla, nb = A.nb, nblk = A.nblk;
  const size_t tbase = (size_t)md.Nb * (A.t - 1);
  PROF_T0();
  // the decisions, then the record and flag that release the helpers, are
  // the block's critical chain: their waves issue ahead of their SIMD
  // partners (the fold waves)
  if (wave == 0) __builtin_amdgcn_s_setprio(3);
  else if (wave == 1) __builtin_amdgcn_s_setprio(2);
  int iW = 0, i11 = 1;
  bool ok = true;
  // a backward pass wraps every column with the exponentials before the
  // update; none may change until all helpers have read them
  bool ekv_free = A.wrap != 2;
  // Proposals of block 0 (wave 0) while the other waves gather block 0's
  // window from G itself, W = G[S_0, S_0]: E is block diagonal over the
  // bonds, so on a backward pass the helpers' wrap E^-1 G E restricted to
  // S_0 is E^-1[S_0] G[S_0, S_0] E[S_0], formed here with their operations
  // in their order (columns, then rows), which gives their values.
  if (tid < min(nb, npla)) {
    if (A.pre)
      L.prop[0][tid] = A.pre[tid];
    else
      ps_propose(A, A.f, A.t, A.u, A.fbfi, A.fcross, A.fph, tid, L.prop[0][tid]);
  }
  if (tid == 0) {
    L.abort = 0;
    L.npub = 0;
    L.done = 0;
    L.fbar = 0;
    L.wdone = 0;
    L.nacc = 1;
  }
  {
    const int n0 = ps_nsites(A, 0), Ns = md.Ns;
    zd *W = L.buf[iW];
    for (int e = tid - 64; e >= 0 && e < PS_BS * PS_BS; e += PS_THREADS - 64) {
      const int r = e >> 5, c = e & 31;
      W[r * PS_BS + c] = (r < n0 && c < n0) ? A.G[A.fsites[r] + (size_t)Ns * A.fsites[c]] : zmk(0.0, 0.0);
    }
    __syncthreads();
    if (A.wrap == 2) {
      const ekv_t *ek = A.ekv + ((size_t)(A.t - 1) * 4 + A.f) * npla;
      const int nb0 = n0 / 4;
      for (int e = tid; e < nb0 * n0; e += PS_THREADS) {  // columns: G E
        const int bl = e / n0, r = e % n0;
        ekv_t E;
        ld_coh_ekv(ek + bl, E);
        zd v[4], y[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = W[r * PS_BS + 4 * bl + c];
        ekv_apply_right(E, false, v, y);
#pragma unroll
        for (int c = 0; c < 4; ++c) W[r * PS_BS + 4 * bl + c] = y[c];
      }
      __syncthreads();
      for (int e = tid; e < nb0 * n0; e += PS_THREADS) {  // rows: E^-1 G
        const int bb = e / n0, q = e % n0;
        ekv_t E;
        ld_coh_ekv(ek + bb, E);
        zd v[4], y[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) v[a] = W[(4 * bb + a) * PS_BS + q];
        ekv_apply(E, true, v, y);
#pragma unroll
        for (int a = 0; a < 4; ++a) W[(4 * bb + a) * PS_BS + q] = y[a];
      }
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
      // The accepted exponentials are read by the helpers' forward wrap after
      // the last block (acc flags and fields only by later launches): this
      // wave's stores are complete before wave 1 publishes the last block.
      if (!has_next) ps_release();
      if (lane == 0) {
        L.nacc = m;
        __hip_atomic_store(&L.done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      PROF_ACC(0);
    } else if (wave == 1) {
      // ---- the record of block k, accept by accept: wave 0's 4 x 4 blocks
      // X_m, T[m][m'] = u_m'[S_m] and V[m'][m] = v_m'[S_m] (m' < m),
      // which are final once accept m is handed over ----
      int md_ = 0;
      ps_rec *R = A.rec + k;
      for (;;) {
        PROF_ACC(0);
        const int np = ps_lds_progress(A, L, &L.npub, &L.done, md_);
        PROF_ACC(0);
        if (np <= md_) break;
        for (; md_ < np; ++md_) {
          const int m = md_, j4 = 4 * L.li[m];
          if (lane == 0) st_coh_i(&R->li[m], L.li[m]);
          for (int e = lane; e < 16 * (2 * m + 1); e += 64) {
            const int blk = e >> 4, r = (e >> 2) & 3, c = e & 3;
            if (blk == 0) {
              st_coh(&R->AT[(4 * m + r) * PS_BS + 4 * m + c], L.sX[m][r + 4 * c]);
            } else if (blk <= m) {
              const int mp = blk - 1;
              st_coh(&R->AT[(4 * m + r) * PS_BS + 4 * mp + c], L.sU[mp][(j4 + r) * 4 + c]);
            } else {
              const int mp = blk - 1 - m;
              st_coh(&R->AT[(4 * mp + r) * PS_BS + 4 * m + c], L.sV[mp][r * PS_BS + j4 + c]);
            }
          }
        }
        PROF_ACC(1);
      }
      // ---- publish block k as soon as it is decided: the record went out
      // accept by accept; its stores are complete, then the flag.  The
      // helpers do not wait for this workgroup's fold of the next window. ----
      const int a = L.nacc;
      if (!L.abort) {
        ps_release();
        if (lane == 0) ps_st(ps_flag(A, k), (unsigned long long)(a + 1));
      }
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
    } else if (has_next && wave <= 5) {
      // ---- waves 2..5: the fold tiles after step k-1, then block k's accepts one by one ----
      bool wok = true;
      if (lane == 0)
        wok = ps_wait(A, ps_pub(A, k - 1, 0), ps_owners(A, k), ps_pub(A, k - 1, 1), ps_owners(A, k + 1)) != 0;
      wok = __shfl(wok ? 1 : 0, 0, 64) != 0;
      PROF_ACC(1);
      if (wok) {
        ps_tiles T;
        ps_tiles_load(A, k, T);
        PROF_ACC(7);
        int md_ = 0;
        for (;;) {
          const int np = ps_lds_progress(A, L, &L.npub, &L.done, md_);
          if (np <= md_) break;
          for (; md_ < np; ++md_) {
            ps_tiles_share(L, md_, T);
            // one barrier per accept: the shared column and row are double-buffered by parity
            ps_lds_bar_to(A, L, &L.fbar, 4 * (md_ + 1));
            ps_tiles_fold(L, md_, nsk, T);
          }
        }
        ps_tiles_store(L.buf[i11], T);
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
        PROF_ACC(0);
        for (; md_ < np; ++md_) {
          ps_w_offdiag(L, L.buf[iW], md_, nsk);
          if (lane == 0) __hip_atomic_store(&L.wdone, md_ + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        PROF_ACC(1);
      }
      PROF_ACC(0);
    } else if (has_next && wave == 7) {
      // ---- wave 7: the next block's proposals ----
      const int p1 = (k + 1) * nb, nb1 = ps_nsites(A, k + 1) / 4;
      if (lane < nb1) ps_propose(A, A.f, A.t, A.u, A.fbfi, A.fcross, A.fph, p1 + lane, L.prop[(k + 1) & 1][lane]);
      PROF_ACC(1);
    } else if (!has_next && wave == 7 && A.nxt) {
      // ---- wave 7, last block: the next pass's proposals of its block 0 ----
      if (lane < min(nb, npla)) ps_propose(A, A.nx_f, A.nx_t, A.nx_u, A.nx_fbfi, A.nx_fcross, A.nx_fph, lane, A.nxt[lane]);
      PROF_ACC(1);
    }
    if (tid == 0 && ps_ld(&A.sync[2]) != 0) L.abort = 1;
    __syncthreads();
    if (L.abort) { ok = false; break; }
    PROF_ACC(2);
    if (!has_next) break;
    if (wave == 0)
      PROF_ACC(1);
    else
      PROF_ACC(6);
    if (tid == 0) {
      L.npub = 0;
      L.done = 0;
        L.fbar = 0;
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
  zd sAT[PS_BS * PS_BS];       // the block's record (X, T and V blocks, stride PS_BS)
  zd sB[PS_BS * PS_BS];        // v, then s, per own column (stride PS_BS)
  int li[PS_NB];
  int pos[PS_NB];              // position of block-local bond b among the accepts, or -1
  int a, ok, sink;
};
// after the header: sw, sy [4 bph][PS_BS], sC [4 bph][Ns], then the family's
// sites [4 npla] (int)

// Own columns: row r of local column q at sC[q * Ns + r].
static __device__ __forceinline__ void ps_wrap_cols(const ps_args &A, zd *sC, int b0, int b1, bool forward) {
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

// After step kstep the tile slot holds the columns of blocks kstep+1, kstep+2
// at the rows of those two blocks (the leader's fold tiles), and the slot the
// columns of block kstep+1 (the pivot columns of step kstep+1).  The tiles go
// out first: the leader waits for them, the helpers' next step for the slot.
static __device__ __forceinline__ void ps_publish_tiles(const ps_args &A, const zd *sC, const int *sS, int b0, int b1,
                                                       int kstep) {
  const int Ns = A.md.Ns, tid = threadIdx.x;
  for (int b = b0; b < b1; ++b) {
    const int blk = b / A.nb;
    const int mb = blk - (kstep + 1);
    if (mb < 0 || mb > 1) continue;
    const int cl0 = 4 * (b - blk * A.nb);
    const zd *col = sC + 4 * (b - b0) * Ns;
    for (int e = tid; e < 4 * 2 * PS_BS; e += PS_THREADS) {
      const int c = e >> 6, i = e & 63, j = kstep + 1 + (i >> 5), r = i & 31;
      if (j < A.nblk && r < ps_nsites(A, j)) st_coh(ps_tcol(A, kstep, mb, cl0 + c) + i, col[c * Ns + sS[4 * A.nb * j + r]]);
    }
  }
}
static __device__ __forceinline__ void ps_publish_slot(const ps_args &A, const zd *sC, int b0, int b1, int kstep) {
  const int Ns = A.md.Ns, tid = threadIdx.x;
  for (int b = b0; b < b1; ++b) {
    const int blk = b / A.nb;
    if (blk != kstep + 1) continue;
    const int cl0 = 4 * (b - blk * A.nb);
    const zd *col = sC + 4 * (b - b0) * Ns;
    for (int e = tid; e < 4 * Ns; e += PS_THREADS) {
      const int c = e / Ns, r = e % Ns;
      st_coh(ps_scol(A, kstep, 0, cl0 + c) + r, col[c * Ns + r]);
    }
  }
}
// Publish after step kstep (all threads): tiles, their count, slot, its count.
static __device__ __forceinline__ void ps_announce(const ps_args &A, int b0, int b1, int kstep);
static __device__ __forceinline__ void ps_publish_cols(const ps_args &A, const zd *sC, const int *sS, int b0, int b1,
                                                      int kstep) {
  bool tiles = false, slot = false;
  for (int b = b0; b < b1; ++b) {
    const int mb = b / A.nb - (kstep + 1);
    tiles = tiles || mb == 0 || mb == 1;
    slot = slot || mb == 0;
  }
  if (!tiles) return;  // uniform
  ps_publish_tiles(A, sC, sS, b0, b1, kstep);
  ps_release();
  __syncthreads();
  if (threadIdx.x == 0) ps_announce(A, b0, b1, kstep);
  if (slot) {
    ps_publish_slot(A, sC, b0, b1, kstep);
    ps_release();
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(ps_spub(A, kstep), 1ull);
  }
}

// One thread: count this helper among the publishers of the blocks
// kstep+1, kstep+2 that it owns bonds of (after its stores are complete).
static __device__ __forceinline__ void ps_announce(const ps_args &A, int b0, int b1, int kstep) {
  bool own0 = false, own1 = false;
  for (int b = b0; b < b1; ++b) {
    const int mb = b / A.nb - (kstep + 1);
    own0 = own0 || mb == 0;
    own1 = own1 || mb == 1;
  }
  if (own0) atomicAdd(ps_pub(A, kstep, 0), 1ull);
  if (own1) atomicAdd(ps_pub(A, kstep, 1), 1ull);
}

// y[q] = (y_0, .., y_{a-1}) of the own columns q from the record (the 4 x 4
// blocks X, T and V of ps_rec):
//     v_m = w_m - sum_{m'<m} T[m][m'] v_m',    w_m = (I - G)[S_m, q],
//     y_m = X_m z_m,  z_m = v_m + sum_{m'>m} V[m][m'] y_m'   (m = a-1 .. 0),
// so that G[:, q] += G[:, S_A] y[q] is the block's rank-4 updates in turn.
// Wave w takes columns 4w..4w+3; lane (ql, c, d) = (l / 16, l / 4 % 4, l % 4)
// forms component c of column 4w+ql, its sums over d reduced in the quad.
// No workgroup barrier: the waves' columns are independent.  The vectors
// stay in registers between the steps: a quad sum leaves component c of v_m
// in lane (ql, c, d), one shuffle gives each lane the component d that its
// next products need, and the backward step's row reduction leaves y_m[d]
// where the next backward products take it.  v goes to sv[q * PS_BS + 4m + c]
// for the backward half, y to sy[q * PS_BS + 4m + d] for the pivot update.
static __device__ __forceinline__ void ps_helper_y(const ps_args &A, const ps_helper_hdr &Hh, int k, int a, int ncol,
                                                   int b0, const zd *sC, const int *sS, zd *sv, zd *sy) {
  const int Ns = A.md.Ns, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (4 * wv >= ncol) return;  // uniform
  const int ql = lane >> 4, c = (lane >> 2) & 3, d = lane & 3, q = 4 * wv + ql;
  const bool qok = q < ncol, lead = qok && d == 0;
  const int *Sk = sS + 4 * A.nb * k;
  const int siteq = qok ? sS[4 * b0 + q] : -1;
  const zd *M = Hh.sAT;
  // after the quad sum lane (ql, c, d) holds component c; lane (ql, d, c) holds d
  const int tsrc = (lane & 48) | (d << 2) | c;
  zd vd[PS_NB], yd[PS_NB];  // v_m[d] and y_m[d] of this lane's column
#pragma unroll
  for (int m = 0; m < PS_NB; ++m) vd[m] = yd[m] = zmk(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < PS_NB; ++m) {
    if (m < a) {  // uniform
      zd acc = zmk(0.0, 0.0);
      if (lead) {
        const int st = Sk[4 * Hh.li[m] + c];
        acc = zsub(zmk(siteq == st ? 1.0 : 0.0, 0.0), sC[q * Ns + st]);
      }
      zd acc1 = zmk(0.0, 0.0);
      if (qok) {
#pragma unroll
        for (int mp = 0; mp < m; ++mp) {
          const zd t = zmul(M[(4 * m + c) * PS_BS + 4 * mp + d], vd[mp]);
          if (mp & 1)
            acc1 = zsub(acc1, t);
          else
            acc = zsub(acc, t);
        }
      }
      acc = zadd(acc, acc1);
      acc.x = quad_sum(acc.x);
      acc.y = quad_sum(acc.y);
      if (lead) sv[q * PS_BS + 4 * m + c] = acc;  // read by the backward half only
      vd[m] = shfl_z(acc, tsrc);
    }
  }
  wave_sync();
  // z_m = v_m + sum_{m'>m} V[m][m'] y_m', then y_m = X_m z_m (the record holds V above the diagonal)
#pragma unroll
  for (int m = PS_NB - 1; m >= 0; --m) {
    if (m < a) {  // uniform
      zd acc0 = zmk(0.0, 0.0), acc1 = acc0;
      if (qok) {
        if (lead) acc0 = sv[q * PS_BS + 4 * m + c];
#pragma unroll
        for (int mp = m + 1; mp < PS_NB; mp += 2) {
          if (mp < a) {  // uniform
            acc1 = zadd(acc1, zmul(M[(4 * m + c) * PS_BS + 4 * mp + d], yd[mp]));
            if (mp + 1 < a) acc0 = zadd(acc0, zmul(M[(4 * m + c) * PS_BS + 4 * mp + 4 + d], yd[mp + 1 < PS_NB ? mp + 1 : mp]));
          }
        }
      }
      const zd xdc = M[(4 * m + d) * PS_BS + 4 * m + c];  // X_m[d][c]
      zd zv = zadd(acc0, acc1);
      zv.x = quad_sum(zv.x);
      zv.y = quad_sum(zv.y);
      // lane (ql, c, d) holds z_m[c]; y_m[d] = sum_c X_m[d][c] z_m[c] over the row's lanes of equal d,
      // which leaves y_m[d] where the next steps use it
      zd yv = zmul(xdc, zv);
      yv.x = row4_sum(yv.x);
      yv.y = row4_sum(yv.y);
      yd[m] = yv;
      if (qok && c == 0) sy[q * PS_BS + 4 * m + d] = yv;
    }
  }
  wave_sync();
}

// G[r, q] += sum_g G_k[r, S_A(g)] y[q][g] for the own columns, four at a
// time.  All pivot loads of a row are issued before the first use (entries
// past 4a read one fixed word, a single request per wave, and are not used);
// y is staged in registers four pivots at a time.
static __device__ __forceinline__ void ps_pivot_update(const ps_args &A, const ps_helper_hdr &Hh, int kstep, int n4, int ncol,
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
// columns G_k[:, S_A] and Y[g][q] = y[q][g], complex as three real products.
// Lane l supplies P[t0 + l%16][g0 + l/16] and Y[g0 + l/16][l%16] and holds C
// rows l/16 + 4i of column l%16.  All pivot loads of a tile are issued
// before the first product.
static __device__ __forceinline__ void ps_pivot_update_mfma(const ps_args &A, const ps_helper_hdr &Hh, int kstep, int n4,
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
    // three real products accumulated over the chunks: re = P1 - P2,
    // im = P3 - P1 - P2
    double4_t p1 = {0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      if (4 * ch < n4) {  // uniform
        p1 = mfma_f64(p[j][ch].x, y[ch].x, p1);
        p2 = mfma_f64(p[j][ch].y, y[ch].y, p2);
        p3 = mfma_f64(p[j][ch].x + p[j][ch].y, y[ch].x + y[ch].y, p3);
      }
    }
    const double4_t cr = p1 - p2, ci = (p3 - p1) - p2;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = t0 + lk + 4 * i;
      if (row < Ns) sC[li * Ns + row] = zadd(sC[li * Ns + row], zmk(cr[i], ci[i]));
    }
  }
}

template <int BPH, bool PRE>
static __device__ __forceinline__ void ps_helper(const ps_args &A, ps_helper_hdr &Hh, zd *sC, zd *sw, zd *sy, int *sS) {
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
  // this workgroup has read the exponentials its backward wrap needs
  if (tid == 0) atomicAdd(ps_wrapped(A), 1ull);
  PROF_ACC(0);
  for (int k = 0; k < A.nblk; ++k) {
    // Block k's columns are in the slot once their owners have published
    // after step k-1; the leader's decision of block k comes later.
    // Meanwhile the helpers of an XCD (workgroups are spread over the 8 XCDs
    // round-robin) touch those columns in slices, so that the pivot loads of
    // step k hit in L2.
    if (tid == 0) Hh.ok = ps_wait(A, ps_spub(A, k - 1), ps_owners(A, k), nullptr, 0) != 0;
    __syncthreads();
    if (!Hh.ok) break;
    // With one bond per helper and one row per thread, each thread loads its
    // row of all the block's columns G_{k-1}[r, S_k] into registers while the
    // leader decides the block, so that the update after the decision has no
    // memory latency left (entries past the block read column 0 and are not
    // used).  Otherwise the helpers of an XCD touch the columns in L2.
    constexpr bool pre = PRE;
    zd Pk[PRE ? PS_BS : 1];
    if constexpr (pre) {
      const int nsk = ps_nsites(A, k), r = tid < Ns ? tid : 0;
#pragma unroll
      for (int g = 0; g < (PRE ? PS_BS : 1); ++g) Pk[g] = ps_scol(A, k - 1, 0, g < nsk ? g : 0)[r];
    } else {
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
      if (tid < PS_NB) Hh.pos[tid] = -1;
      for (int e = tid; e < n4 * n4; e += PS_THREADS) {
        const int r = e / n4, c = e % n4;
        Hh.sAT[r * PS_BS + c] = R->AT[r * PS_BS + c];
      }
      __syncthreads();
      PROF_ACC(2);
      if (tid < a) Hh.pos[Hh.li[tid]] = tid;
      ps_helper_y(A, Hh, k, a, ncol, b0, sC, sS, Hh.sB, sy);
      __syncthreads();
      PROF_ACC(3);
      // G[:, q] += G_k[:, S_A] y[q]; pivot columns from the slot of step k-1
      if constexpr (BPH == 4) {
        ps_pivot_update_mfma(A, Hh, k - 1, n4, sy, sC);
      } else if constexpr (pre) {
        // the accepted bonds' columns from the registers (uniform branches)
        if (tid < Ns) {
          // one accumulator per (column, pivot within the bond): short FMA chains
          zd acc[NC][4];
#pragma unroll
          for (int q = 0; q < NC; ++q)
#pragma unroll
            for (int d = 0; d < 4; ++d) acc[q][d] = zmk(0.0, 0.0);
          int pos[PS_NB];
#pragma unroll
          for (int b = 0; b < PS_NB; ++b) pos[b] = Hh.pos[b];
#pragma unroll
          for (int b = 0; b < (PRE ? PS_NB : 0); ++b) {
            const int m = pos[b];
            if (m >= 0) {
#pragma unroll
              for (int d = 0; d < 4; ++d)
#pragma unroll
                for (int q = 0; q < NC; ++q) acc[q][d] = zfma(acc[q][d], Pk[4 * b + d], sy[q * PS_BS + 4 * m + d]);
            }
          }
#pragma unroll
          for (int q = 0; q < NC; ++q)
            sC[q * Ns + tid] = zadd(sC[q * Ns + tid], zadd(zadd(acc[q][0], acc[q][1]), zadd(acc[q][2], acc[q][3])));
        }
      } else {
        ps_pivot_update(A, Hh, k - 1, n4, NC, sy, sC);
      }
      __syncthreads();
    }
    PROF_ACC(5);
    ps_publish_cols(A, sC, sS, b0, b1, k);
    PROF_ACC(6);
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
    // one bond per helper with one row per thread keeps the pivot rows in registers
    if (A.bph == 1 && A.md.Ns <= PS_THREADS)
      ps_helper<1, true>(A, Hh, sC, sw, sy, sS);
    else if (A.bph == 1)
      ps_helper<1, false>(A, Hh, sC, sw, sy, sS);
    else if (A.bph == 2)
      ps_helper<2, false>(A, Hh, sC, sw, sy, sS);
    else
      ps_helper<4, false>(A, Hh, sC, sw, sy, sS);
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
      for (int k = 0; k < A.nblk; ++k) {
        ps_st(ps_flag(A, k), 0ull);
        ps_st(ps_spub(A, k - 1), 0ull);
      }
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
