// One family pass at one time slice: the local phonon-field updates of
// update_phonon_field (src/mod_update.f90:13-201), fused with the two-way
// wrap of the same family (right_evolve before it on a backward sweep,
// left_evolve after it on a forward sweep, src/mod_evolution.f90:236-328).
//
// The reference visits the family's bonds in order and, after each accepted
// bond i, applies the rank-4 update
//     G <- G + u_i v_i,   u_i = G[:, S_i] X_i,   v_i = (I - G)[S_i, :],
//     X_i = Delta_i (I + G[S_i, S_i] Delta_i)^-1.
// Here the bonds are cut into decision blocks of nb consecutive bonds and
// the work is split between one leader workgroup and H helper workgroups:
//
//  * Helper h owns the columns of bph bonds and keeps them in LDS for the
//    whole pass.  After block k is decided it applies the block's accepts
//    to its columns in one step: with S_A the sites of the a accepted bonds,
//        G[:, q] += G_k[:, S_A] (A v[q]),
//    where G_k is G before block k, v[q] stacks v_m[q] (m = 1..a) and the
//    4a x 4a block-upper-triangular A folds the in-block sequence:
//        A[:, m] = (sel_m + sum_{m'<m} A[:, m'] v_{m'}[S_m]) X_m,
//        v_m[q]  = e_q(S_m) - G_k[S_m, q] - sum_{m'<m} u_{m'}[S_m] v_{m'}[q].
//    The pivot columns G_k[:, S_A] are read from a slot that the owners of
//    block k's columns filled after step k-1.
//  * The leader decides the bonds of block k in one wave from an LDS copy
//    W = G[S_k, S_k] that it keeps current with the rank-4 updates (only the
//    rows and columns of the block's later bonds change).  It publishes the
//    accepted bonds with X_m and the 4x4 blocks u_m[S_j], v_m[S_j] of every
//    later bond j, which is all a helper needs.  Before block k+1 it needs
//    G[S_{k+1}, S_{k+1}] after block k: it takes G[S_{k+1}, S_k],
//    G[S_k, S_{k+1}] and G[S_{k+1}, S_{k+1}] after block k-1 from the slot
//    of step k-1 (filled while it decided block k) and replays block k's
//    accepts on them.  The helpers therefore lag the leader by one block.
//
// Accept test: the reference rejects when prob < u with prob =
// min(1, exp(min(10, w))), w = dS + ncopy*(1+TR)*log|det(I + G_SS Delta)|
// (cal_prob, :475-500).  As log u < 0 this is exactly w < log u, i.e.
// |det|^2 < exp(2 (log u - dS) / (ncopy*(1+TR))); the right-hand side does
// not depend on G and is prepared before the bond's turn.
#include "pqmc_common.h"

#include "k_pass.h"

struct ps_prop {
  zd D[16];  // Delta = E(old)^-1 E(new) - I, column-major
  ekv_t en;
  double thr, xnew;
  int fld, pad[3];
};

// Data exchanged between the workgroups of a pass (records, slot columns,
// updated exponentials) is written and read only with device-scope
// (sc1) accesses, which are coherent across the XCDs' L2 caches.  A
// publisher therefore needs no L2 write-back: it waits for its own stores
// to complete (vmcnt 0) and then raises a flag; a reader that saw the flag
// issues its coherent loads afterwards.  Agent-scope fences, which write
// back and invalidate the whole L2 of the XCD, are avoided on purpose.
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

// Single-lane wait until *p >= v (and *q >= w when q is set).  Bounded by
// wall-clock time and by the abort word.
static __device__ bool ps_wait(const ps_args &A, unsigned long long *p, unsigned long long v,
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
      return true;
    }
    if ((it & 31) == 0) {
      if (ps_ld(&A.sync[2]) != 0) return false;
      if (wall_clock64() - t0 > PS_WAIT_TICKS) {
        __hip_atomic_store(&A.sync[2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(&A.stats[2], 1ull);
        return false;
      }
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

static __device__ __forceinline__ unsigned long long *ps_arr(const ps_args &A, int k) { return &A.sync[8 + k + 1]; }
static __device__ __forceinline__ int ps_sidx(int k) { return (k + 1) & 1; }  // slot written after step k
static __device__ __forceinline__ zd *ps_scol(const ps_args &A, int s, int mb, int cl) {
  return A.slot + ((size_t)((s * 2 + mb) * PS_BS + cl)) * A.md.Ns;
}
static __device__ __forceinline__ int ps_nsites(const ps_args &A, int k) {
  return 4 * min(A.nb, A.md.npla - k * A.nb);
}

static __device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

static __device__ __forceinline__ zd shfl_z(zd v, int src) {
  return zmk(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}
static __device__ __forceinline__ zd shfl_xor_z(zd v, int m) {
  return zmk(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

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

#define PROF_T0() unsigned long long _pt = A.prof ? wall_clock64() : 0ull
#define PROF_ACC(k)                                                         \
  do {                                                                      \
    if (A.prof && threadIdx.x == 0) {                                       \
      const unsigned long long _now = wall_clock64();                       \
      A.prof[blockIdx.x * 8 + (k)] += _now - _pt;                           \
      _pt = _now;                                                           \
    }                                                                       \
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

// Gather a tile G[rows of block rb, cols of block cb] from the slot holding
// blocks (base, base+1) after a step.  T is row-major with stride PS_BS.
static __device__ void ps_load_tile(const ps_args &A, int s, int base, int rb, int cb, zd *T, int t0, int nt) {
  const int nr = ps_nsites(A, rb), nc = ps_nsites(A, cb);
  const int *rs = A.fsites + 4 * A.nb * rb;
  for (int e = t0; e < PS_BS * PS_BS; e += nt) {
    const int r = e / PS_BS, c = e % PS_BS;
    zd v = zmk(0.0, 0.0);
    if (r < nr && c < nc) v = ld_coh(ps_scol(A, s, cb - base, c) + rs[r]);
    T[e] = v;
  }
}

// --------------------------------------------------------------------------
// Leader: decisions, records, replay.
// --------------------------------------------------------------------------
struct ps_leader_lds {
  zd buf[4][PS_BS * PS_BS];     // W, T10, T01, T11 (rotating)
  zd sU[PS_NB][PS_BS * 4];      // u_m at the block's rows (row*4 + comp)
  zd sV[PS_NB][4 * PS_BS];      // v_m at the block's cols (comp*BS + col)
  zd sUN[PS_BS * 4], sVN[4 * PS_BS];
  zd sX[PS_NB][16];
  zd sM[16];
  ps_prop prop[2][PS_NB];
  int li[PS_NB];
  int nacc, abort;
};

static __device__ void ps_leader(const ps_args &A, ps_leader_lds &L) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const model_t &md = A.md;
  const int npla = md.npla, nb = A.nb, nblk = A.nblk;
  const size_t tbase = (size_t)md.Nb * (A.t - 1);
  PROF_T0();
  int iW = 0, i10 = 1, i01 = 2, i11 = 3;
  bool ok = true;
  // proposals of block 0; the window after the helpers' load
  if (tid < min(nb, npla)) ps_propose(A, tid, L.prop[0][tid]);
  if (tid == 0) {
    L.abort = 0;
    L.nacc = ps_wait(A, ps_arr(A, -1), (unsigned long long)A.H, nullptr, 0) ? 1 : 0;
  }
  __syncthreads();
  ok = L.nacc != 0;
  if (ok) ps_load_tile(A, ps_sidx(-1), 0, 0, 0, L.buf[iW], tid, PS_THREADS);
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
      for (int j = 0; j < nbk; ++j) {
        // M = I + G_SS Delta (src/mod_update.f90:97-100), one entry per lane
        if (lane < 16) {
          const int a = lane & 3, c = lane >> 2;
          zd mv = zmk(a == c ? 1.0 : 0.0, 0.0);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) mv = zadd(mv, zmul(W[(4 * j + a) * PS_BS + 4 * j + kk], pp[j].D[kk + 4 * c]));
          L.sM[lane] = mv;
        }
        wave_sync();
        zd cof = zmk(0.0, 0.0);
        if (lane < 16) {
          const int ri = lane & 3, cj = lane >> 2;
          cof = minor3(L.sM, ri, cj);
          if ((ri + cj) & 1) cof = zneg(cof);
        }
        // det = sum_j M(0,j) C(0,j): lanes 0, 4, 8, 12
        zd term = (lane < 16 && (lane & 3) == 0) ? zmul(L.sM[lane], cof) : zmk(0.0, 0.0);
        term = zadd(term, shfl_xor_z(term, 4));
        term = zadd(term, shfl_xor_z(term, 8));
        const zd det = shfl_z(term, 0);
        const bool acc = !((det.x * det.x + det.y * det.y) < pp[j].thr);
        if (lane == 0) A.acc[p0 + j] = (uint8_t)acc;
        PROF_ACC(4);
        if (!acc) {
          ++total_rej;
          continue;
        }
        // X = Delta M^-1 with M^-1 = adj(M)/det; lane k + 4c holds X(k, c)
        const int ia = lane & 3, ic = (lane >> 2) & 3;
        const zd inv = zdiv(shfl_z(cof, ic + 4 * ia), det);
        zd xv = zmk(0.0, 0.0);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) xv = zadd(xv, zmul(pp[j].D[ia + 4 * kk], shfl_z(inv, kk + 4 * ic)));
        if (lane < 16) {
          L.sX[m][lane] = xv;
          st_coh(&A.rec[k].X[m][lane], xv);
        }
        if (lane == 0) {
          L.li[m] = j;
          const int p = p0 + j;
          A.x[pp[j].fld + tbase] = pp[j].xnew;
          st_coh_ekv(&A.ekv[((size_t)(A.t - 1) * 4 + A.f) * npla + p], pp[j].en);
          // ln_cw += log(real(R_elec)) (src/mod_update.f90:142)
          zd rel = det;
          if (md.tr_weight) rel = zmul(rel, zconj(rel));
          zd rp = rel;
          for (int c = 1; c < md.ncopy; ++c) rp = zmul(rp, rel);
          atomicAdd(A.lncw_acc, log(rp.x));
        }
        wave_sync();
        PROF_ACC(5);
        // u_m, v_m at the later sites of the block; then the window update
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
        PROF_ACC(6);
        for (int e = lane; e < Ln * 4; e += 64) {
          // ub[m][jj][r*4 + c] = u_m[S_jj(r)][c]; vb[m][jj][c*4 + s] = v_m[c][S_jj(s)]
          const int r = o + (e >> 2), c = e & 3;
          st_coh(&A.rec[k].ub[m][r >> 2][(r & 3) * 4 + c], L.sU[m][r * 4 + c]);
          const int ca = e / Ln, cc = o + e % Ln;
          st_coh(&A.rec[k].vb[m][cc >> 2][ca * 4 + (cc & 3)], L.sV[m][ca * PS_BS + cc]);
        }
        for (int e = lane; e < Ln * Ln; e += 64) {
          const int r = o + e / Ln, c = o + e % Ln;
          zd s = W[r * PS_BS + c];
#pragma unroll
          for (int a = 0; a < 4; ++a) s = zadd(s, zmul(L.sU[m][r * 4 + a], L.sV[m][a * PS_BS + c]));
          W[r * PS_BS + c] = s;
        }
        wave_sync();
        PROF_ACC(7);
        ++m;
        ++total_acc;
      }
      if (lane == 0) {
        st_coh_i(&A.rec[k].a, m);
        for (int q = 0; q < m; ++q) st_coh_i(&A.rec[k].li[q], L.li[q]);
        L.nacc = m;
      }
      ps_release();
      if (lane == 0) ps_st(&A.sync[0], (unsigned long long)(k + 1));
      PROF_ACC(1);
    } else if (has_next) {
      // ---- waves 1..7: next block's proposals and tiles after step k-1 ----
      const int t1 = tid - 64, nt1 = PS_THREADS - 64;
      const int p1 = (k + 1) * nb, nb1 = ps_nsites(A, k + 1) / 4;
      if (t1 < nb1) ps_propose(A, p1 + t1, L.prop[(k + 1) & 1][t1]);
      bool wok = true;
      if (lane == 0) wok = ps_wait(A, ps_arr(A, k - 1), (unsigned long long)A.H, nullptr, 0);
      wok = __shfl(wok ? 1 : 0, 0, 64) != 0;
      if (wok) {
        const int s = ps_sidx(k - 1);
        ps_load_tile(A, s, k, k + 1, k, L.buf[i10], t1, nt1);
        ps_load_tile(A, s, k, k, k + 1, L.buf[i01], t1, nt1);
        ps_load_tile(A, s, k, k + 1, k + 1, L.buf[i11], t1, nt1);
      } else if (lane == 0) {
        L.abort = 1;
      }
    }
    if (tid == 0 && ps_ld(&A.sync[2]) != 0) L.abort = 1;
    __syncthreads();
    if (L.abort) { ok = false; break; }
    PROF_ACC(2);
    if (!has_next) break;
    // ---- replay block k's accepts on the tiles of block k+1 ----
    const int m_acc = L.nacc;
    const int nn = ps_nsites(A, k + 1);
    zd *T10 = L.buf[i10], *T01 = L.buf[i01], *T11 = L.buf[i11];
    for (int mm = 0; mm < m_acc; ++mm) {
      const int j = L.li[mm];
      const int o = 4 * (j + 1), Ln = nsk - o;
      for (int e = tid; e < nn * 4; e += PS_THREADS) {
        const int r = e >> 2, c = e & 3;
        zd s = zmk(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < 4; ++a) s = zadd(s, zmul(T10[r * PS_BS + 4 * j + a], L.sX[mm][a + 4 * c]));
        L.sUN[r * 4 + c] = s;
        const int ca = e / nn, cc = e % nn;
        L.sVN[ca * PS_BS + cc] = zneg(T01[(4 * j + ca) * PS_BS + cc]);
      }
      __syncthreads();
      const int n10 = nn * Ln, n01 = Ln * nn, n11 = nn * nn;
      for (int e = tid; e < n10 + n01 + n11; e += PS_THREADS) {
        if (e < n10) {
          const int r = e / Ln, c = o + e % Ln;
          zd s = T10[r * PS_BS + c];
#pragma unroll
          for (int a = 0; a < 4; ++a) s = zadd(s, zmul(L.sUN[r * 4 + a], L.sV[mm][a * PS_BS + c]));
          T10[r * PS_BS + c] = s;
        } else if (e < n10 + n01) {
          const int e2 = e - n10;
          const int r = o + e2 / nn, c = e2 % nn;
          zd s = T01[r * PS_BS + c];
#pragma unroll
          for (int a = 0; a < 4; ++a) s = zadd(s, zmul(L.sU[mm][r * 4 + a], L.sVN[a * PS_BS + c]));
          T01[r * PS_BS + c] = s;
        } else {
          const int e2 = e - n10 - n01;
          const int r = e2 / nn, c = e2 % nn;
          zd s = T11[r * PS_BS + c];
#pragma unroll
          for (int a = 0; a < 4; ++a) s = zadd(s, zmul(L.sUN[r * 4 + a], L.sVN[a * PS_BS + c]));
          T11[r * PS_BS + c] = s;
        }
      }
      __syncthreads();
    }
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
}

// --------------------------------------------------------------------------
// Helpers: own columns, wraps, block steps.
// --------------------------------------------------------------------------
struct ps_helper_hdr {
  zd sA[PS_BS * PS_BS];        // A (row-major, 4a x 4a)
  zd sT[PS_BS * 4];            // A-recurrence temporaries
  zd sX[PS_NB][16];
  zd sUB[PS_NB * (PS_NB - 1) / 2][16];  // u_{m'}[S_m] for m' < m
  zd sVB[PS_NB * (PS_NB - 1) / 2][16];  // v_{m'}[S_m]
  int li[PS_NB];
  int a, ok;
};

static __device__ __forceinline__ int ps_pair(int mp, int m) { return m * (m - 1) / 2 + mp; }

// Own columns: row r of local column q at sC[q * Ns + r]; sv/sy are [ncol][PS_BS].
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

static __device__ void ps_publish_cols(const ps_args &A, const zd *sC, int b0, int b1, int kstep) {
  // after step kstep the slot holds the columns of blocks kstep+1, kstep+2
  const int Ns = A.md.Ns, tid = threadIdx.x;
  const int s = ps_sidx(kstep);
  for (int b = b0; b < b1; ++b) {
    const int blk = b / A.nb;
    const int mb = blk - (kstep + 1);
    if (mb < 0 || mb > 1) continue;
    const int cl0 = 4 * (b - blk * A.nb);
    for (int e = tid; e < 4 * Ns; e += PS_THREADS) {
      const int c = e / Ns, r = e % Ns;
      st_coh(ps_scol(A, s, mb, cl0 + c) + r, sC[(4 * (b - b0) + c) * Ns + r]);
    }
  }
}

static __device__ void ps_helper(const ps_args &A, ps_helper_hdr &Hh, zd *sC, zd *sv, zd *sy) {
  const int tid = threadIdx.x, lane = tid & 63;
  const model_t &md = A.md;
  const int Ns = md.Ns, npla = md.npla;
  const int h = blockIdx.x - 1;
  const int b0 = h * A.bph, b1 = min(npla, b0 + A.bph);
  const int ncol = 4 * (b1 - b0);
  PROF_T0();
  // ---- load own columns (natural column = site), optional backward wrap ----
  for (int e = tid; e < ncol * Ns; e += PS_THREADS) {
    const int q = e / Ns, r = e % Ns;
    sC[e] = A.G[r + (size_t)Ns * A.fsites[4 * b0 + q]];
  }
  __syncthreads();
  if (A.wrap == 2) ps_wrap_cols(A, sC, b0, b1, false);
  ps_publish_cols(A, sC, b0, b1, -1);
  ps_release();
  __syncthreads();
  if (tid == 0) atomicAdd(ps_arr(A, -1), 1ull);
  PROF_ACC(0);
  for (int k = 0; k < A.nblk; ++k) {
    if (tid == 0) {
      Hh.ok = ps_wait(A, &A.sync[0], (unsigned long long)(k + 1), ps_arr(A, k - 1), (unsigned long long)A.H) ? 1 : 0;
      Hh.a = Hh.ok ? ld_coh_i(&A.rec[k].a) : 0;
    }
    __syncthreads();
    if (!Hh.ok) break;
    PROF_ACC(1);
    const int a = Hh.a;
    if (a > 0) {
      const ps_rec *R = A.rec + k;
      if (tid < a) Hh.li[tid] = ld_coh_i(&R->li[tid]);
      if (tid < 16 * a) Hh.sX[tid >> 4][tid & 15] = ld_coh(&R->X[tid >> 4][tid & 15]);
      __syncthreads();
      const int npair = a * (a - 1) / 2;
      for (int e = tid; e < npair * 16; e += PS_THREADS) {
        const int pr = e >> 4, q = e & 15;
        int m = 1;
        while (m * (m + 1) / 2 <= pr) ++m;
        const int mp = pr - m * (m - 1) / 2;
        Hh.sUB[pr][q] = ld_coh(&R->ub[mp][Hh.li[m]][q]);
        Hh.sVB[pr][q] = ld_coh(&R->vb[mp][Hh.li[m]][q]);
      }
      __syncthreads();
      const int n4 = 4 * a;
      const int *Sk = A.fsites + 4 * A.nb * k;
      // v for own columns: threads (q, c), sequential over m (one wave covers all)
      if (tid < 4 * ncol) {
        const int q = tid >> 2, c = tid & 3;
        const int qsite = A.fsites[4 * b0 + q];
        for (int m = 0; m < a; ++m) {
          const int s = Sk[4 * Hh.li[m] + c];
          zd val = zsub(zmk(qsite == s ? 1.0 : 0.0, 0.0), sC[q * Ns + s]);
          for (int mp = 0; mp < m; ++mp) {
            const zd *ub = Hh.sUB[ps_pair(mp, m)];
#pragma unroll
            for (int d = 0; d < 4; ++d) val = zsub(val, zmul(ub[c * 4 + d], sv[q * PS_BS + 4 * mp + d]));
          }
          sv[q * PS_BS + 4 * m + c] = val;
          wave_sync();
        }
      }
      // A: threads (e, kk), e < 4a; block column m = (sel_m + sum A[:, m'] v_{m'}[S_m]) X_m
      for (int m = 0; m < a; ++m) {
        if (tid < n4 * 4) {
          const int e = tid >> 2, kk = tid & 3;
          zd t = zmk(e == 4 * m + kk ? 1.0 : 0.0, 0.0);
          for (int mp = 0; mp < m; ++mp) {
            const zd *vb = Hh.sVB[ps_pair(mp, m)];
#pragma unroll
            for (int d = 0; d < 4; ++d) t = zadd(t, zmul(Hh.sA[e * PS_BS + 4 * mp + d], vb[d * 4 + kk]));
          }
          Hh.sT[e * 4 + kk] = t;
        }
        __syncthreads();
        if (tid < n4 * 4) {
          const int e = tid >> 2, c = tid & 3;
          zd s = zmk(0.0, 0.0);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) s = zadd(s, zmul(Hh.sT[e * 4 + kk], Hh.sX[m][kk + 4 * c]));
          Hh.sA[e * PS_BS + 4 * m + c] = s;
        }
        __syncthreads();
      }
      // y[q] = A v[q]
      for (int e = tid; e < ncol * n4; e += PS_THREADS) {
        const int q = e / n4, r = e % n4;
        zd s = zmk(0.0, 0.0);
        for (int f = r & ~3; f < n4; ++f) s = zadd(s, zmul(Hh.sA[r * PS_BS + f], sv[q * PS_BS + f]));
        sy[q * PS_BS + r] = s;
      }
      __syncthreads();
      PROF_ACC(2);
      // G[:, q] += G_k[:, S_A] y[q]; pivot columns from the slot of step k-1
      const int s = ps_sidx(k - 1);
      for (int r = tid; r < Ns; r += PS_THREADS) {
        zd accq[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) accq[q] = zmk(0.0, 0.0);
        for (int f0 = 0; f0 < n4; f0 += 16) {
          zd P[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int f = f0 + i;
            P[i] = (f < n4) ? ld_coh(ps_scol(A, s, 0, 4 * Hh.li[f >> 2] + (f & 3)) + r) : zmk(0.0, 0.0);
          }
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int f = f0 + i;
            if (f < n4) {
#pragma unroll
              for (int q = 0; q < 8; ++q)
                if (q < ncol) accq[q] = zadd(accq[q], zmul(P[i], sy[q * PS_BS + f]));
            }
          }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (q < ncol) sC[q * Ns + r] = zadd(sC[q * Ns + r], accq[q]);
      }
      __syncthreads();
    }
    PROF_ACC(3);
    ps_publish_cols(A, sC, b0, b1, k);
    PROF_ACC(5);
    ps_release();
    PROF_ACC(6);
    __syncthreads();
    if (tid == 0) atomicAdd(ps_arr(A, k), 1ull);
    PROF_ACC(7);
  }
  (void)lane;
  if (A.wrap == 1) ps_wrap_cols(A, sC, b0, b1, true);
  for (int e = tid; e < ncol * Ns; e += PS_THREADS) {
    const int q = e / Ns, r = e % Ns;
    A.G[r + (size_t)Ns * A.fsites[4 * b0 + q]] = sC[e];
  }
  PROF_ACC(4);
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
    zd *base = reinterpret_cast<zd *>(ps_lds + sizeof(ps_helper_hdr));
    const int ncol_max = 4 * A.bph;
    zd *sv = base;
    zd *sy = sv + ncol_max * PS_BS;
    zd *sC = sy + ncol_max * PS_BS;
    ps_helper(A, Hh, sC, sv, sy);
  }
  // the last workgroup to finish clears the sync words for the next launch
  __syncthreads();
  if (threadIdx.x == 0) {
    ps_release();
    const unsigned long long done = atomicAdd(&A.sync[1], 1ull);
    if (done == gridDim.x - 1) {
      ps_st(&A.sync[0], 0ull);
      ps_st(&A.sync[2], 0ull);
      for (int k = -1; k < A.nblk; ++k) ps_st(ps_arr(A, k), 0ull);
      ps_st(&A.sync[1], 0ull);
    }
  }
}

extern "C" size_t pqmc_pass_lds(int Ns, int bph) {
  const size_t lead = sizeof(ps_leader_lds);
  const size_t help = sizeof(ps_helper_hdr) + sizeof(zd) * (size_t)4 * bph * (2 * PS_BS + Ns);
  return lead > help ? lead : help;
}
extern "C" size_t pqmc_pass_rec_bytes(void) { return sizeof(ps_rec); }

extern "C" hipError_t pqmc_launch_pass(const ps_args *a, hipStream_t s) {
  const size_t lds = pqmc_pass_lds(a->md.Ns, a->bph);
  static size_t attr = 0;
  if (lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_family_pass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
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
