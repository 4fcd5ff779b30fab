// Local phonon-field update of one bond family at one time slice
// (update_phonon_field, src/mod_update.f90:13-201) as a delayed rank-4k
// Sherman-Morrison-Woodbury chain.
//
// G0 (the Green's function at the start of the family pass) is read-only.
// The pass keeps the accumulated change as G = G0 + U V with V = VT^T,
// where accepted bond j contributes four columns:
//   U[:, 4j+a'] = (G[:, S_j] X_j)[:, a'],   VT[:, 4j+c] = ((I - G)[S_j, :])^T
// evaluated with the updates of bonds accepted before j (the reference's
// immediate update :152-187, written in delayed form).  U and VT are stored
// row-major with leading dimension Ns, rows in family order (row 4b+a is
// site S_b[a]), so one row's history is contiguous.
//
// The family's bonds are split into P contiguous chunks, one workgroup each.
// The owner of the current chunk decides its bonds in order, in one wave,
// from an LDS-resident copy of the 4x4 blocks G[S_k, S_k] of its undecided
// bonds; the whole workgroup joins only when a bond is accepted.  Every
// accept is published as a record (bond, X); each workgroup applies every
// record to its own rows, so when a chunk's turn comes its blocks are
// current.  The chunk owner applies a record first to the rows of its still
// undecided bonds and fills in its earlier rows after its chunk is done.
//
// Accept test: the reference rejects when prob < u with prob =
// min(1, exp(min(10, w))), w = dS + ncopy*(1+TR)*log|det(I + G_SS Delta)|
// (cal_prob, :475-500).  Since log u < 0 this is exactly w < log u, i.e.
// |det|^2 < exp(2 (log u - dS) / (ncopy*(1+TR))); the right-hand side does
// not depend on G and is computed for all bonds up front, so the serial
// decision needs no transcendental function.
#include "pqmc_common.h"

#define UPD_THREADS 512
#define UPD_MAXCHUNK 128
#define UPD_SPIN_LIMIT (1u << 26)
#define UPD_SUB 16  // lanes per row in the history sums

struct upd_rec {
  int p;
  int pad[3];
  zd relec;
  zd X[16];
};

struct upd_args {
  model_t md;
  int f;      // family, 0-based
  int t;      // time slice, 1-based
  int P;      // workgroups
  int chunk;  // bonds per workgroup
  const int *fsites;   // [npla*4] this family, family order
  const int *fbfi;     // [npla]
  const zd *fph;       // [npla*4] BC_phases(1:2,3:4) for this family
  const int *fcross;   // [npla]
  double *x;           // boson_field(Nb, Lt)
  ekv_t *ekv;          // table, entry ((t-1)*4 + f)*npla + b
  const float *u;      // 2*npla ran2 draws: (ran_sysm source, rands) per bond
  const zd *G;         // Ns x Ns
  zd *U, *VT;          // row-major Ns x Ns each
  uint8_t *acc;        // [npla]
  upd_rec *recs;       // [npla]
  unsigned long long *sync;  // [0] accept log word, [1] decided word: (epoch<<32)|value
  unsigned epoch;
  int *rank_out;
  double *lncw_acc;
  unsigned long long *stats;  // [0] accepted, [1] rejected, [2] sync timeouts
  unsigned long long *prof;   // optional [P][8] wall-clock accumulators (null: off)
};

static __device__ __forceinline__ int word_val(unsigned long long w, unsigned epoch) {
  return ((unsigned)(w >> 32) == epoch) ? (int)(unsigned)(w & 0xffffffffu) : 0;
}
static __device__ __forceinline__ unsigned long long ld_acq(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ unsigned long long ld_rlx(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait (thread 0) until the accept log holds more than `processed` records
// or the decided word reaches `dec_target`.  Polls with relaxed loads, which
// leave the caches alone, and issues one acquire fence once something
// changed.  Returns false on timeout.
static __device__ bool wait_log(const upd_args &A, int processed, int dec_target, int *cnt_out, int *dec_out) {
  for (unsigned spins = 0; spins < UPD_SPIN_LIMIT; ++spins) {
    const int dec = word_val(ld_rlx(&A.sync[1]), A.epoch);
    const int cnt = word_val(ld_rlx(&A.sync[0]), A.epoch);
    if (cnt > processed || dec >= dec_target) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // re-read after the fence: the decided word first, then the count
      *dec_out = word_val(ld_rlx(&A.sync[1]), A.epoch);
      *cnt_out = word_val(ld_rlx(&A.sync[0]), A.epoch);
      return true;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return false;
}
static __device__ __forceinline__ void st_rel(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Optional phase timers: prof[w*8 + k] += wall-clock ticks of phase k
// (0 set-up, 1 wait before own chunk, 2 decisions, 3 own-row updates in
// the chunk, 4 fill-in of earlier rows, 5 later chunks' accepts).
#define PROF_T0() unsigned long long _pt = A.prof ? wall_clock64() : 0ull
#define PROF_ACC(k)                                                         \
  do {                                                                      \
    if (A.prof && threadIdx.x == 0) {                                       \
      const unsigned long long _now = wall_clock64();                       \
      A.prof[blockIdx.x * 8 + (k)] += _now - _pt;                           \
      _pt = _now;                                                           \
    }                                                                       \
  } while (0)

// modi of src/mod_matrixlib.f90:66-77
static __device__ __forceinline__ int modi_d(int n, int p) {
  int r = n % p;
  if (n < 0) return p + r;
  if (r == 0) return p;
  return r;
}

static __device__ __forceinline__ zd shfl_z(zd v, int src) {
  return zmk(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}
static __device__ __forceinline__ zd shfl_xor_z(zd v, int m) {
  return zmk(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
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

// Apply accept record j (bond p, rank before = 4j) to family-order rows
// [lo, hi) of U/VT (row-major, ld Ns); update the LDS blocks sG of this
// workgroup's bonds with index >= kmin (bond b maps to sG[b - b0]).
// sSh holds 8 * Ns complex: rows 4p..4p+3 of VT, then of U.
// All global loads of a pass (own history rows, G entries, the staged rows
// of bond p) are issued before the first use so their latencies overlap.
// No fence here: the publishing workgroup fences once before its release.
static __device__ void apply_record(const upd_args &A, int j, int lo, int hi, int b0, int kmin, zd *sG,
                                    zd *sSh, const zd *X) {
  const int Ns = A.md.Ns;
  const int tid = threadIdx.x;
  const int p = A.recs[j].p;
  const int rj = 4 * j;
  int Sp[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) Sp[a] = A.fsites[4 * p + a];
  const zd *shV = sSh, *shU = sSh + 4 * Ns;
  const int rows_per_pass = UPD_THREADS / UPD_SUB;
  bool staged = false;
  for (int base = lo; base < hi; base += rows_per_pass) {
    const int i = base + tid / UPD_SUB;
    const int sub = tid % UPD_SUB;
    const bool active = i < hi;
    const int si = active ? A.fsites[i] : 0;
    // first batch of this row's history + the G entries, in flight together
    zd ub[8], vb[8], g_col[4], g_row[4];
    const zd *urow = A.U + (size_t)(active ? i : 0) * Ns;
    const zd *vrow = A.VT + (size_t)(active ? i : 0) * Ns;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = sub + k * UPD_SUB;
      const bool ok = active && q < rj;
      ub[k] = ok ? urow[q] : zmk(0.0, 0.0);
      vb[k] = ok ? vrow[q] : zmk(0.0, 0.0);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      g_col[a] = active ? A.G[si + (size_t)Ns * Sp[a]] : zmk(0.0, 0.0);
      g_row[a] = active ? A.G[Sp[a] + (size_t)Ns * si] : zmk(0.0, 0.0);
    }
    if (!staged) {
      // rows 4p..4p+3 of VT and U, columns 0..rj-1 (contiguous)
      for (int e = tid; e < 8 * rj; e += UPD_THREADS) {
        const int row = e / rj, q = e % rj;
        sSh[row * Ns + q] =
            (row < 4) ? A.VT[(size_t)(4 * p + row) * Ns + q] : A.U[(size_t)(4 * p + row - 4) * Ns + q];
      }
      __syncthreads();
      staged = true;
    }
    zd cs[4], vs[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) { cs[a] = zmk(0.0, 0.0); vs[a] = zmk(0.0, 0.0); }
    for (int q0 = 0;;) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int q = q0 + sub + k * UPD_SUB;
        if (q < rj) {
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            cs[a] = zadd(cs[a], zmul(ub[k], shV[a * Ns + q]));
            vs[a] = zadd(vs[a], zmul(shU[a * Ns + q], vb[k]));
          }
        }
      }
      q0 += 8 * UPD_SUB;
      if (q0 >= rj) break;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int q = q0 + sub + k * UPD_SUB;
        const bool ok = active && q < rj;
        ub[k] = ok ? urow[q] : zmk(0.0, 0.0);
        vb[k] = ok ? vrow[q] : zmk(0.0, 0.0);
      }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
      for (int o = 1; o < UPD_SUB; o <<= 1) {
        cs[a] = zadd(cs[a], shfl_xor_z(cs[a], o));
        vs[a] = zadd(vs[a], shfl_xor_z(vs[a], o));
      }
    }
    zd un[4], vn[4];
    {
      zd col[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        col[a] = zadd(g_col[a], cs[a]);
        vn[a] = zsub(zsub(zmk(si == Sp[a] ? 1.0 : 0.0, 0.0), g_row[a]), vs[a]);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        zd s = zmk(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < 4; ++a) s = zadd(s, zmul(col[a], X[a + 4 * c]));
        un[c] = s;
      }
    }
    if (active && sub < 4) {
      A.U[(size_t)i * Ns + rj + sub] = un[sub];
      A.VT[(size_t)i * Ns + rj + sub] = vn[sub];
    }
    // block of bond k = i/4: its 4 rows are the 64 lanes of one wave; lane
    // (row a, sub c < 4) adds sum_a' un[row a][a'] vn[row c][a'] to sG_k(a, c)
    const int src = (sub & 3) * UPD_SUB;  // a lane holding row c = sub & 3
    zd upd = zmk(0.0, 0.0);
#pragma unroll
    for (int a2 = 0; a2 < 4; ++a2) upd = zadd(upd, zmul(un[a2], shfl_z(vn[a2], src)));
    if (active && sub < 4) {
      const int k = i / 4, a = i & 3;
      if (k >= kmin) {
        zd &gref = sG[(k - b0) * 16 + a + 4 * sub];
        gref = zadd(gref, upd);
      }
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(UPD_THREADS) void k_update_family(upd_args A) {
  const int w = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const model_t &md = A.md;
  const int npla = md.npla, Ns = md.Ns, Nb = md.Nb;
  const int b0 = w * A.chunk;
  const int b1 = min(npla, b0 + A.chunk);
  extern __shared__ zd sSh[];  // 8 * Ns
  __shared__ zd sG[UPD_MAXCHUNK * 16];
  __shared__ zd sDelta[UPD_MAXCHUNK * 16];
  __shared__ double sThr[UPD_MAXCHUNK];
  __shared__ double sXnew[UPD_MAXCHUNK];
  __shared__ ekv_t sEnew[UPD_MAXCHUNK];
  __shared__ zd sX[16];
  __shared__ zd sM[16];
  __shared__ int sInt[4];
  if (b0 >= b1) return;
  PROF_T0();
  const int nmine = b1 - b0;
  const size_t tbase = (size_t)Nb * (A.t - 1);
  const double kpow = (double)md.ncopy * (md.tr_weight ? 2.0 : 1.0);

  // ---- per-bond proposal data (independent of the accept sequence) ----
  if (tid < nmine) {
    const int b = b0 + tid;
    const int fld = A.fbfi[b];
    const double xold = A.x[fld + tbase];
    // ran_sysm = ran2*2-1 in REAL(SP) (src/mod_ranstate.f90:233-236); jump in DP
    const float rs = A.u[2 * b] * 2.0f - 1.0f;
    const double jump = md.jump * (double)rs;
    // cal_energy_difference (src/mod_update.f90:203-228)
    const double be = A.x[fld + (size_t)Nb * (modi_d(A.t - 1, md.Lt) - 1)];
    const double af = A.x[fld + (size_t)Nb * (modi_d(A.t + 1, md.Lt) - 1)];
    const double nw = xold + jump;
    const double dt2 = md.dtau * md.dtau;
    const double e1 = 0.5 * md.M * ((be - xold) * (be - xold) + (af - xold) * (af - xold)) / dt2 +
                      0.5 * md.D * (xold * xold) - md.biased * xold;
    const double e2 = 0.5 * md.M * ((be - nw) * (be - nw) + (af - nw) * (af - nw)) / dt2 +
                      0.5 * md.D * (nw * nw) - md.biased * nw;
    const double dS = -md.dtau * (e2 - e1);
    sThr[tid] = exp(2.0 * (log((double)A.u[2 * b + 1]) - dS) / kpow);
    sXnew[tid] = nw;
    // calculate_Delta (:230-245): Delta = expKV_inv(old) expKV(new) - I
    ekv_t en;
    expkv_compact(md, A.f, nw, A.fcross[b] ? &A.fph[4 * b] : nullptr, en);
    sEnew[tid] = en;
    const ekv_t eo = A.ekv[((size_t)(A.t - 1) * 4 + A.f) * npla + b];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int ii = 0; ii < 4; ++ii) {
        zd s = zmk(0.0, 0.0);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) s = zadd(s, zmul(ekv_entry(eo, true, ii, kk), ekv_entry(en, false, kk, jj)));
        if (ii == jj) s.x -= 1.0;
        sDelta[tid * 16 + ii + 4 * jj] = s;
      }
  }
  for (int e = tid; e < nmine * 16; e += UPD_THREADS) {
    const int k = e / 16, q = e % 16, a = q & 3, c = q >> 2;
    const int *S = &A.fsites[4 * (b0 + k)];
    sG[e] = A.G[S[a] + (size_t)Ns * S[c]];
  }
  __syncthreads();
  PROF_ACC(0);

  int processed = 0;
  bool timed_out = false;
  const int lo_all = 4 * b0, hi_all = 4 * b1;

  // ---- phase A: apply the accepts of earlier chunks as they appear ----
  if (w > 0) {
    int cnt = 0, dec = 0;
    while (true) {
      if (processed >= cnt && dec < b0) {
        if (tid == 0) {
          int c2 = 0, d2 = 0;
          const bool ok = wait_log(A, processed, b0, &c2, &d2);
          sInt[0] = c2;
          sInt[1] = d2;
          sInt[3] = ok;
        }
        __syncthreads();
        cnt = sInt[0];
        dec = sInt[1];
        const bool ok = sInt[3];
        __syncthreads();
        if (!ok) { timed_out = true; break; }
      }
      if (processed < cnt) {
        if (tid < 16) sX[tid] = A.recs[processed].X[tid];
        __syncthreads();
        apply_record(A, processed, lo_all, hi_all, b0, b0, sG, sSh, sX);
        ++processed;
        continue;
      }
      if (dec >= b0) break;
    }
  }
  PROF_ACC(1);

  // ---- phase B: decide this chunk's bonds ----
  const int first_own = processed;
  int p = b0;
  int nrej = 0;
  while (p < b1 && !timed_out) {
    // wave 0 decides bonds p, p+1, ... until one is accepted
    if (wave == 0) {
      int accepted = -1;
      for (int q = p; q < b1; ++q) {
        const int k = q - b0;
        // M = I + G_SS Delta (src/mod_update.f90:97-100), one entry per lane
        if (lane < 16) {
          const int a = lane & 3, c = lane >> 2;
          zd mval = zmk(0.0, 0.0);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) mval = zadd(mval, zmul(sG[k * 16 + a + 4 * kk], sDelta[k * 16 + kk + 4 * c]));
          if (a == c) mval.x += 1.0;
          sM[lane] = mval;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // cofactors C(i,j) = (-1)^(i+j) det(minor_ij); lane i + 4j
        zd cof = zmk(0.0, 0.0);
        if (lane < 16) {
          const int ri = lane & 3, cj = lane >> 2;
          cof = minor3(sM, ri, cj);
          if ((ri + cj) & 1) cof = zneg(cof);
        }
        // det = sum_j M(0,j) C(0,j): lanes 0,4,8,12
        zd term = (lane < 16 && (lane & 3) == 0) ? zmul(sM[lane], cof) : zmk(0.0, 0.0);
        term = zadd(term, shfl_xor_z(term, 4));
        term = zadd(term, shfl_xor_z(term, 8));
        const zd det = shfl_z(term, 0);
        // reject iff |det|^2 < exp(2 (log u - dS) / k)
        const bool acc = !((det.x * det.x + det.y * det.y) < sThr[k]);
        if (lane == 0) A.acc[q] = (uint8_t)acc;
        if (!acc) {
          ++nrej;
          continue;
        }
        // X = Delta (I + G_SS Delta)^-1 with the inverse = adj(M)/det:
        // inv(k, c) = C(c, k) / det, held by lane k + 4c after the exchange
        const int ia = lane & 3, ic = (lane >> 2) & 3;
        const zd cT = shfl_z(cof, ic + 4 * ia);
        const zd inv = zdiv(cT, det);
        zd xv = zmk(0.0, 0.0);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) xv = zadd(xv, zmul(sDelta[k * 16 + ia + 4 * kk], shfl_z(inv, kk + 4 * ic)));
        if (lane < 16) {
          sX[lane] = xv;
          A.recs[processed].X[lane] = xv;
        }
        if (lane == 0) {
          A.recs[processed].p = q;
          zd rel = det;
          if (md.tr_weight) rel = zmul(rel, zconj(rel));
          zd rp = rel;
          for (int c = 1; c < md.ncopy; ++c) rp = zmul(rp, rel);
          A.recs[processed].relec = rp;
        }
        accepted = q;
        break;
      }
      if (lane == 0) sInt[2] = accepted;
    }
    // every wave: make its earlier U/VT rows (and wave 0 the record)
    // visible device-wide before the release below
    __threadfence();
    __syncthreads();
    const int accepted = sInt[2];
    if (accepted < 0) break;
    if (tid == 0 && A.P > 1) st_rel(&A.sync[0], ((unsigned long long)A.epoch << 32) | (unsigned)(processed + 1));
    PROF_ACC(2);
    apply_record(A, processed, 4 * accepted, hi_all, b0, accepted + 1, sG, sSh, sX);
    ++processed;
    p = accepted + 1;
    PROF_ACC(3);
  }
  const int last_own = processed;
  __threadfence();
  __syncthreads();
  if (tid == 0) st_rel(&A.sync[1], ((unsigned long long)A.epoch << 32) | (unsigned)b1);
  PROF_ACC(2);

  // ---- phase C: accepted bonds' fields and exponentials, ln_cw, earlier rows ----
  for (int j = first_own + tid; j < last_own; j += UPD_THREADS) {
    const int q = A.recs[j].p;
    const int k = q - b0;
    A.x[A.fbfi[q] + tbase] = sXnew[k];
    A.ekv[((size_t)(A.t - 1) * 4 + A.f) * npla + q] = sEnew[k];
    // ln_cw += log(real(R_elec)) (src/mod_update.f90:142)
    atomicAdd(A.lncw_acc, log(A.recs[j].relec.x));
  }
  if (tid == 0) {
    atomicAdd(&A.stats[0], (unsigned long long)(last_own - first_own));
    atomicAdd(&A.stats[1], (unsigned long long)nrej);
  }
  for (int j = first_own; j < last_own; ++j) {
    const int q = A.recs[j].p;
    if (q == b0) continue;
    if (tid < 16) sX[tid] = A.recs[j].X[tid];
    __syncthreads();
    apply_record(A, j, lo_all, 4 * q, b0, npla, sG, sSh, sX);
  }
  PROF_ACC(4);

  // ---- phase D: accepts of later chunks, all rows of this chunk ----
  if (b1 < npla) {
    int cnt = processed, dec = b1;
    while (!timed_out) {
      if (processed >= cnt && dec < npla) {
        if (tid == 0) {
          int c2 = 0, d2 = 0;
          const bool ok = wait_log(A, processed, npla, &c2, &d2);
          sInt[0] = c2;
          sInt[1] = d2;
          sInt[3] = ok;
        }
        __syncthreads();
        cnt = sInt[0];
        dec = sInt[1];
        const bool ok = sInt[3];
        __syncthreads();
        if (!ok) { timed_out = true; break; }
      }
      if (processed < cnt) {
        if (tid < 16) sX[tid] = A.recs[processed].X[tid];
        __syncthreads();
        apply_record(A, processed, lo_all, hi_all, b0, npla, sG, sSh, sX);
        ++processed;
        continue;
      }
      if (dec >= npla) break;
    }
  } else if (tid == 0) {
    *A.rank_out = 4 * processed;
  }
  PROF_ACC(5);
  if (timed_out && tid == 0) atomicAdd(&A.stats[2], 1ull);
}

// Host-side launcher.
extern "C" hipError_t pqmc_launch_update(const upd_args *a, hipStream_t s) {
  const size_t lds = sizeof(zd) * 8 * (size_t)a->md.Ns;
  static size_t attr = 48 * 1024;
  if (lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_update_family, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL(k_update_family, dim3(a->P), dim3(UPD_THREADS), lds, s, *a);
  return hipGetLastError();
}
