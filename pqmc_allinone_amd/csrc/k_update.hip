// Local phonon-field update of one bond family at one time slice
// (update_phonon_field, src/mod_update.f90:13-201) as a delayed rank-4k
// Sherman-Morrison-Woodbury chain.
//
// G0 (the Green's function at the start of the family pass) is read-only.
// The pass keeps the accumulated change as G = G0 + U V with V = VT^T,
// where accepted bond j contributes four columns:
//   U[:, 4j+a'] = (G[:, S_j] X_j)[:, a'],   VT[:, 4j+c] = ((I - G)[S_j, :])^T
// evaluated with the updates of bonds accepted before j (the reference's
// immediate update :152-187, written in delayed form).  Rows of U and VT
// are stored in family order (row 4b+a is site S_b[a]).
//
// The family's bonds are split into P contiguous chunks, one workgroup each.
// The owner of the current chunk decides its bonds in order from an
// LDS-resident copy of the 4x4 blocks G[S_k, S_k] of its undecided bonds.
// Every accept is published as a record (bond, X); each workgroup applies
// every record to its own rows, so when a chunk's turn comes its blocks are
// current.  The chunk owner applies a record first to the rows of its still
// undecided bonds and fills in its earlier rows after its chunk is done.
#include "pqmc_common.h"

#define UPD_THREADS 512
#define UPD_MAXCHUNK 64
#define UPD_KCH 128
#define UPD_SPIN_LIMIT (1u << 26)

struct upd_rec {
  int p;
  int pad[3];
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
  zd *U, *VT;          // Ns x Ns each
  uint8_t *acc;        // [npla]
  upd_rec *recs;       // [npla]
  unsigned long long *sync;  // [0] accept log word, [1] decided word: (epoch<<32)|value
  unsigned epoch;
  int *rank_out;
  double *lncw_acc;
  unsigned long long *stats;  // [0] accepted, [1] rejected, [2] sync timeouts
};

static __device__ __forceinline__ unsigned long long ld_acq(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void st_rel(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ int word_val(unsigned long long w, unsigned epoch) {
  return ((unsigned)(w >> 32) == epoch) ? (int)(unsigned)(w & 0xffffffffu) : 0;
}

// 4x4 inverse by Gauss-Jordan with partial pivoting on |Re|+|Im| (ZGETRF's
// izamax), used for (I + G_SS Delta)^-1 (src/mod_update.f90:171).  All
// indices are compile-time after unrolling; the pivot row swap is a
// conditional move so nothing lands in scratch memory.
static __device__ __forceinline__ void inv4(zd (&a)[16]) {
  zd inv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) inv[q] = zmk((q % 5) == 0 ? 1.0 : 0.0, 0.0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double best = fabs(a[j + 4 * j].x) + fabs(a[j + 4 * j].y);
    int p = j;
#pragma unroll
    for (int i = j + 1; i < 4; ++i) {
      const double v = fabs(a[i + 4 * j].x) + fabs(a[i + 4 * j].y);
      if (v > best) { best = v; p = i; }
    }
#pragma unroll
    for (int i = j + 1; i < 4; ++i) {
      const bool sw = (p == i);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const zd tj = a[j + 4 * c], ti = a[i + 4 * c];
        a[j + 4 * c] = sw ? ti : tj;
        a[i + 4 * c] = sw ? tj : ti;
        const zd uj = inv[j + 4 * c], ui = inv[i + 4 * c];
        inv[j + 4 * c] = sw ? ui : uj;
        inv[i + 4 * c] = sw ? uj : ui;
      }
    }
    const zd r = zdiv(zmk(1.0, 0.0), a[j + 4 * j]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[j + 4 * c] = zmul(a[j + 4 * c], r);
      inv[j + 4 * c] = zmul(inv[j + 4 * c], r);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i == j) continue;
      const zd fct = a[i + 4 * j];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        a[i + 4 * c] = zsub(a[i + 4 * c], zmul(fct, a[j + 4 * c]));
        inv[i + 4 * c] = zsub(inv[i + 4 * c], zmul(fct, inv[j + 4 * c]));
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = inv[q];
}

// modi of src/mod_matrixlib.f90:66-77
static __device__ __forceinline__ int modi_d(int n, int p) {
  int r = n % p;
  if (n < 0) return p + r;
  if (r == 0) return p;
  return r;
}

// Apply accept record j (bond p, rank before = 4j) to family-order rows
// [lo, hi) of U/VT; update the LDS blocks sG of this workgroup's bonds
// with index >= kmin (bonds b0..b1-1 map to sG[b - b0]).
static __device__ void apply_record(const upd_args &A, int j, int lo, int hi, int b0, int kmin,
                                    zd *sG, zd *sSh, const zd *X) {
  const int Ns = A.md.Ns;
  const int tid = threadIdx.x;
  const upd_rec &rec = A.recs[j];
  const int p = rec.p;
  const int rj = 4 * j;
  int Sp[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) Sp[a] = A.fsites[4 * p + a];
  for (int base = lo; base < hi; base += UPD_THREADS / 4) {
    const int i = base + tid / 4;
    const int sub = tid & 3;
    const bool active = i < hi;
    zd cs[4], vs[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) { cs[a] = zmk(0.0, 0.0); vs[a] = zmk(0.0, 0.0); }
    for (int q0 = 0; q0 < rj; q0 += UPD_KCH) {
      const int nq = min(UPD_KCH, rj - q0);
      __syncthreads();
      for (int e = tid; e < 8 * nq; e += UPD_THREADS) {
        int row = e / nq, q = e % nq;
        // rows 0..3: VT[4p+a, q]; rows 4..7: U[4p+c, q]
        sSh[row * UPD_KCH + q] = (row < 4) ? A.VT[(4 * p + row) + (size_t)Ns * (q0 + q)]
                                           : A.U[(4 * p + row - 4) + (size_t)Ns * (q0 + q)];
      }
      __syncthreads();
      if (active) {
#pragma unroll 2
        for (int q = sub; q < nq; q += 4) {
          zd uiq = A.U[i + (size_t)Ns * (q0 + q)];
          zd viq = A.VT[i + (size_t)Ns * (q0 + q)];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            cs[a] = zadd(cs[a], zmul(uiq, sSh[a * UPD_KCH + q]));
            vs[a] = zadd(vs[a], zmul(sSh[(4 + a) * UPD_KCH + q], viq));
          }
        }
      }
    }
    // reduce the four q-strided partial sums of this row
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
      for (int o = 1; o < 4; o <<= 1) {
        cs[a].x += __shfl_xor(cs[a].x, o, 64);
        cs[a].y += __shfl_xor(cs[a].y, o, 64);
        vs[a].x += __shfl_xor(vs[a].x, o, 64);
        vs[a].y += __shfl_xor(vs[a].y, o, 64);
      }
    }
    zd un[4], vn[4];
    if (active) {
      const int si = A.fsites[i];
      zd col[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        col[a] = zadd(A.G[si + (size_t)Ns * Sp[a]], cs[a]);
        zd g = A.G[Sp[a] + (size_t)Ns * si];
        vn[a] = zsub(zsub(zmk(si == Sp[a] ? 1.0 : 0.0, 0.0), g), vs[a]);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        zd s = zmk(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < 4; ++a) s = zadd(s, zmul(col[a], X[a + 4 * c]));
        un[c] = s;
      }
      if (sub == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          A.U[i + (size_t)Ns * (rj + c)] = un[c];
          A.VT[i + (size_t)Ns * (rj + c)] = vn[c];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) { un[c] = zmk(0.0, 0.0); vn[c] = zmk(0.0, 0.0); }
    }
    // block update of bond k = i/4 (rows 4k..4k+3 are 16 consecutive lanes):
    // lane (row a, sub c) adds sum_a' un[row a][a'] * vn[row c][a'] to sG_k(a, c)
    const int lane = tid & 63;
    const int src_c = (lane & ~15) + 4 * sub;  // lane holding row c = sub (sub 0)
    zd upd = zmk(0.0, 0.0);
#pragma unroll
    for (int a2 = 0; a2 < 4; ++a2) {
      zd vc;
      vc.x = __shfl(vn[a2].x, src_c, 64);
      vc.y = __shfl(vn[a2].y, src_c, 64);
      upd = zadd(upd, zmul(un[a2], vc));
    }
    if (active) {
      const int k = i / 4, a = i & 3;
      if (k >= kmin && k >= b0 && k < b0 + UPD_MAXCHUNK) {
        zd &gref = sG[(k - b0) * 16 + a + 4 * sub];
        gref = zadd(gref, upd);
      }
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(UPD_THREADS) void k_update_family(upd_args A) {
  const int w = blockIdx.x;
  const int tid = threadIdx.x;
  const model_t &md = A.md;
  const int npla = md.npla, Ns = md.Ns, Nb = md.Nb;
  const int b0 = w * A.chunk;
  const int b1 = min(npla, b0 + A.chunk);
  __shared__ zd sG[UPD_MAXCHUNK * 16];
  __shared__ zd sDelta[UPD_MAXCHUNK * 16];
  __shared__ double sdS[UPD_MAXCHUNK];
  __shared__ double sxnew[UPD_MAXCHUNK];
  __shared__ zd sSh[8 * UPD_KCH];
  __shared__ zd sX[16];
  __shared__ zd sM[16];
  __shared__ int sInt[4];
  __shared__ double sDbl[4];
  if (b0 >= b1) return;
  const int nmine = b1 - b0;
  const size_t tbase = (size_t)Nb * (A.t - 1);

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
    sdS[tid] = -md.dtau * (e2 - e1);
    sxnew[tid] = nw;
    // calculate_Delta (:230-245): Delta = expKV_inv(old) expKV(new) - I
    ekv_t en;
    expkv_compact(md, A.f, nw, A.fcross[b] ? &A.fph[4 * b] : nullptr, en);
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

  int processed = 0;
  unsigned spins = 0;
  bool timed_out = false;
  const int lo_all = 4 * b0, hi_all = 4 * b1;

  // ---- phase A: apply the accepts of earlier chunks as they appear ----
  if (w > 0) {
    while (true) {
      if (tid == 0) {
        int cnt = word_val(ld_acq(&A.sync[0]), A.epoch);
        int dec = word_val(ld_acq(&A.sync[1]), A.epoch);
        sInt[0] = cnt;
        sInt[1] = dec;
      }
      __syncthreads();
      const int cnt = sInt[0], dec = sInt[1];
      __syncthreads();
      if (processed < cnt) {
        if (tid < 16) sX[tid] = A.recs[processed].X[tid];
        __syncthreads();
        apply_record(A, processed, lo_all, hi_all, b0, b0, sG, sSh, sX);
        ++processed;
        spins = 0;
        continue;
      }
      if (dec >= b0) break;
      if (++spins > UPD_SPIN_LIMIT) { timed_out = true; break; }
      __builtin_amdgcn_s_sleep(2);
    }
  }

  // ---- phase B: decide this chunk's bonds ----
  const int first_own = processed;
  for (int p = b0; p < b1 && !timed_out; ++p) {
    const int k = p - b0;
    if (tid < 16) {
      // M = I + G_SS Delta (src/mod_update.f90:97-100)
      const int a = tid & 3, c = tid >> 2;
      zd s = zmk(0.0, 0.0);
      for (int kk = 0; kk < 4; ++kk) s = zadd(s, zmul(sG[k * 16 + a + 4 * kk], sDelta[k * 16 + kk + 4 * c]));
      if (a == c) s.x += 1.0;
      sM[tid] = s;
    }
    __syncthreads();
    if (tid < 64) {
      // det(M) by Laplace expansion over the 2x2 minors of rows (1,2)
      const int lane = tid;
      zd term = zmk(0.0, 0.0);
      if (lane < 6) {
        // complementary 2x2 minors of rows (1,2) and (3,4); pairs in
        // lexicographic order with Laplace signs + - + + - +
        int j1, j2, k1, k2;
        double sg;
        switch (lane) {
          case 0: j1 = 0; j2 = 1; k1 = 2; k2 = 3; sg = 1.0; break;
          case 1: j1 = 0; j2 = 2; k1 = 1; k2 = 3; sg = -1.0; break;
          case 2: j1 = 0; j2 = 3; k1 = 1; k2 = 2; sg = 1.0; break;
          case 3: j1 = 1; j2 = 2; k1 = 0; k2 = 3; sg = 1.0; break;
          case 4: j1 = 1; j2 = 3; k1 = 0; k2 = 2; sg = -1.0; break;
          default: j1 = 2; j2 = 3; k1 = 0; k2 = 1; sg = 1.0; break;
        }
        zd top = zsub(zmul(sM[0 + 4 * j1], sM[1 + 4 * j2]), zmul(sM[0 + 4 * j2], sM[1 + 4 * j1]));
        zd bot = zsub(zmul(sM[2 + 4 * k1], sM[3 + 4 * k2]), zmul(sM[2 + 4 * k2], sM[3 + 4 * k1]));
        term = zscale(zmul(top, bot), sg);
      }
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) {
        term.x += __shfl_xor(term.x, o, 64);
        term.y += __shfl_xor(term.y, o, 64);
      }
      if (lane == 0) {
        zd relec = term;
        if (md.tr_weight) relec = zmul(relec, zconj(relec));
        zd rp = relec;
        for (int c = 1; c < md.ncopy; ++c) rp = zmul(rp, relec);
        relec = rp;
        // cal_prob (:475-500), Metropolis: prob = min(1, exp(min(10, Re w)))
        const double wre = sdS[k] + log(hypot(relec.x, relec.y));
        const double prob = fmin(1.0, exp(fmin(10.0, wre)));
        const int accept = !(prob < (double)A.u[2 * p + 1]);
        sInt[2] = accept;
        sDbl[0] = relec.x;
      }
    }
    __syncthreads();
    const int accept = sInt[2];
    if (tid == 0) A.acc[p] = (uint8_t)accept;
    if (!accept) {
      if (tid == 0) atomicAdd(&A.stats[1], 1ull);
      continue;
    }
    if (tid == 0) {
      // X = Delta (I + G_SS Delta)^-1 (src/mod_update.f90:166-173)
      zd m[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) m[q] = sM[q];
      inv4(m);
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          zd s = zmk(0.0, 0.0);
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) s = zadd(s, zmul(sDelta[k * 16 + a + 4 * kk], m[kk + 4 * c]));
          sX[a + 4 * c] = s;
        }
      atomicAdd(&A.stats[0], 1ull);
      // ln_cw += log(real(R_elec)) (:142)
      atomicAdd(A.lncw_acc, log(sDbl[0]));
    }
    if (tid == 1) {
      // accepted: new field and expKV (:126-127)
      const int fld = A.fbfi[p];
      A.x[fld + tbase] = sxnew[k];
      ekv_t en;
      expkv_compact(md, A.f, sxnew[k], A.fcross[p] ? &A.fph[4 * p] : nullptr, en);
      A.ekv[((size_t)(A.t - 1) * 4 + A.f) * npla + p] = en;
    }
    __syncthreads();
    if (tid < 16) A.recs[processed].X[tid] = sX[tid];
    if (tid == 0) A.recs[processed].p = p;
    __threadfence();
    __syncthreads();
    if (tid == 0) st_rel(&A.sync[0], ((unsigned long long)A.epoch << 32) | (unsigned)(processed + 1));
    // rows of the still undecided bonds first
    apply_record(A, processed, 4 * p, hi_all, b0, p + 1, sG, sSh, sX);
    ++processed;
  }
  const int last_own = processed;
  __threadfence();
  __syncthreads();
  if (tid == 0) st_rel(&A.sync[1], ((unsigned long long)A.epoch << 32) | (unsigned)b1);

  // ---- phase C: earlier rows of this chunk for its own accepts ----
  for (int j = first_own; j < last_own; ++j) {
    const int p = A.recs[j].p;
    if (p == b0) continue;
    if (tid < 16) sX[tid] = A.recs[j].X[tid];
    __syncthreads();
    apply_record(A, j, lo_all, 4 * p, b0, npla, sG, sSh, sX);
  }

  // ---- phase D: accepts of later chunks, all rows of this chunk ----
  if (b1 < npla) {
    spins = 0;
    while (!timed_out) {
      if (tid == 0) {
        sInt[0] = word_val(ld_acq(&A.sync[0]), A.epoch);
        sInt[1] = word_val(ld_acq(&A.sync[1]), A.epoch);
      }
      __syncthreads();
      const int cnt = sInt[0], dec = sInt[1];
      __syncthreads();
      if (processed < cnt) {
        if (tid < 16) sX[tid] = A.recs[processed].X[tid];
        __syncthreads();
        apply_record(A, processed, lo_all, hi_all, b0, npla, sG, sSh, sX);
        ++processed;
        spins = 0;
        continue;
      }
      if (dec >= npla) break;
      if (++spins > UPD_SPIN_LIMIT) { timed_out = true; break; }
      __builtin_amdgcn_s_sleep(2);
    }
  } else if (tid == 0) {
    *A.rank_out = 4 * processed;
  }
  if (timed_out && tid == 0) atomicAdd(&A.stats[2], 1ull);
}

// Host-side launcher.
extern "C" hipError_t pqmc_launch_update(const upd_args *a, hipStream_t s) {
  hipLaunchKernelGGL(k_update_family, dim3(a->P), dim3(UPD_THREADS), 0, s, *a);
  return hipGetLastError();
}
