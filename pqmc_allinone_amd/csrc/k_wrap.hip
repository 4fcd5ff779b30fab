// Checkerboard propagation kernels:
//  * k_flush_wrap: G <- W(G + U V), the delayed rank-r flush of a family
//    pass fused with the next two-way wrap (left_evolve / right_evolve with
//    two_way, src/mod_evolution.f90:236-328).  The rank-r product runs on
//    FP64 MFMA; the bond exponentials are applied per 4x4 block in LDS.
//  * k_propagate: one-way B-string propagation X(S,:) <- E X(S,:) over a run
//    of (slice, family) stages (init_g_T0 / get_g_scratch_T0 loops,
//    src/mod_evolution.f90:88-102, 137-162).
//  * k_ekv_build: the expKV table for all (slice, family, bond)
//    (init_expKV, src/mod_phonon_field.f90:423-449).
#include "pqmc_common.h"

PQ_DEV zd bond_field_grad(const model_t &md, int f, double xv, const zd *p, const ekv_t &e, const zd g[16]);

#define FW_TB 8   // bonds per tile side
#define FW_TS 32  // sites per tile side
#define FW_KC 16  // rank chunk



__global__ __launch_bounds__(256) void k_flush_wrap(fw_args A) {
  const int npla = A.md.npla, Ns = A.md.Ns;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int bi0 = blockIdx.x * FW_TB, bj0 = blockIdx.y * FW_TB;
  const int nbr = min(FW_TB, npla - bi0), nbc = min(FW_TB, npla - bj0);
  const int nr = 4 * nbr, nc = 4 * nbc;
  const int r = A.rank ? *A.rank : 0;
  if (r == 0 && A.wrap == 0) return;
  __shared__ int rs[FW_TS], cs[FW_TS], rpos[FW_TS], cpos[FW_TS];
  __shared__ double sAr[FW_KC][FW_TS], sAi[FW_KC][FW_TS], sBr[FW_KC][FW_TS], sBi[FW_KC][FW_TS];
  __shared__ zd sT[FW_TS][FW_TS + 1];
  if (tid < FW_TS) {
    const int ii = tid;
    rs[ii] = ii < nr ? A.tsites[4 * bi0 + ii] : 0;
    cs[ii] = ii < nc ? A.tsites[4 * bj0 + ii] : 0;
    rpos[ii] = (ii < nr && A.pos_ffl) ? A.pos_ffl[rs[ii]] : 0;
    cpos[ii] = (ii < nc && A.pos_ffl) ? A.pos_ffl[cs[ii]] : 0;
  }
  __syncthreads();
  const int wi = wv >> 1, wj = wv & 1;
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  for (int q0 = 0; q0 < r; q0 += FW_KC) {
    for (int e = tid; e < 2 * FW_TS * FW_KC; e += 256) {
      const int which = e / (FW_TS * FW_KC);
      const int idx = e % (FW_TS * FW_KC);
      const int ii = idx % FW_TS, k = idx / FW_TS;
      const int q = q0 + k;
      zd v = zmk(0.0, 0.0);
      if (which == 0) {
        if (q < r && ii < nr) v = A.U[(size_t)rpos[ii] * Ns + q];
        sAr[k][ii] = v.x;
        sAi[k][ii] = v.y;
      } else {
        if (q < r && ii < nc) v = A.VT[(size_t)cpos[ii] * Ns + q];
        sBr[k][ii] = v.x;
        sBi[k][ii] = v.y;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < FW_KC / 4; ++kk) {
      const int krow = kk * 4 + (lane >> 4);
      const double ar = sAr[krow][wi * 16 + (lane & 15)], ai = sAi[krow][wi * 16 + (lane & 15)];
      const double br = sBr[krow][wj * 16 + (lane & 15)], bi = sBi[krow][wj * 16 + (lane & 15)];
      cr = mfma_f64(ar, br, cr);
      cr = mfma_f64(-ai, bi, cr);
      ci = mfma_f64(ar, bi, ci);
      ci = mfma_f64(ai, br, ci);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = wi * 16 + (lane >> 4) + 4 * i;
    const int col = wj * 16 + (lane & 15);
    sT[row][col] = zmk(cr[i], ci[i]);
  }
  __syncthreads();
  for (int e = tid; e < FW_TS * FW_TS; e += 256) {
    const int ii = e % FW_TS, jj = e / FW_TS;
    if (ii < nr && jj < nc) sT[ii][jj] = zadd(sT[ii][jj], A.G[rs[ii] + (size_t)Ns * cs[jj]]);
  }
  __syncthreads();
  if (A.grad) {  // uniform: a kernel argument
    // a diagonal tile holds each of its bonds' G(S, S) before the wrap
    if (blockIdx.x == blockIdx.y && tid < nbr) {
      const int b = bi0 + tid, fb = A.f_wrap * npla + b;
      const size_t idx = (size_t)A.bfi[fb] + (size_t)A.md.Nb * (A.t_wrap - 1);
      zd g[16];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) g[i + 4 * j] = sT[4 * tid + i][4 * tid + j];
      const zd tr = bond_field_grad(A.md, A.f_wrap, A.x[idx], A.cross[fb] ? &A.ph[4 * fb] : nullptr, A.ekv_w[b], g);
      A.grad[idx] = zadd(A.grad[idx], tr);
    }
    __syncthreads();
  }
  if (A.wrap) {
    const bool left = (A.wrap == 1);
    {  // rows: thread (bond bi, column jj): v = B(:, jj) of block row bi -> E_i v (left) / E_i^-1 v
      const int b = tid / FW_TS, jj = tid % FW_TS;
      if (b < nbr && jj < nc) {
        const ekv_t e = A.ekv_w[bi0 + b];
        zd v[4], y[4];
        for (int a = 0; a < 4; ++a) v[a] = sT[4 * b + a][jj];
        ekv_apply(e, !left, v, y);
        for (int a = 0; a < 4; ++a) sT[4 * b + a][jj] = y[a];
      }
    }
    __syncthreads();
    {  // columns: thread (row ii, bond bj): v = B(ii, :) -> v E_j^-1 (left) / v E_j
      const int ii = tid % FW_TS, b = tid / FW_TS;
      if (b < nbc && ii < nr) {
        const ekv_t e = A.ekv_w[bj0 + b];
        zd v[4], y[4];
        for (int a = 0; a < 4; ++a) v[a] = sT[ii][4 * b + a];
        ekv_apply_right(e, left, v, y);
        for (int a = 0; a < 4; ++a) sT[ii][4 * b + a] = y[a];
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < FW_TS * FW_TS; e += 256) {
    const int ii = e % FW_TS, jj = e / FW_TS;
    if (ii < nr && jj < nc) A.G[rs[ii] + (size_t)Ns * cs[jj]] = sT[ii][jj];
  }
}

extern "C" hipError_t pqmc_launch_flush_wrap(const fw_args *a, hipStream_t s) {
  const int nt = (a->md.npla + FW_TB - 1) / FW_TB;
  hipLaunchKernelGGL(k_flush_wrap, dim3(nt, nt), dim3(256), 0, s, *a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------

struct prop_args {
  model_t md;
  zd *X;               // Ns x ncols, ld Ns
  int ncols;
  const int *sites;    // [4][npla*4]
  const ekv_t *ekv;    // full table
  int t_first;         // first slice (1-based)
  int nslices;
  int backward;        // 0: t ascending, families 1..4; 1: t descending, families 4..1
  int fam_only;        // >= 0: only this family, one stage per slice
};

#define PROP_COLS 4

__global__ __launch_bounds__(1024) void k_propagate(prop_args A) {
  extern __shared__ zd sX[];  // PROP_COLS x Ns
  const int Ns = A.md.Ns, npla = A.md.npla;
  const int c0 = blockIdx.x * PROP_COLS;
  const int ncw = min(PROP_COLS, A.ncols - c0);
  const int tid = threadIdx.x;
  for (int e = tid; e < ncw * Ns; e += blockDim.x) {
    const int c = e / Ns, i = e % Ns;
    sX[c * Ns + i] = A.X[i + (size_t)Ns * (c0 + c)];
  }
  __syncthreads();
  const int c = tid / npla, b = tid % npla;
  const int per = (A.fam_only >= 0) ? 1 : 4;
  const int nst = per * A.nslices;
  for (int s = 0; s < nst; ++s) {
    const int t = A.backward ? A.t_first - s / per : A.t_first + s / per;
    const int f = (A.fam_only >= 0) ? A.fam_only : (A.backward ? 3 - (s % 4) : (s % 4));
    if (c < ncw) {
      const int *S = &A.sites[f * npla * 4 + 4 * b];
      const ekv_t e = A.ekv[((size_t)(t - 1) * 4 + f) * npla + b];
      zd v[4], y[4];
      for (int a = 0; a < 4; ++a) v[a] = sX[c * Ns + S[a]];
      ekv_apply(e, false, v, y);
      for (int a = 0; a < 4; ++a) sX[c * Ns + S[a]] = y[a];
    }
    __syncthreads();
  }
  for (int e = tid; e < ncw * Ns; e += blockDim.x) {
    const int cc = e / Ns, i = e % Ns;
    A.X[i + (size_t)Ns * (c0 + cc)] = sX[cc * Ns + i];
  }
}

extern "C" hipError_t pqmc_launch_propagate(const prop_args *a, hipStream_t s) {
  const int nblk = (a->ncols + PROP_COLS - 1) / PROP_COLS;
  int threads = PROP_COLS * a->md.npla;
  threads = ((threads + 63) / 64) * 64;
  if (threads > 1024) return hipErrorInvalidValue;
  const size_t lds = sizeof(zd) * PROP_COLS * a->md.Ns;
  hipLaunchKernelGGL(k_propagate, dim3(nblk), dim3(threads), lds, s, *a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------

__global__ void k_ekv_build(model_t md, const double *x, const int *bfi /*[4*npla]*/,
                            const int *cross /*[4*npla]*/, const zd *ph /*[4*npla*4]*/, ekv_t *ekv,
                            int t_lo, int t_hi) {
  const int npla = md.npla;
  const long n = (long)(t_hi - t_lo + 1) * 4 * npla;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int b = e % npla;
    const int f = (e / npla) % 4;
    const int t = t_lo + (int)(e / (4 * npla));
    const int fb = f * npla + b;
    const double xv = x[bfi[fb] + (size_t)md.Nb * (t - 1)];
    ekv_t out;
    expkv_compact(md, f, xv, cross[fb] ? &ph[4 * fb] : nullptr, out);
    ekv[((size_t)(t - 1) * 4 + f) * npla + b] = out;
  }
}

// The EPSOC model's 2x2 bond exponentials for n (family, field, phase)
// triples: out[4 i .. 4 i + 3] is E (or E^-1) of triple i, column-major.
__global__ void k_ekv_epsoc(double dtau, const zd *kcoe, const zd *vcoe, const zd *k12, const zd *v12, int nfam,
                            long n, const int *fam, const double *x, const int *cross, const zd *ph, int inv,
                            zd *out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int f = fam[i];
    zd e[4];
    if (f < 0 || f >= nfam) {  // refused on the host; never read outside the tables
      e[0] = e[1] = e[2] = e[3] = zmk(0.0, 0.0);
    } else {
      expkv_epsoc(dtau, kcoe[f], vcoe[f], k12[f], v12[f], x[i], (ph && cross[i]) ? &ph[i] : nullptr, inv != 0, e);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) out[4 * i + q] = e[q];
  }
}

extern "C" hipError_t pqmc_launch_ekv_epsoc(double dtau, const zd *kcoe, const zd *vcoe, const zd *k12,
                                            const zd *v12, int nfam, long n, const int *fam, const double *x,
                                            const int *cross, const zd *ph, int inv, zd *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_ekv_epsoc, dim3((unsigned)blocks), dim3(256), 0, s, dtau, kcoe, vcoe, k12, v12, nfam, n, fam,
                     x, cross, ph, inv, out);
  return hipGetLastError();
}

extern "C" hipError_t pqmc_launch_ekv_build(const model_t *md, const double *x, const int *bfi,
                                            const int *cross, const zd *ph, ekv_t *ekv, int t_lo,
                                            int t_hi, hipStream_t s) {
  const long n = (long)(t_hi - t_lo + 1) * 4 * md->npla;
  const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_ekv_build, dim3(blocks), dim3(256), 0, s, *md, x, bfi, cross, ph, ekv, t_lo, t_hi);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Field gradient of ln_cw for one family at one slice (no counterpart in the
// reference).  With G positioned so that E_f(t) is the outermost factor of
// its right string (just before the backward sweep's wrap E^-1 G E),
// d ln det / dx = Tr[G dE E^-1], restricted to the bond's four sites:
// sum_ij (de e^-1)(i,j) G(S_j, S_i), one thread per bond of a diagonal tile
// of k_flush_wrap, from the tile's G before the wrap.  With
// n = delt sqrt(|K|^2 + (|V| x)^2), s = sinh(n)/n and nd = delt^2 |V|^2 x:
// d cosh(n)/dx = s nd, d s/dx = r nd with r = (n cosh n - sinh n)/n^3 (its
// series 1/3 + n^2/30 below n = 1e-3), dh/dx = -delt V_coe V12.

// d ln det / dx of one bond from its G(S, S) block g (column-major 4x4),
// tr_weight doubling and ncopy applied.
PQ_DEV zd bond_field_grad(const model_t &md, int f, double xv, const zd *p, const ekv_t &e, const zd g[16]) {
  const double ak = hypot(md.kcoe[f].x, md.kcoe[f].y);
  const double av = hypot(md.vcoe[f].x, md.vcoe[f].y);
  const double n = md.dtau * sqrt(ak * ak + (av * xv) * (av * xv));
  const double nd = md.dtau * md.dtau * av * av * xv;
  const double s = n > 0.0 ? sinh(n) / n : 1.0;
  const double r = n < 1e-3 ? 1.0 / 3.0 + n * n / 30.0 : (n * cosh(n) - sinh(n)) / (n * n * n);
  const zd vx = zscale(md.vcoe[f], xv);
  // dE = dc I + [[0, dm],[dm^H, 0]]
  ekv_t de;
  de.c = s * nd;
  de.pad = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    zd h = zscale(zadd(zmul(md.kcoe[f], md.k12[f][q]), zmul(vx, md.v12[f][q])), -md.dtau);
    zd dh = zscale(zmul(md.vcoe[f], md.v12[f][q]), -md.dtau);
    if (p) {
      h = zmul(h, p[q]);
      dh = zmul(dh, p[q]);
    }
    de.m[q] = zadd(zscale(dh, s), zscale(h, r * nd));
  }
  zd tr = zmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      zd mij = zmk(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 4; ++k) mij = zadd(mij, zmul(ekv_entry(de, false, i, k), ekv_entry(e, true, k, j)));
      tr = zadd(tr, zmul(mij, g[j + 4 * i]));
    }
  if (md.tr_weight) tr = zmk(2.0 * tr.x, 0.0);
  return zscale(tr, (double)md.ncopy);
}
