// Dense complex FP64 kernels for the QR stabilization and the Green's
// function from scratch:
//  * k_zgemm<opA,opB>: C = alpha op(A) op(B) + beta C on FP64 MFMA (the
//    ZGEMM calls of cal_g_and_det_T0, src/mod_evolution.f90:344-350).
//  * k_qr_panel + k_apply_refl: blocked Householder QR with the LAPACK
//    (zgeqrf/zungqr) real-beta convention, as ZGEQRF/ZUNGQR in zqr
//    (src/mod_matrixlib.f90:452-467).
//  * k_trsm_upper: Y = R^-1 Z, used for (L R)^-1 through the QR of L R.
//  * small reductions: log-weight, max |G_fast - G| (err_fast).
#include <mutex>

#include "pqmc_common.h"

// ---------------------------------------------------------------------------
// GEMM

#define GM_T 32
#define GM_KC 32

// One 32 x 32 tile of C per workgroup (four waves, a 16 x 16 quarter each,
// v_mfma_f64_16x16x4f64), K in chunks of 32 through LDS.  Every thread holds
// four elements of A's and four of B's chunk: the loads of the next chunk are
// in flight while the MFMAs consume the current one.  The complex product
// takes three real products: P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi),
// re = P1 - P2, im = P3 - P1 - P2.
template <int OPA, int OPB>
static __device__ __forceinline__ void zgemm_fetch(int M, int N, int K, const zd *A, int lda, const zd *B, int ldb,
                                                   int m0, int n0, int k0, zd (&va)[4], zd (&vb)[4]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + 256 * u;
    int i, k;
    if (OPA == 0) {
      i = e % GM_T; k = e / GM_T;
      va[u] = (m0 + i < M && k0 + k < K) ? A[(m0 + i) + (size_t)lda * (k0 + k)] : zmk(0.0, 0.0);
    } else {
      k = e % GM_KC; i = e / GM_KC;
      va[u] = (m0 + i < M && k0 + k < K) ? zconj(A[(k0 + k) + (size_t)lda * (m0 + i)]) : zmk(0.0, 0.0);
    }
    int j;
    if (OPB == 0) {
      k = e % GM_KC; j = e / GM_KC;
      vb[u] = (n0 + j < N && k0 + k < K) ? B[(k0 + k) + (size_t)ldb * (n0 + j)] : zmk(0.0, 0.0);
    } else {
      j = e % GM_T; k = e / GM_T;
      vb[u] = (n0 + j < N && k0 + k < K) ? zconj(B[(n0 + j) + (size_t)ldb * (k0 + k)]) : zmk(0.0, 0.0);
    }
  }
}

template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, zd alpha, const zd *A, int lda,
                                               const zd *B, int ldb, zd beta, zd *C, int ldc) {
  // rows padded to GM_T + 1: a transposed operand's stores (32 threads along
  // k) then fall in different banks
  __shared__ double sAr[GM_KC][GM_T + 1], sAi[GM_KC][GM_T + 1], sBr[GM_KC][GM_T + 1], sBi[GM_KC][GM_T + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.x * GM_T, n0 = blockIdx.y * GM_T;
  const int wi = wv >> 1, wj = wv & 1;
  double4_t p1 = {0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
  zd va[4], vb[4];
  zgemm_fetch<OPA, OPB>(M, N, K, A, lda, B, ldb, m0, n0, 0, va, vb);
  for (int k0 = 0; k0 < K; k0 += GM_KC) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      const int ia = OPA == 0 ? e % GM_T : e / GM_KC, ka = OPA == 0 ? e / GM_T : e % GM_KC;
      sAr[ka][ia] = va[u].x;
      sAi[ka][ia] = va[u].y;
      const int jb = OPB == 0 ? e / GM_KC : e % GM_T, kb = OPB == 0 ? e % GM_KC : e / GM_T;
      sBr[kb][jb] = vb[u].x;
      sBi[kb][jb] = vb[u].y;
    }
    __syncthreads();
    if (k0 + GM_KC < K) zgemm_fetch<OPA, OPB>(M, N, K, A, lda, B, ldb, m0, n0, k0 + GM_KC, va, vb);
#pragma unroll
    for (int kk = 0; kk < GM_KC / 4; ++kk) {
      const int krow = kk * 4 + (lane >> 4);
      const double ar = sAr[krow][wi * 16 + (lane & 15)], ai = sAi[krow][wi * 16 + (lane & 15)];
      const double br = sBr[krow][wj * 16 + (lane & 15)], bi = sBi[krow][wj * 16 + (lane & 15)];
      p1 = mfma_f64(ar, br, p1);
      p2 = mfma_f64(ai, bi, p2);
      p3 = mfma_f64(ar + ai, br + bi, p3);
    }
    __syncthreads();
  }
  const double4_t cr = p1 - p2, ci = (p3 - p1) - p2;
  const bool b0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = m0 + wi * 16 + (lane >> 4) + 4 * i;
    const int col = n0 + wj * 16 + (lane & 15);
    if (row < M && col < N) {
      zd v = zmul(alpha, zmk(cr[i], ci[i]));
      zd &cref = C[row + (size_t)ldc * col];
      if (!b0) v = zadd(v, zmul(beta, cref));
      cref = v;
    }
  }
}

extern "C" hipError_t pqmc_zgemm(int opA, int opB, int M, int N, int K, zd alpha, const zd *A, int lda,
                                 const zd *B, int ldb, zd beta, zd *C, int ldc, hipStream_t s) {
  dim3 grid((M + GM_T - 1) / GM_T, (N + GM_T - 1) / GM_T);
  if (opA == 0 && opB == 0)
    hipLaunchKernelGGL((k_zgemm<0, 0>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (opA == 1 && opB == 0)
    hipLaunchKernelGGL((k_zgemm<1, 0>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (opA == 0 && opB == 1)
    hipLaunchKernelGGL((k_zgemm<0, 1>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else
    hipLaunchKernelGGL((k_zgemm<1, 1>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Householder panel: rows [j0, m), columns [j0, j0+nb) of A.
// zgeqr2 / zlarfg: beta = -sign(Re alpha) |(alpha, x)|, tau = ((beta-Re a)/beta,
// -Im a/beta), v = x/(alpha-beta); H^H = I - conj(tau) v v^H applied from the
// left.  T of the block reflector H = I - V T V^H as zlarft('F','C').
//
// One workgroup of 16 waves holds the panel (mp <= 1024 rows) in LDS.  Per
// column: wave 0 forms |x|^2 of the column below the diagonal while wave w
// forms x^H a_{jj+w} of the unscaled column with column jj+w (all in
// parallel); then 16 threads derive the reflector and the row-vector
// w = tau^* v^H A, and every thread updates one row.  Three barriers per
// column, no serial thread-0 loops.

#define QR_THREADS 1024
#define QR_WAVES (QR_THREADS / 64)
#define QR_CPW (16 / QR_WAVES)  // panel columns per wave in the column products

__global__ __launch_bounds__(QR_THREADS) void k_qr_panel(zd *A, int lda, int m, int j0, int nb, zd *V,
                                                          int ldv, zd *T, zd *tau, double *vn2,
                                                          zd *beta_out) {
  extern __shared__ zd sP[];  // mp x nb, ld mp
  __shared__ zd sDot[16], sA0[16], sW[16];
  __shared__ zd sT[16 * 16], sZ[16 * 16];
  __shared__ double sZr[4][16][17], sZi[4][16][17];
  __shared__ zd sTaus[16], sBetas[16], sScal[2];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int mp = m - j0;
  // load the panel with 16 loads in flight per thread
  for (int e0 = tid; e0 < mp * nb; e0 += 16 * QR_THREADS) {
    zd v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = e0 + u * QR_THREADS;
      v[u] = (e < mp * nb) ? A[(j0 + e % mp) + (size_t)lda * (j0 + e / mp)] : zmk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = e0 + u * QR_THREADS;
      if (e < mp * nb) sP[e] = v[u];
    }
  }
  __syncthreads();
  for (int jj = 0; jj < nb; ++jj) {
    // phase A: wave wv forms the products of column jj (rows > jj) with
    // columns jj + wv + QR_WAVES q; for q = 0 on wave 0 that is |x|^2.  Rows come in
    // batches of four per lane with every load issued before the products.
    {
      double pr[QR_CPW], pim[QR_CPW];
#pragma unroll
      for (int q = 0; q < QR_CPW; ++q) pr[q] = pim[q] = 0.0;
      for (int r0 = jj + 1; r0 < mp; r0 += 64 * 4) {
        zd xs[4], as[QR_CPW][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = r0 + lane + 64 * i;
          xs[i] = (r < mp) ? sP[r + mp * jj] : zmk(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < QR_CPW; ++q) {
            const int c = jj + wv + QR_WAVES * q;
            as[q][i] = (r < mp && c < nb) ? sP[r + mp * c] : zmk(0.0, 0.0);
          }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int q = 0; q < QR_CPW; ++q) {
            pr[q] += xs[i].x * as[q][i].x + xs[i].y * as[q][i].y;  // conj(x) a
            pim[q] += xs[i].x * as[q][i].y - xs[i].y * as[q][i].x;
          }
      }
#pragma unroll
      for (int q = 0; q < QR_CPW; ++q) {
        const int c = jj + wv + QR_WAVES * q;
        if (c < nb) {  // uniform per wave
          const double sr = wave_sum(pr[q]), si = wave_sum(pim[q]);
          if (lane == 0) {
            sDot[c - jj] = zmk(sr, si);
            sA0[c - jj] = sP[jj + mp * c];
          }
        }
      }
    }
    __syncthreads();
    // phase B1: reflector scalars and w_c = tau^* (a_{jj,c} + sc^* x^H a_c)
    if (tid < 16) {
      const zd alpha = sA0[0];
      const double tot = sDot[0].x;
      zd t, sc, bt;
      double v2;
      if (tot == 0.0 && alpha.y == 0.0) {
        t = zmk(0.0, 0.0);
        sc = zmk(0.0, 0.0);
        bt = alpha;
        v2 = 1.0;
      } else {
        const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + tot), alpha.x);
        t = zmk((beta - alpha.x) / beta, -alpha.y / beta);
        sc = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
        bt = zmk(beta, 0.0);
        v2 = 1.0 + tot * (sc.x * sc.x + sc.y * sc.y);
      }
      const int cc = jj + tid;
      if (tid > 0 && cc < nb) sW[tid] = zmul(zconj(t), zadd(sA0[tid], zmul(zconj(sc), sDot[tid])));
      if (tid == 0) {
        tau[jj] = t;
        beta_out[jj] = bt;
        vn2[jj] = v2;
        sTaus[jj] = t;
        sBetas[jj] = bt;
        sScal[0] = sc;
        sScal[1] = t;
      }
    }
    __syncthreads();
    // phase B2: rows: a_rc -= v_r w_c (c > jj), v_r = x_r sc (r > jj), 1 (r = jj)
    {
      const zd sc = sScal[0], t = sScal[1];
      const bool has = (t.x != 0.0 || t.y != 0.0);
      for (int r = jj + tid; r < mp; r += QR_THREADS) {
        const zd x = sP[r + mp * jj];
        const zd v = (r == jj) ? zmk(1.0, 0.0) : zmul(x, sc);
        for (int c2 = jj + 1; c2 < nb; ++c2) sP[r + mp * c2] = zsub(sP[r + mp * c2], zmul(v, sW[c2 - jj]));
        if (r > jj && has) sP[r + mp * jj] = v;
      }
    }
    __syncthreads();
  }
  // panel back to A: R above (beta on the diagonal), V below; V explicit
  for (int e = tid; e < mp * nb; e += QR_THREADS) {
    const int i = e % mp, c = e / mp;
    const zd p = sP[e];
    A[(j0 + i) + (size_t)lda * (j0 + c)] = (i == c) ? sBetas[c] : p;
    const zd v = (i < c) ? zmk(0.0, 0.0) : (i == c ? zmk(1.0, 0.0) : p);
    V[(j0 + i) + (size_t)ldv * (j0 + c)] = v;
    sP[e] = v;
  }
  __syncthreads();
  // Gram V^H V on MFMA: wave wv < 4 takes row chunks 4 (wv + 4q)
  if (wv < 4) {
    const int li = lane & 15, lk = lane >> 4;
    double4_t zr = {0.0, 0.0, 0.0, 0.0}, zi = {0.0, 0.0, 0.0, 0.0};
    for (int r0 = 4 * wv; r0 < mp; r0 += 16) {
      const int r = r0 + lk;
      zd a = zmk(0.0, 0.0);
      if (r < mp && li < nb) a = sP[r + mp * li];
      // conj(V)^T V: A operand conj(V[r][i]) (lane i = li, k = lk), B operand V[r][j] (k = lk, j = li)
      zr = mfma_f64(a.x, a.x, zr);
      zr = mfma_f64(a.y, a.y, zr);
      zi = mfma_f64(a.x, a.y, zi);
      zi = mfma_f64(-a.y, a.x, zi);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      sZr[wv][lk + 4 * i][li] = zr[i];
      sZi[wv][lk + 4 * i][li] = zi[i];
    }
  }
  __syncthreads();
  // combine the four partial Gram tiles: z(l, i) into sZ
  if (tid < 256) {
    const int l = tid >> 4, i = tid & 15;
    sZ[l + 16 * i] = zmk(sZr[0][l][i] + sZr[1][l][i] + sZr[2][l][i] + sZr[3][l][i],
                         sZi[0][l][i] + sZi[1][l][i] + sZi[2][l][i] + sZi[3][l][i]);
  }
  __syncthreads();
  // T of zlarft (T(0:i, i) = -tau_i T(0:i, 0:i) z(0:i, i), T(i, i) = tau_i)
  // in closed form: T^-1 = diag(1/tau) + striu(Z), so T = (I + N)^-1 diag(tau)
  // with N = diag(tau) striu(Z) strictly upper, N^16 = 0, and
  // (I + N)^-1 = (I - N)(I + N^2)(I + N^4)(I + N^8): six 16 x 16 products,
  // one entry per thread, instead of 16 dependent steps (and no 1/tau, so
  // tau = 0 needs no care).  The partial-Gram scratch holds N's powers and W.
  {
    zd *sPw = reinterpret_cast<zd *>(&sZr[0][0][0]);  // N, then its powers
    zd *sWk = sPw + 256;                              // the product so far
    const int i = tid & 15, k = (tid >> 4) & 15;
    const bool own = tid < 256;  // one entry of the 16 x 16 products per thread
    if (own) {
      const zd ti = i < nb ? sTaus[i] : zmk(0.0, 0.0);
      const zd nik = (i < k && k < nb) ? zmul(ti, sZ[i + 16 * k]) : zmk(0.0, 0.0);
      sPw[i + 16 * k] = nik;
      sWk[i + 16 * k] = zsub(zmk(i == k ? 1.0 : 0.0, 0.0), nik);
    }
    __syncthreads();
    for (int s = 0; s < 3; ++s) {
      // P <- P P (N^2, N^4, N^8), then W <- W + W P
      zd pp = zmk(0.0, 0.0);
      if (own)
        for (int l = 0; l < 16; ++l) pp = zadd(pp, zmul(sPw[i + 16 * l], sPw[l + 16 * k]));
      __syncthreads();
      if (own) sPw[i + 16 * k] = pp;
      __syncthreads();
      zd wp = zmk(0.0, 0.0);
      if (own)
        for (int l = 0; l < 16; ++l) wp = zadd(wp, zmul(sWk[i + 16 * l], sPw[l + 16 * k]));
      __syncthreads();
      if (own) sWk[i + 16 * k] = zadd(sWk[i + 16 * k], wp);
      __syncthreads();
    }
    if (own) sT[i + 16 * k] = k < nb ? zmul(sWk[i + 16 * k], sTaus[k]) : zmk(0.0, 0.0);
  }
  __syncthreads();
  for (int e = tid; e < nb * nb; e += QR_THREADS) {
    const int k = e % nb, i = e / nb;
    T[k + nb * i] = sT[k + 16 * i];
  }
}

extern "C" hipError_t pqmc_qr_panel(zd *A, int lda, int m, int j0, int nb, zd *V, int ldv, zd *T,
                                    zd *tau, double *vn2, zd *beta_out, hipStream_t s) {
  if (m - j0 > 1024 || nb > 16) return hipErrorInvalidValue;
  const size_t lds = sizeof(zd) * (size_t)(m - j0) * nb;
  static std::mutex attr_mu;  // launches may come from several host threads
  static size_t attr = 48 * 1024;
  {
    std::lock_guard<std::mutex> lk(attr_mu);
    if (lds > attr) {
      hipError_t e = hipFuncSetAttribute((const void *)k_qr_panel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = lds;
    }
  }
  hipLaunchKernelGGL(k_qr_panel, dim3(1), dim3(QR_THREADS), lds, s, A, lda, m, j0, nb, V, ldv, T, tau, vn2,
                     beta_out);
  return hipGetLastError();
}

// A(mp x ncols) -= V op(T) (V^H A); op(T) = T^H (trans=1, applies H^H) or T.
// One workgroup (4 waves) per 16 columns of A, both products on FP64 MFMA
// (v_mfma_f64_16x16x4f64, complex as four real products):
//   W = V^H A_blk   (nb x 16, the K = mp rows split over the 4 waves),
//   W2 = op(T) W,   A_blk -= V W2 (16-row tiles over the waves).
#define RF_THREADS 256
#define RF_WAVES (RF_THREADS / 64)

__global__ __launch_bounds__(RF_THREADS) void k_apply_refl(zd *A, int lda, int mp, int ncols, const zd *V, int ldv,
                                                           const zd *T, int nb, int trans) {
  __shared__ double sWr[RF_WAVES][16][17], sWi[RF_WAVES][16][17];
  __shared__ double sW2r[16][17], sW2i[16][17];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c0 = blockIdx.x * 16;
  const int li = lane & 15, lk = lane >> 4;
  const int acol = c0 + li;
  const bool acol_ok = acol < ncols;
  // ---- W = V^H A_blk: wave wv takes row chunks 4 (wv + RF_WAVES q), 16 chunks per batch ----
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  for (int r0 = 4 * wv; r0 < mp; r0 += 4 * RF_WAVES * 16) {
    zd vv[16], xx[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int r = r0 + 4 * RF_WAVES * u + lk;  // chunk r0 + 4 RF_WAVES u, row lk of it
      vv[u] = (r < mp && li < nb) ? V[r + (size_t)ldv * li] : zmk(0.0, 0.0);
      xx[u] = (r < mp && acol_ok) ? A[r + (size_t)lda * acol] : zmk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      // conj(V)^T A: re = Vr Ar + Vi Ai, im = Vr Ai - Vi Ar
      cr = mfma_f64(vv[u].x, xx[u].x, cr);
      cr = mfma_f64(vv[u].y, xx[u].y, cr);
      ci = mfma_f64(vv[u].x, xx[u].y, ci);
      ci = mfma_f64(-vv[u].y, xx[u].x, ci);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    sWr[wv][lk + 4 * i][li] = cr[i];
    sWi[wv][lk + 4 * i][li] = ci[i];
  }
  __syncthreads();
  // ---- W2 = op(T) W ----
  if (tid < 256) {
    const int kk = tid >> 4, cc = tid & 15;
    double wr = 0.0, wi = 0.0;
    if (kk < nb) {
      for (int k = 0; k < nb; ++k) {
        const zd tk = trans ? zconj(T[k + nb * kk]) : T[kk + nb * k];
        double xr = 0.0, xi = 0.0;
#pragma unroll
        for (int w = 0; w < RF_WAVES; ++w) {
          xr += sWr[w][k][cc];
          xi += sWi[w][k][cc];
        }
        wr += tk.x * xr - tk.y * xi;
        wi += tk.x * xi + tk.y * xr;
      }
    }
    sW2r[kk][cc] = wr;
    sW2i[kk][cc] = wi;
  }
  __syncthreads();
  // ---- A_blk -= V W2, 16-row tiles ----
  double w2r[4], w2i[4];
#pragma unroll
  for (int k4 = 0; k4 < 4; ++k4) {
    w2r[k4] = sW2r[4 * k4 + lk][li];
    w2i[k4] = sW2i[4 * k4 + lk][li];
  }
  // 16-row tiles t0 = 16 (wv + RF_WAVES q), four tiles per batch: V and A loads first
  for (int tb = 16 * wv; tb < mp; tb += 4 * 16 * RF_WAVES) {
    zd vv[4][4], aa[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t0 = tb + 16 * RF_WAVES * q;
      const int r = t0 + li;
#pragma unroll
      for (int k4 = 0; k4 < 4; ++k4) {
        const int kc = 4 * k4 + lk;
        vv[q][k4] = (r < mp && kc < nb) ? V[r + (size_t)ldv * kc] : zmk(0.0, 0.0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = t0 + lk + 4 * i;
        aa[q][i] = (row < mp && acol_ok) ? A[row + (size_t)lda * acol] : zmk(0.0, 0.0);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t0 = tb + 16 * RF_WAVES * q;
      double4_t pr = {0.0, 0.0, 0.0, 0.0}, pi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k4 = 0; k4 < 4; ++k4) {
        // (Vr + i Vi)(Wr + i Wi)
        pr = mfma_f64(vv[q][k4].x, w2r[k4], pr);
        pr = mfma_f64(-vv[q][k4].y, w2i[k4], pr);
        pi = mfma_f64(vv[q][k4].x, w2i[k4], pi);
        pi = mfma_f64(vv[q][k4].y, w2r[k4], pi);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = t0 + lk + 4 * i;
        if (row < mp && acol_ok) A[row + (size_t)lda * acol] = zsub(aa[q][i], zmk(pr[i], pi[i]));
      }
    }
  }
}

extern "C" hipError_t pqmc_apply_refl(zd *A, int lda, int mp, int ncols, const zd *V, int ldv, const zd *T,
                                      int nb, int trans, hipStream_t s) {
  if (ncols <= 0) return hipSuccess;
  if (nb > 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_apply_refl, dim3((ncols + 15) / 16), dim3(RF_THREADS), 0, s, A, lda, mp, ncols, V, ldv, T,
                     nb, trans);
  return hipGetLastError();
}

// Q = [I; 0] (m x n)
__global__ void k_eye(zd *Q, int ldq, int m, int n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)m * n; e += (long)gridDim.x * blockDim.x) {
    const int i = e % m, j = e / m;
    Q[i + (size_t)ldq * j] = zmk(i == j ? 1.0 : 0.0, 0.0);
  }
}

extern "C" hipError_t pqmc_eye(zd *Q, int ldq, int m, int n, hipStream_t s) {
  long tot = (long)m * n;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_eye, dim3(blocks), dim3(256), 0, s, Q, ldq, m, n);
  return hipGetLastError();
}

// D = diag(R) and the unit upper R' = D^-1 R of qdr (src/mod_matrixlib.f90:323-334)
// from the factored A (R in its upper triangle).  conjt stores R'^H (ldq) and
// D = conj(diag(R)): the factors of ldq (:349-360) of A^H, whose diagonal
// ZGELQF leaves conjugated, (beta, -0).
__global__ void k_qdr_finish(const zd *A, int lda, int n, zd *Rout, zd *D, int conjt) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    const zd d = A[i + (size_t)lda * i];
    zd v = (i <= j) ? zdiv(A[i + (size_t)lda * j], d) : zmk(0.0, 0.0);
    if (conjt) Rout[j + (size_t)n * i] = zconj(v);
    else Rout[i + (size_t)n * j] = v;
    if (i == j) D[i] = conjt ? zconj(d) : d;
  }
}

extern "C" hipError_t pqmc_qdr_finish(const zd *A, int lda, int n, zd *Rout, zd *D, int conjt, hipStream_t s) {
  int blocks = (n * n + 255) / 256;
  hipLaunchKernelGGL(k_qdr_finish, dim3(blocks), dim3(256), 0, s, A, lda, n, Rout, D, conjt);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Y = R^-1 Z, R upper triangular (n x n, ldr), Z (n x ncols, ldz), in place.
// The 16 x 16 diagonal blocks of R are inverted first (k_trinv_diag, all
// blocks at once); the solve then walks the blocks from the bottom with no
// sequential division chain: y_b = R_bb^-1 z_b, z_above -= R(above, b) y_b.

#define TS_COLS 4
#define TS_KB 16

// Rinv[b] (16 x 16, column-major) = inverse of diagonal block b; lane j < 16
// back-substitutes column j.
__global__ __launch_bounds__(64) void k_trinv_diag(const zd *R, int ldr, int n, zd *Rinv) {
  __shared__ zd sR[TS_KB * TS_KB];
  const int b = blockIdx.x, k0 = TS_KB * b, nk = min(TS_KB, n - k0);
  const int lane = threadIdx.x;
  for (int e = lane; e < TS_KB * TS_KB; e += 64) {
    const int i = e % TS_KB, k = e / TS_KB;
    sR[e] = (i < nk && k < nk) ? R[(k0 + i) + (size_t)ldr * (k0 + k)] : zmk(i == k ? 1.0 : 0.0, 0.0);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < TS_KB) {
    const int j = lane;
    zd x[TS_KB];
#pragma unroll
    for (int i = 0; i < TS_KB; ++i) x[i] = zmk(0.0, 0.0);
#pragma unroll
    for (int i = TS_KB - 1; i >= 0; --i) {
      if (i <= j) {
        zd s = zmk(i == j ? 1.0 : 0.0, 0.0);
#pragma unroll
        for (int k = i + 1; k < TS_KB; ++k)
          if (k <= j) s = zsub(s, zmul(sR[i + TS_KB * k], x[k]));
        x[i] = zdiv(s, sR[i + TS_KB * i]);
      }
    }
#pragma unroll
    for (int i = 0; i < TS_KB; ++i) Rinv[(size_t)b * TS_KB * TS_KB + i + TS_KB * j] = x[i];
  }
}

__global__ __launch_bounds__(256) void k_trsm_upper(const zd *R, int ldr, int n, zd *Z, int ldz, int ncols,
                                                    const zd *Rinv) {
  extern __shared__ zd sm[];
  zd *sZ = sm;                // TS_COLS x n, column c at sZ[c * n + i]
  zd *sR = sm + TS_COLS * n;  // n x TS_KB: sR[i + n * kk] = R(i, klo + kk)
  __shared__ zd sRi[TS_KB * TS_KB], sY[TS_COLS * TS_KB];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * TS_COLS;
  const int ncw = min(TS_COLS, ncols - c0);
  for (int e = tid; e < TS_COLS * n; e += 256) {
    const int c = e / n, i = e % n;
    sZ[e] = (c < ncw) ? Z[i + (size_t)ldz * (c0 + c)] : zmk(0.0, 0.0);
  }
  const int nblk = (n + TS_KB - 1) / TS_KB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int klo = TS_KB * b, nk = min(TS_KB, n - klo);
    __syncthreads();
    for (int e = tid; e < klo * nk; e += 256) {
      const int i = e % klo, kk = e / klo;
      sR[i + n * kk] = R[i + (size_t)ldr * (klo + kk)];
    }
    for (int e = tid; e < TS_KB * TS_KB; e += 256) sRi[e] = Rinv[(size_t)b * TS_KB * TS_KB + e];
    __syncthreads();
    if (tid < TS_COLS * TS_KB) {
      const int c = tid / TS_KB, i = tid % TS_KB;
      zd y = zmk(0.0, 0.0);
      if (i < nk) {
#pragma unroll
        for (int k = 0; k < TS_KB; ++k)
          if (k >= i && k < nk) y = zadd(y, zmul(sRi[i + TS_KB * k], sZ[c * n + klo + k]));
      }
      sY[tid] = y;
    }
    __syncthreads();
    if (tid < TS_COLS * TS_KB) {
      const int c = tid / TS_KB, i = tid % TS_KB;
      if (i < nk) sZ[c * n + klo + i] = sY[tid];
    }
    for (int i = tid; i < klo; i += 256) {
      zd rr[TS_KB];
#pragma unroll
      for (int kk = 0; kk < TS_KB; ++kk) rr[kk] = (kk < nk) ? sR[i + n * kk] : zmk(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < TS_COLS; ++c) {
        zd acc = sZ[c * n + i];
#pragma unroll
        for (int kk = 0; kk < TS_KB; ++kk)
          if (kk < nk) acc = zsub(acc, zmul(rr[kk], sY[c * TS_KB + kk]));
        sZ[c * n + i] = acc;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < ncw * n; e += 256) {
    const int c = e / n, i = e % n;
    Z[i + (size_t)ldz * (c0 + c)] = sZ[c * n + i];
  }
}

// The same for n <= 16 + 256: thread i < klo keeps row i of R's column
// panel [klo, klo + 16) in registers, and the next panel's row (with the
// next diagonal inverse) is loaded while the current one is used, so that
// no step of the block back substitution waits on a load it issues.
#define TS_NMAX_R (TS_KB + 256)

struct ts_panel {
  zd r[TS_KB];
  zd ri;
};

static __device__ __forceinline__ void ts_fetch(const zd *R, int ldr, int n, const zd *Rinv, int b, ts_panel &p) {
  const int tid = threadIdx.x, klo = TS_KB * b, nk = min(TS_KB, n - klo);
#pragma unroll
  for (int kk = 0; kk < TS_KB; ++kk) p.r[kk] = (tid < klo && kk < nk) ? R[tid + (size_t)ldr * (klo + kk)] : zmk(0.0, 0.0);
  p.ri = Rinv[(size_t)b * TS_KB * TS_KB + tid];
}

// One block step: y_b = Rinv_b z_b, then rows above b less R(:, b) y_b, with
// panel b in cur; panel b - 1 is fetched into nxt meanwhile.
static __device__ __forceinline__ void ts_step(const zd *R, int ldr, int n, const zd *Rinv, zd *sZ, zd *sRi, zd *sY, int b,
                                               const ts_panel &cur, ts_panel &nxt) {
  const int tid = threadIdx.x, klo = TS_KB * b, nk = min(TS_KB, n - klo);
  sRi[tid] = cur.ri;
  __syncthreads();
  if (b > 0) ts_fetch(R, ldr, n, Rinv, b - 1, nxt);
  if (tid < TS_COLS * TS_KB) {
    const int c = tid / TS_KB, i = tid % TS_KB;
    zd y = zmk(0.0, 0.0);
    if (i < nk) {
#pragma unroll
      for (int k = 0; k < TS_KB; ++k)
        if (k >= i && k < nk) y = zadd(y, zmul(sRi[i + TS_KB * k], sZ[c * n + klo + k]));
    }
    sY[tid] = y;
  }
  __syncthreads();
  if (tid < TS_COLS * TS_KB) {
    const int c = tid / TS_KB, i = tid % TS_KB;
    if (i < nk) sZ[c * n + klo + i] = sY[tid];
  }
  if (tid < klo) {
#pragma unroll
    for (int c = 0; c < TS_COLS; ++c) {
      zd acc = sZ[c * n + tid];
#pragma unroll
      for (int kk = 0; kk < TS_KB; ++kk)
        if (kk < nk) acc = zsub(acc, zmul(cur.r[kk], sY[c * TS_KB + kk]));
      sZ[c * n + tid] = acc;
    }
  }
}

__global__ __launch_bounds__(256) void k_trsm_upper_r(const zd *R, int ldr, int n, zd *Z, int ldz, int ncols,
                                                      const zd *Rinv) {
  extern __shared__ zd sm[];
  zd *sZ = sm;  // TS_COLS x n, column c at sZ[c * n + i]
  __shared__ zd sRi[TS_KB * TS_KB], sY[TS_COLS * TS_KB];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * TS_COLS;
  const int ncw = min(TS_COLS, ncols - c0);
  const int nz = TS_COLS * n;  // <= 5 * 256
  {
    zd t[5];
#pragma unroll
    for (int u = 0; u < 5; ++u) {
      const int e = tid + 256 * u, c = e / n;
      t[u] = (e < nz && c < ncw) ? Z[e % n + (size_t)ldz * (c0 + c)] : zmk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 5; ++u)
      if (tid + 256 * u < nz) sZ[tid + 256 * u] = t[u];
  }
  const int nblk = (n + TS_KB - 1) / TS_KB;
  ts_panel pa, pb;
  ts_fetch(R, ldr, n, Rinv, nblk - 1, pa);
  for (int b = nblk - 1; b >= 0; b -= 2) {
    ts_step(R, ldr, n, Rinv, sZ, sRi, sY, b, pa, pb);
    if (b == 0) break;
    ts_step(R, ldr, n, Rinv, sZ, sRi, sY, b - 1, pb, pa);
  }
  __syncthreads();
  for (int e = tid; e < ncw * n; e += 256) {
    const int c = e / n, i = e % n;
    Z[i + (size_t)ldz * (c0 + c)] = sZ[c * n + i];
  }
}

extern "C" hipError_t pqmc_trsm_upper(const zd *R, int ldr, int n, zd *Z, int ldz, int ncols, zd *work,
                                      hipStream_t s) {
  const int nblk = (n + TS_KB - 1) / TS_KB;
  hipLaunchKernelGGL(k_trinv_diag, dim3(nblk), dim3(64), 0, s, R, ldr, n, work);
  if (n <= TS_NMAX_R) {  // TS_COLS n entries of LDS: at most 17 KB
    hipLaunchKernelGGL(k_trsm_upper_r, dim3((ncols + TS_COLS - 1) / TS_COLS), dim3(256), sizeof(zd) * TS_COLS * (size_t)n,
                       s, R, ldr, n, Z, ldz, ncols, (const zd *)work);
    return hipGetLastError();
  }
  const size_t lds = sizeof(zd) * (size_t)n * (TS_COLS + TS_KB);
  static std::mutex attr_mu;  // launches may come from several host threads
  static size_t attr = 48 * 1024;
  {
    std::lock_guard<std::mutex> lk(attr_mu);
    if (lds > attr) {
      hipError_t e = hipFuncSetAttribute((const void *)k_trsm_upper, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = lds;
    }
  }
  hipLaunchKernelGGL(k_trsm_upper, dim3((ncols + TS_COLS - 1) / TS_COLS), dim3(256), lds, s, R, ldr, n, Z, ldz,
                     ncols, (const zd *)work);
  return hipGetLastError();
}

// sum_i log(D(i)) of one D_string block (principal complex logs), into *out
__global__ void k_dlog(const zd *D, int n, int conj, zd *out) {
  __shared__ double sr[4], si[4];
  const int tid = threadIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int j = tid; j < n; j += 256) {
    const zd l = zlog(conj ? zconj(D[j]) : D[j]);
    ar += l.x;
    ai += l.y;
  }
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  if ((tid & 63) == 0) {
    sr[tid >> 6] = ar;
    si[tid >> 6] = ai;
  }
  __syncthreads();
  if (tid == 0) *out = zmk(sr[0] + sr[1] + sr[2] + sr[3], si[0] + si[1] + si[2] + si[3]);
}

extern "C" hipError_t pqmc_dlog(const zd *D, int n, zd *out, hipStream_t s) {
  hipLaunchKernelGGL(k_dlog, dim3(1), dim3(256), 0, s, D, n, 0, out);
  return hipGetLastError();
}

// sum_i log conj(D_i): the diagonal of an LQ (ZGELQF) taken from the QR of
// the conjugate transpose; a negative real entry then adds -i pi
extern "C" hipError_t pqmc_dlog_conj(const zd *D, int n, zd *out, hipStream_t s) {
  hipLaunchKernelGGL(k_dlog, dim3(1), dim3(256), 0, s, D, n, 1, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ln_cw of cal_g_and_det_T0 (src/mod_evolution.f90:345-356):
//   log det(L R) + sum(log D_string); TR doubling; *ncopy; Im mod 2*pi_SP.
// det(L R) = det(Q) det(R) of its Householder QR: det(Q) = prod(1 - tau_i |v_i|^2),
// det(R) = prod(beta_i); the log is reduced to the principal branch like
// log(det(.)) of the reference.

#define SP_TWO_PI 6.28318548202514648438  // 2*PI in REAL(SP), src/mod_nrtype.f90:42
#define DP_PI 3.14159265358979323846

// With drift_call set, the fast log-weight, the value *lncw_out held before
// plus the accepts' log(real(R_elec)) terms *acc since then, is compared with
// the new one: |Re(new - fast)| goes into *drift_call and (as a running
// maximum) *drift_run, uint64 bits as in k_maxdiff (get_g_scratch_T0,
// src/mod_evolution.f90:187-192).  An accepted ratio with a negative real
// part makes the fast value NaN (log of a negative real,
// src/mod_update.f90:142); the reference's comparison then fails and the
// check is skipped, and so it is here.  A non-finite new value counts as +inf.
__global__ void k_lncw(int ne, const zd *betaM, const zd *tauM, const double *vn2M, const zd *dlog, int nblk,
                       int tr_weight, int ncopy, zd *lncw_out, const double *acc, unsigned long long *drift_call,
                       unsigned long long *drift_run, int conjq) {
  __shared__ double sr[4], si[4];
  const int tid = threadIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int j = tid; j < ne; j += 256) {
    const zd lb = zlog(betaM[j]);
    const zd dq = zsub(zmk(1.0, 0.0), zscale(tauM[j], vn2M[j]));
    const zd lq = zlog(dq);
    ar += lb.x + lq.x;
    ai += lb.y + lq.y;
  }
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  if ((tid & 63) == 0) {
    sr[tid >> 6] = ar;
    si[tid >> 6] = ai;
  }
  __syncthreads();
  if (tid == 0) {
    // conjq: the QR was of M^H, whose determinant is the conjugate of det M
    const double ldr = sr[0] + sr[1] + sr[2] + sr[3], ldi_q = si[0] + si[1] + si[2] + si[3];
    const double ldi_raw = conjq ? -ldi_q : ldi_q;
    double dr = 0.0, di = 0.0;
    for (int b = 0; b < nblk; ++b) {
      dr += dlog[b].x;
      di += dlog[b].y;
    }
    // principal branch of log(det)
    const double ldi = atan2(sin(ldi_raw), cos(ldi_raw));
    zd lw = zmk(ldr + dr, ldi + di);
    if (tr_weight) lw = zadd(lw, zconj(lw));
    lw = zscale(lw, (double)ncopy);
    lw.y = fmod(lw.y, SP_TWO_PI);
    if (drift_call) {
      const double fast = lncw_out->x + *acc;
      double e = isfinite(lw.x) ? fabs(lw.x - fast) : INFINITY;
      if (e == e) {
        atomicMax(drift_call, (unsigned long long)__double_as_longlong(e));
        atomicMax(drift_run, (unsigned long long)__double_as_longlong(e));
      }
    }
    *lncw_out = lw;
  }
}

extern "C" hipError_t pqmc_lncw(int ne, const zd *betaM, const zd *tauM, const double *vn2M, const zd *dlog, int nblk,
                                int tr_weight, int ncopy, zd *out, const double *acc, unsigned long long *drift_call,
                                unsigned long long *drift_run, int conjq, hipStream_t s) {
  hipLaunchKernelGGL(k_lncw, dim3(1), dim3(256), 0, s, ne, betaM, tauM, vn2M, dlog, nblk, tr_weight, ncopy, out, acc,
                     drift_call, drift_run, conjq);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Finite temperature (proj = .false.): the scale-split Green's function of
// cal_g_and_det_FT (src/mod_evolution.f90:494-648).  A scale d splits into a
// big part db = (|d| > 1 ? d : 1) and a small part ds = (|d| > 1 ? 1 : d).

PQ_DEV zd ft_big(zd d) { return (d.x * d.x + d.y * d.y > 1.0) ? d : zmk(1.0, 0.0); }
PQ_DEV zd ft_small(zd d) { return (d.x * d.x + d.y * d.y > 1.0) ? zmk(1.0, 0.0) : d; }

// Scales of B(beta,0) reach exp(beta * band edge) (about exp(400) at
// beta = 100), so nothing here squares one: a / d rescales d by a power of
// two first, and the string steps factor X Dh with D = Dh 2^e, |Dh| in
// [0.5, 1), restoring the powers of two exactly afterwards.
PQ_DEV zd zdiv_scaled(zd a, zd d) {
  int e;
  frexp(fmax(fabs(d.x), fabs(d.y)), &e);
  const zd q = zdiv(a, zmk(ldexp(d.x, -e), ldexp(d.y, -e)));
  return zmk(ldexp(q.x, -e), ldexp(q.y, -e));
}

// X (n x n): row i divided by the big part of D[i]
__global__ void k_ft_rowdiv(zd *X, int n, const zd *D) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  X[e] = zdiv_scaled(X[e], ft_big(D[e % n]));
}

// Pre-pivoting of one string step (one workgroup of 1024; Bai, Lee, Li, Xu,
// Linear Algebra Appl. 435 (2011) 659-673): column j of X D has squared
// norm 4^e_j v_j with |D_j| = a_j 2^e_j (frexp) and v_j = a_j^2 sum_i
// |X_ij|^2; perm orders the columns by it, decreasing, ties by index
// (compared through ldexp, exact); ep[k] = e_perm[k].
__global__ __launch_bounds__(1024) void k_ft_prep(const zd *X, int n, const zd *D, int *perm, int *ep) {
  __shared__ double v[1024];
  __shared__ int ex[1024];
  const int tid = threadIdx.x, lane = tid & 63, nw = blockDim.x >> 6;
  for (int j = tid >> 6; j < n; j += nw) {
    double ss = 0.0;
    for (int i = lane; i < n; i += 64) {
      const zd x = X[i + (size_t)n * j];
      ss += x.x * x.x + x.y * x.y;
    }
    ss = wave_sum(ss);
    if (lane == 0) {
      int e;
      const double a = frexp(hypot(D[j].x, D[j].y), &e);
      v[j] = a * a * ss;
      ex[j] = e;
    }
  }
  __syncthreads();
  for (int j = tid; j < n; j += blockDim.x) {
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const double vi = ldexp(v[i], 2 * (ex[i] - ex[j]));
      rank += (vi > v[j]) || (vi == v[j] && i < j);
    }
    perm[rank] = j;
    ep[rank] = ex[j];
  }
}

// Xp(:, k) = X(:, perm[k]) Dh_perm[k], Dh = D 2^-e (its conjugate if conj)
__global__ void k_ft_gather_scale(const zd *X, zd *Xp, int n, const zd *D, const int *perm, const int *ep,
                                  int conj) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e % n), k = (int)(e / n), j = perm[k];
  zd dh = zmk(ldexp(D[j].x, -ep[k]), ldexp(D[j].y, -ep[k]));
  if (conj) dh = zconj(dh);
  Xp[e] = zmul(X[i + (size_t)n * j], dh);
}

// The powers of two back after the QR of Xp: D[k] *= 2^ep[k]; the unit
// triangular factor T(r, c) *= 2^(ep[c] - ep[r]) (conjt, its conjugate
// transpose: 2^(ep[r] - ep[c]))
__global__ void k_ft_unscale(zd *T, zd *D, int n, const int *ep, int conjt) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const int s = conjt ? ep[r] - ep[c] : ep[c] - ep[r];
  T[e] = zmk(ldexp(T[e].x, s), ldexp(T[e].y, s));
  if (e < n) D[e] = zmk(ldexp(D[e].x, ep[e]), ldexp(D[e].y, ep[e]));
}

// rows: dst(k, :) = src(perm[k], :); else dst(:, k) = src(:, perm[k])
__global__ void k_ft_permute(const zd *src, zd *dst, int n, const int *perm, int rows) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e % n), k = (int)(e / n);
  dst[e] = rows ? src[perm[i] + (size_t)n * k] : src[i + (size_t)n * perm[k]];
}

// M = db_r^-1 QQ db_l^-1 + ds_r TT ds_l  (n x n, ld n)
__global__ void k_ft_mix(zd *M, const zd *QQ, const zd *TT, const zd *Dr, const zd *Dl, int n) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n);
  const zd a = zdiv_scaled(zdiv_scaled(QQ[e], ft_big(Dr[i])), ft_big(Dl[j]));
  const zd b = zmul(zmul(ft_small(Dr[i]), TT[e]), ft_small(Dl[j]));
  M[e] = zadd(a, b);
}

// out[0] = sum log db_r, out[1] = sum log db_l (principal logs), out[2] =
// *Lr, out[3] = *Ll (log det of the two unitary factors)
__global__ void k_ft_dlog(const zd *Dr, const zd *Dl, const zd *Lr, const zd *Ll, int n, zd *out) {
  __shared__ double s[4][4];
  const int tid = threadIdx.x;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = tid; j < n; j += 256) {
    const zd lr = zlog(ft_big(Dr[j])), ll = zlog(ft_big(Dl[j]));
    a[0] += lr.x;
    a[1] += lr.y;
    a[2] += ll.x;
    a[3] += ll.y;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double v = wave_sum(a[k]);
    if ((tid & 63) == 0) s[tid >> 6][k] = v;
  }
  __syncthreads();
  if (tid == 0) {
    out[0] = zmk(s[0][0] + s[1][0] + s[2][0] + s[3][0], s[0][1] + s[1][1] + s[2][1] + s[3][1]);
    out[1] = zmk(s[0][2] + s[1][2] + s[2][2] + s[3][2], s[0][3] + s[1][3] + s[2][3] + s[3][3]);
    out[2] = *Lr;
    out[3] = *Ll;
  }
}

// *out = log det Q of a Householder QR (det Q = prod_i (1 - tau_i |v_i|^2)),
// principal branch, or its conjugate (the determinant of Q^H)
__global__ void k_ft_qlog(const zd *tau, const double *vn2, int n, int conj, zd *out) {
  __shared__ double sr[4], si[4];
  const int tid = threadIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int j = tid; j < n; j += 256) {
    const zd l = zlog(zsub(zmk(1.0, 0.0), zscale(tau[j], vn2[j])));
    ar += l.x;
    ai += l.y;
  }
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  if ((tid & 63) == 0) {
    sr[tid >> 6] = ar;
    si[tid >> 6] = ai;
  }
  __syncthreads();
  if (tid == 0) {
    const double raw = si[0] + si[1] + si[2] + si[3];
    const double im = atan2(sin(raw), cos(raw));
    *out = zmk(sr[0] + sr[1] + sr[2] + sr[3], conj ? -im : im);
  }
}

// D[0..n) = 1
__global__ void k_ft_ones(zd *D, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) D[e] = zmk(1.0, 0.0);
}

// G <- 1 - G (g_h from (1 + B)^-1)
__global__ void k_ft_finish(zd *G, int n) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n);
  G[e] = zsub(zmk(i == j ? 1.0 : 0.0, 0.0), G[e]);
}

static inline int ft_grid(long n) { return (int)((n + 255) / 256); }

extern "C" hipError_t pqmc_ft_rowdiv(zd *X, int n, const zd *D, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_rowdiv, dim3(ft_grid((long)n * n)), dim3(256), 0, s, X, n, D);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_prep(const zd *X, int n, const zd *D, int *perm, int *ep, hipStream_t s) {
  if (n > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ft_prep, dim3(1), dim3(1024), 0, s, X, n, D, perm, ep);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_gather_scale(const zd *X, zd *Xp, int n, const zd *D, const int *perm, const int *ep,
                                           int conj, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_gather_scale, dim3(ft_grid((long)n * n)), dim3(256), 0, s, X, Xp, n, D, perm, ep, conj);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_unscale(zd *T, zd *D, int n, const int *ep, int conjt, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_unscale, dim3(ft_grid((long)n * n)), dim3(256), 0, s, T, D, n, ep, conjt);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_permute(const zd *src, zd *dst, int n, const int *perm, int rows, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_permute, dim3(ft_grid((long)n * n)), dim3(256), 0, s, src, dst, n, perm, rows);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_mix(zd *M, const zd *QQ, const zd *TT, const zd *Dr, const zd *Dl, int n,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_ft_mix, dim3(ft_grid((long)n * n)), dim3(256), 0, s, M, QQ, TT, Dr, Dl, n);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_dlog(const zd *Dr, const zd *Dl, const zd *Lr, const zd *Ll, int n, zd *out,
                                   hipStream_t s) {
  hipLaunchKernelGGL(k_ft_dlog, dim3(1), dim3(256), 0, s, Dr, Dl, Lr, Ll, n, out);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_qlog(const zd *tau, const double *vn2, int n, int conj, zd *out, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_qlog, dim3(1), dim3(256), 0, s, tau, vn2, n, conj, out);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_ones(zd *D, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_ones, dim3(ft_grid(n)), dim3(256), 0, s, D, n);
  return hipGetLastError();
}
extern "C" hipError_t pqmc_ft_finish(zd *G, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_ft_finish, dim3(ft_grid((long)n * n)), dim3(256), 0, s, G, n);
  return hipGetLastError();
}

// max |a - b| over n complex entries into *out, and into *run when set (as
// uint64 bits of a nonnegative double, so that an unsigned atomic max is the
// max of the values; a NaN is taken as +inf)
__global__ void k_maxdiff(const zd *a, const zd *b, long n, unsigned long long *out, unsigned long long *run) {
  double m = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const zd d = zsub(a[e], b[e]);
    const double v = hypot(d.x, d.y);
    if (!(v <= m)) m = v;  // NaN propagates
  }
  if (m != m) m = INFINITY;  // fmax in the wave reduction would drop a NaN
  m = wave_max(m);
  // one atomic per workgroup: thousands on one address serialize
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      if (sm[w] > m) m = sm[w];
    }
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
    if (run) atomicMax(run, (unsigned long long)__double_as_longlong(m));
  }
}

extern "C" hipError_t pqmc_maxdiff(const zd *a, const zd *b, long n, unsigned long long *out, unsigned long long *run,
                                   hipStream_t s) {
  int blocks = (int)((n + 1023) / 1024);
  if (blocks > 256) blocks = 256;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_maxdiff, dim3(blocks), dim3(256), 0, s, a, b, n, out, run);
  return hipGetLastError();
}

// C = A^H (A: m x n, lda) -> C: n x m (ldc)
__global__ void k_conjt(const zd *A, int lda, int m, int n, zd *C, int ldc) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)m * n; e += (long)gridDim.x * blockDim.x) {
    const int i = e % m, j = e / m;
    C[j + (size_t)ldc * i] = zconj(A[i + (size_t)lda * j]);
  }
}

extern "C" hipError_t pqmc_conjt(const zd *A, int lda, int m, int n, zd *C, int ldc, hipStream_t s) {
  long tot = (long)m * n;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_conjt, dim3(blocks), dim3(256), 0, s, A, lda, m, n, C, ldc);
  return hipGetLastError();
}

// MFMA operand-layout self test: C(16x16) = A(16x4) B(4x16), one wave.
__global__ void k_mfma_probe(const double *A, const double *B, double *C) {
  const int l = threadIdx.x;
  const double a = A[(l & 15) + 16 * (l >> 4)];  // A[i][k] at i + 16k
  const double b = B[(l >> 4) + 4 * (l & 15)];   // B[k][j] at k + 4j
  double4_t c = {0.0, 0.0, 0.0, 0.0};
  c = mfma_f64(a, b, c);
  for (int i = 0; i < 4; ++i) C[((l >> 4) + 4 * i) + 16 * (l & 15)] = c[i];
}

extern "C" hipError_t pqmc_mfma_probe(const double *A, const double *B, double *C, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, C);
  return hipGetLastError();
}
