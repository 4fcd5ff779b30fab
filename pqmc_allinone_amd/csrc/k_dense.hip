// Dense complex FP64 kernels for the QR stabilization and the Green's
// function from scratch:
//  * k_zgemm<opA,opB>: C = alpha op(A) op(B) + beta C on FP64 MFMA (the
//    ZGEMM calls of cal_g_and_det_T0, src/mod_evolution.f90:344-350).
//  * k_qr_panel + k_apply_refl: blocked Householder QR with the LAPACK
//    (zgeqrf/zungqr) real-beta convention, as ZGEQRF/ZUNGQR in zqr
//    (src/mod_matrixlib.f90:452-467).
//  * k_trsm_upper: Y = R^-1 Z, used for (L R)^-1 through the QR of L R.
//  * small reductions: log-weight, max |G_fast - G| (err_fast).
#include "pqmc_common.h"

// ---------------------------------------------------------------------------
// GEMM

#define GM_T 32
#define GM_KC 16

template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, zd alpha, const zd *A, int lda,
                                               const zd *B, int ldb, zd beta, zd *C, int ldc) {
  __shared__ double sAr[GM_KC][GM_T], sAi[GM_KC][GM_T], sBr[GM_KC][GM_T], sBi[GM_KC][GM_T];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.x * GM_T, n0 = blockIdx.y * GM_T;
  const int wi = wv >> 1, wj = wv & 1;
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += GM_KC) {
    for (int e = tid; e < GM_T * GM_KC; e += 256) {
      int i, k;
      zd v = zmk(0.0, 0.0);
      if (OPA == 0) {
        i = e % GM_T; k = e / GM_T;
        if (m0 + i < M && k0 + k < K) v = A[(m0 + i) + (size_t)lda * (k0 + k)];
      } else {
        k = e % GM_KC; i = e / GM_KC;
        if (m0 + i < M && k0 + k < K) v = zconj(A[(k0 + k) + (size_t)lda * (m0 + i)]);
      }
      sAr[k][i] = v.x;
      sAi[k][i] = v.y;
    }
    for (int e = tid; e < GM_T * GM_KC; e += 256) {
      int j, k;
      zd v = zmk(0.0, 0.0);
      if (OPB == 0) {
        k = e % GM_KC; j = e / GM_KC;
        if (n0 + j < N && k0 + k < K) v = B[(k0 + k) + (size_t)ldb * (n0 + j)];
      } else {
        j = e % GM_T; k = e / GM_T;
        if (n0 + j < N && k0 + k < K) v = zconj(B[(n0 + j) + (size_t)ldb * (k0 + k)]);
      }
      sBr[k][j] = v.x;
      sBi[k][j] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GM_KC / 4; ++kk) {
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

#define QR_THREADS 1024

__global__ __launch_bounds__(QR_THREADS) void k_qr_panel(zd *A, int lda, int m, int j0, int nb, zd *V,
                                                          int ldv, zd *T, zd *tau, double *vn2,
                                                          zd *beta_out) {
  extern __shared__ zd sP[];  // mp x nb, ld mp
  __shared__ double sRed[QR_THREADS / 64];
  __shared__ zd sTau, sScale, sBeta, sZ[16];
  __shared__ zd sT[16 * 16];
  __shared__ zd sTaus[16], sBetas[16];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int mp = m - j0;
  for (int e = tid; e < mp * nb; e += QR_THREADS) {
    const int i = e % mp, c = e / mp;
    sP[e] = A[(j0 + i) + (size_t)lda * (j0 + c)];
  }
  __syncthreads();
  for (int jj = 0; jj < nb; ++jj) {
    double ss = 0.0;
    for (int i = jj + 1 + tid; i < mp; i += QR_THREADS) {
      const zd v = sP[i + mp * jj];
      ss += v.x * v.x + v.y * v.y;
    }
    ss = wave_sum(ss);
    if (lane == 0) sRed[wv] = ss;
    __syncthreads();
    if (tid == 0) {
      double tot = 0.0;
      for (int w = 0; w < QR_THREADS / 64; ++w) tot += sRed[w];
      const zd alpha = sP[jj + mp * jj];
      const double xnorm = sqrt(tot);
      if (xnorm == 0.0 && alpha.y == 0.0) {
        sTau = zmk(0.0, 0.0);
        sScale = zmk(0.0, 0.0);
        sBeta = alpha;
        vn2[jj] = 1.0;
      } else {
        const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + tot), alpha.x);
        sTau = zmk((beta - alpha.x) / beta, -alpha.y / beta);
        const zd sc = zdiv(zmk(1.0, 0.0), zsub(alpha, zmk(beta, 0.0)));
        sScale = sc;
        sBeta = zmk(beta, 0.0);
        vn2[jj] = 1.0 + tot * (sc.x * sc.x + sc.y * sc.y);
      }
      tau[jj] = sTau;
      beta_out[jj] = sBeta;
      sTaus[jj] = sTau;
      sBetas[jj] = sBeta;
    }
    __syncthreads();
    const zd t = sTau;
    if (t.x != 0.0 || t.y != 0.0) {
      for (int i = jj + 1 + tid; i < mp; i += QR_THREADS) sP[i + mp * jj] = zmul(sP[i + mp * jj], sScale);
    }
    __syncthreads();
    if (tid == 0) sP[jj + mp * jj] = zmk(1.0, 0.0);
    __syncthreads();
    // apply (I - conj(tau) v v^H) to columns jj+1..nb-1, one wave per column
    const int c = jj + 1 + wv;
    if (c < nb) {
      double wr = 0.0, wi = 0.0;
      for (int i = jj + lane; i < mp; i += 64) {
        const zd v = sP[i + mp * jj], a = sP[i + mp * c];
        wr += v.x * a.x + v.y * a.y;  // conj(v) a
        wi += v.x * a.y - v.y * a.x;
      }
      wr = wave_sum(wr);
      wi = wave_sum(wi);
      const zd w = zmul(zconj(t), zmk(wr, wi));
      for (int i = jj + lane; i < mp; i += 64) sP[i + mp * c] = zsub(sP[i + mp * c], zmul(sP[i + mp * jj], w));
    }
    __syncthreads();
  }
  // panel back to A: R above (beta on the diagonal), V below
  for (int e = tid; e < mp * nb; e += QR_THREADS) {
    const int i = e % mp, c = e / mp;
    A[(j0 + i) + (size_t)lda * (j0 + c)] = (i == c) ? sBetas[c] : sP[e];
  }
  __syncthreads();
  // V explicit (unit diagonal, zero above) and zlarft T
  for (int e = tid; e < mp * nb; e += QR_THREADS) {
    const int i = e % mp, c = e / mp;
    const zd v = (i < c) ? zmk(0.0, 0.0) : (i == c ? zmk(1.0, 0.0) : sP[e]);
    V[(j0 + i) + (size_t)ldv * (j0 + c)] = v;
    sP[e] = v;
  }
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    // z_k = V(:,k)^H V(:,i), k < i; wave k
    if (wv < i) {
      double zr = 0.0, zi = 0.0;
      for (int r = i + lane; r < mp; r += 64) {
        const zd a = sP[r + mp * wv], b = sP[r + mp * i];
        zr += a.x * b.x + a.y * b.y;
        zi += a.x * b.y - a.y * b.x;
      }
      zr = wave_sum(zr);
      zi = wave_sum(zi);
      if (lane == 0) sZ[wv] = zmk(zr, zi);
    }
    __syncthreads();
    if (tid == 0) {
      const zd ti = sTaus[i];
      for (int k = 0; k < i; ++k) {
        zd s = zmk(0.0, 0.0);
        for (int l = k; l < i; ++l) s = zadd(s, zmul(sT[k + 16 * l], sZ[l]));
        sT[k + 16 * i] = zmul(zneg(ti), s);
      }
      for (int k = i + 1; k < 16; ++k) sT[k + 16 * i] = zmk(0.0, 0.0);
      sT[i + 16 * i] = ti;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += QR_THREADS) {
    const int k = e % nb, i = e / nb;
    T[k + nb * i] = sT[k + 16 * i];
  }
}

extern "C" hipError_t pqmc_qr_panel(zd *A, int lda, int m, int j0, int nb, zd *V, int ldv, zd *T,
                                    zd *tau, double *vn2, zd *beta_out, hipStream_t s) {
  const size_t lds = sizeof(zd) * (size_t)(m - j0) * nb;
  static size_t attr = 48 * 1024;
  if (lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_qr_panel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL(k_qr_panel, dim3(1), dim3(QR_THREADS), lds, s, A, lda, m, j0, nb, V, ldv, T, tau, vn2,
                     beta_out);
  return hipGetLastError();
}

// A(mp x ncols) -= V op(T) (V^H A); op(T) = T^H (trans=1, applies H^H) or T.
#define AR_ROWS 64
__global__ __launch_bounds__(256) void k_apply_refl(zd *A, int lda, int mp, int ncols, const zd *V, int ldv,
                                                    const zd *T, int nb, int trans) {
  __shared__ zd sV[AR_ROWS][17], sA[AR_ROWS][17];
  __shared__ zd sW[16][16], sW2[16][16];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * 16;
  const int ncb = min(16, ncols - c0);
  const int kk = tid / 16, cc = tid % 16;
  zd acc = zmk(0.0, 0.0);
  for (int r0 = 0; r0 < mp; r0 += AR_ROWS) {
    const int nr = min(AR_ROWS, mp - r0);
    for (int e = tid; e < AR_ROWS * 16; e += 256) {
      const int i = e % AR_ROWS, c = e / AR_ROWS;
      sV[i][c] = (i < nr && c < nb) ? V[(r0 + i) + (size_t)ldv * c] : zmk(0.0, 0.0);
      sA[i][c] = (i < nr && c < ncb) ? A[(r0 + i) + (size_t)lda * (c0 + c)] : zmk(0.0, 0.0);
    }
    __syncthreads();
    for (int i = 0; i < nr; ++i) acc = zadd(acc, zcmul(sV[i][kk], sA[i][cc]));
    __syncthreads();
  }
  sW[kk][cc] = acc;
  __syncthreads();
  {
    zd s = zmk(0.0, 0.0);
    if (kk < nb) {
      for (int k = 0; k < nb; ++k) {
        const zd tk = trans ? zconj(T[k + nb * kk]) : T[kk + nb * k];
        s = zadd(s, zmul(tk, sW[k][cc]));
      }
    }
    sW2[kk][cc] = s;
  }
  __syncthreads();
  for (int r0 = 0; r0 < mp; r0 += AR_ROWS) {
    const int nr = min(AR_ROWS, mp - r0);
    for (int e = tid; e < AR_ROWS * 16; e += 256) {
      const int i = e % AR_ROWS, c = e / AR_ROWS;
      sV[i][c] = (i < nr && c < nb) ? V[(r0 + i) + (size_t)ldv * c] : zmk(0.0, 0.0);
    }
    __syncthreads();
    for (int e = tid; e < nr * ncb; e += 256) {
      const int i = e % nr, c = e / nr;
      zd s = zmk(0.0, 0.0);
      for (int k = 0; k < nb; ++k) s = zadd(s, zmul(sV[i][k], sW2[k][c]));
      zd &ar = A[(r0 + i) + (size_t)lda * (c0 + c)];
      ar = zsub(ar, s);
    }
    __syncthreads();
  }
}

extern "C" hipError_t pqmc_apply_refl(zd *A, int lda, int mp, int ncols, const zd *V, int ldv, const zd *T,
                                      int nb, int trans, hipStream_t s) {
  if (ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_refl, dim3((ncols + 15) / 16), dim3(256), 0, s, A, lda, mp, ncols, V, ldv, T, nb,
                     trans);
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
// from the factored A (R in its upper triangle).  conjt stores R'^H (ldq).
__global__ void k_qdr_finish(const zd *A, int lda, int n, zd *Rout, zd *D, int conjt) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    const zd d = A[i + (size_t)lda * i];
    zd v = (i <= j) ? zdiv(A[i + (size_t)lda * j], d) : zmk(0.0, 0.0);
    if (conjt) Rout[j + (size_t)n * i] = zconj(v);
    else Rout[i + (size_t)n * j] = v;
    if (i == j) D[i] = d;
  }
}

extern "C" hipError_t pqmc_qdr_finish(const zd *A, int lda, int n, zd *Rout, zd *D, int conjt, hipStream_t s) {
  int blocks = (n * n + 255) / 256;
  hipLaunchKernelGGL(k_qdr_finish, dim3(blocks), dim3(256), 0, s, A, lda, n, Rout, D, conjt);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Y = R^-1 Z, R upper triangular (n x n, ldr), Z (n x ncols, ldz), in place.

#define TS_COLS 4
#define TS_KB 16
__global__ __launch_bounds__(256) void k_trsm_upper(const zd *R, int ldr, int n, zd *Z, int ldz, int ncols) {
  extern __shared__ zd sm[];
  zd *sZ = sm;                    // TS_COLS x n
  zd *sR = sm + TS_COLS * n;      // n x TS_KB (rows 0..n-1 of a block of columns)
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * TS_COLS;
  const int ncw = min(TS_COLS, ncols - c0);
  for (int e = tid; e < TS_COLS * n; e += 256) {
    const int c = e / n, i = e % n;
    sZ[e] = (c < ncw) ? Z[i + (size_t)ldz * (c0 + c)] : zmk(0.0, 0.0);
  }
  for (int kb = n; kb > 0; kb -= TS_KB) {
    const int klo = max(0, kb - TS_KB);
    const int nk = kb - klo;
    __syncthreads();
    for (int e = tid; e < kb * nk; e += 256) {
      const int i = e % kb, kk = e / kb;
      sR[i + n * kk] = R[i + (size_t)ldr * (klo + kk)];
    }
    __syncthreads();
    for (int k = kb - 1; k >= klo; --k) {
      const int kk = k - klo;
      if (tid < TS_COLS) sZ[tid * n + k] = zdiv(sZ[tid * n + k], sR[k + n * kk]);
      __syncthreads();
      for (int e = tid; e < TS_COLS * k; e += 256) {
        const int c = e / k, i = e % k;
        sZ[c * n + i] = zsub(sZ[c * n + i], zmul(sR[i + n * kk], sZ[c * n + k]));
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < ncw * n; e += 256) {
    const int c = e / n, i = e % n;
    Z[i + (size_t)ldz * (c0 + c)] = sZ[c * n + i];
  }
}

extern "C" hipError_t pqmc_trsm_upper(const zd *R, int ldr, int n, zd *Z, int ldz, int ncols, hipStream_t s) {
  const size_t lds = sizeof(zd) * (size_t)n * (TS_COLS + TS_KB);
  static size_t attr = 48 * 1024;
  if (lds > attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_trsm_upper, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = lds;
  }
  hipLaunchKernelGGL(k_trsm_upper, dim3((ncols + TS_COLS - 1) / TS_COLS), dim3(256), lds, s, R, ldr, n, Z, ldz,
                     ncols);
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

__global__ void k_lncw(int ne, const zd *betaM, const zd *tauM, const double *vn2M, const zd *D, int nD,
                       int tr_weight, int ncopy, zd *lncw_out) {
  __shared__ double sr[256], si[256];
  const int tid = threadIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int j = tid; j < ne; j += 256) {
    const zd lb = zlog(betaM[j]);
    const zd dq = zsub(zmk(1.0, 0.0), zscale(tauM[j], vn2M[j]));
    const zd lq = zlog(dq);
    ar += lb.x + lq.x;
    ai += lb.y + lq.y;
  }
  sr[tid] = ar;
  si[tid] = ai;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) { sr[tid] += sr[tid + o]; si[tid] += si[tid + o]; }
    __syncthreads();
  }
  const double ldr = sr[0], ldi_raw = si[0];
  __syncthreads();
  double dr = 0.0, di = 0.0;
  for (int j = tid; j < nD; j += 256) {
    const zd l = zlog(D[j]);
    dr += l.x;
    di += l.y;
  }
  sr[tid] = dr;
  si[tid] = di;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) { sr[tid] += sr[tid + o]; si[tid] += si[tid + o]; }
    __syncthreads();
  }
  if (tid == 0) {
    // principal branch of log(det)
    double ldi = atan2(sin(ldi_raw), cos(ldi_raw));
    zd lw = zmk(ldr + sr[0], ldi + si[0]);
    if (tr_weight) lw = zadd(lw, zconj(lw));
    lw = zscale(lw, (double)ncopy);
    lw.y = fmod(lw.y, SP_TWO_PI);
    *lncw_out = lw;
  }
}

extern "C" hipError_t pqmc_lncw(int ne, const zd *betaM, const zd *tauM, const double *vn2M, const zd *D, int nD,
                                int tr_weight, int ncopy, zd *out, hipStream_t s) {
  hipLaunchKernelGGL(k_lncw, dim3(1), dim3(256), 0, s, ne, betaM, tauM, vn2M, D, nD, tr_weight, ncopy, out);
  return hipGetLastError();
}

// max |a - b| over n complex entries into *out (as uint64 bits of a nonnegative double)
__global__ void k_maxdiff(const zd *a, const zd *b, long n, unsigned long long *out) {
  double m = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const zd d = zsub(a[e], b[e]);
    const double v = hypot(d.x, d.y);
    if (!(v <= m)) m = v;  // NaN propagates
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

extern "C" hipError_t pqmc_maxdiff(const zd *a, const zd *b, long n, unsigned long long *out, hipStream_t s) {
  int blocks = (int)((n + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_maxdiff, dim3(blocks), dim3(256), 0, s, a, b, n, out);
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
