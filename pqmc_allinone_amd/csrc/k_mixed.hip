// Mixed-precision B-string products: C = P Q between two QR checkpoints,
// where P (M x K) is the product of the ngroup slices' family layers, formed
// in FP64, and Q (K x N) the checkpoint's string (the left_evolve /
// right_evolve loops of get_g_scratch_T0 and init_g_T0,
// src/mod_evolution.f90:88-102, 137-162, taken as one dense product).  The
// complex product runs as real products on split operands:
//
//  * FP32 three-term split (mode 1): every entry, scaled by a power of two
//    per row of P (per column of Q) into [-1, 1], splits as
//        t = t0 + t1 + t2,  t0 on the grid 2^-11, t1 on the grid 2^-22,
//        t2 = fp32(t - t0 - t1),
//    an error-free splitting in the manner of Ozaki, Ogita, Oishi and Rump
//    (Numer. Algorithms 59 (2012) 95-118).  t0 s0 products are integers of at
//    most 22 bits in units of 2^-22 and four of them (one K-step of
//    v_mfma_f32_16x16x4f32) sum exactly in FP32; each K-step's partial
//    products are added to FP64 accumulators at once.  The terms t0 s0,
//    t0 s1 + t1 s0 and t0 s2 + t1 s1 + t2 s0 + t1 s2 + t2 s1 are kept; the
//    dropped t2 s2 is below 2^-46 of scale_row * scale_col, and the FP32
//    roundings of the higher-order sums about 2^-44, per K-step.
//  * BF16x3 (mode 2): every scaled entry, rounded to FP32, splits into three
//    bfloat16 terms b0 + b1 + b2 (b0 = bf16(t), b1 = bf16(t - b0),
//    b2 = bf16(t - b0 - b1)), as in Henry, Tang and Heinecke (ARITH 2019);
//    the six terms of order up to two run on v_mfma_f32_16x16x32_bf16 with
//    FP32 accumulation over each K-step of 32, then FP64 accumulation.  The
//    result is about as accurate as an FP32 product.
//
// QR, D, the (L R)^-1 solve and every determinant ratio stay FP64.
#include "pqmc_common.h"

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// per-row (rows != 0) or per-column power-of-two scales: 2^e with every
// |Re|, |Im| of the row (column) below 2^e, or 1 for a zero row

__global__ __launch_bounds__(256) void k_mx_scales(const zd *A, int lda, int nrow, int ncol, int rows, double *scale) {
  const int lane = threadIdx.x & 63;
  const int idx = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int n = rows ? nrow : ncol, len = rows ? ncol : nrow;
  if (idx >= n) return;  // whole waves
  double m = 0.0;
  for (int k = lane; k < len; k += 64) {
    const zd v = rows ? A[idx + (size_t)lda * k] : A[k + (size_t)lda * idx];
    m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
  }
  m = wave_max(m);
  if (lane == 0) {
    int e = 0;
    if (m > 0.0) {
      frexp(m, &e);  // m = f 2^e with f in [0.5, 1)
    }
    scale[idx] = m > 0.0 ? ldexp(1.0, e) : 1.0;
  }
}

// ---------------------------------------------------------------------------
// splits

struct sp3 {
  float t0, t1, t2;
};
PQ_DEV sp3 split_fp32(double t) {
  sp3 s;
  const double a0 = rint(t * 2048.0) * (1.0 / 2048.0);
  const double r0 = t - a0;
  const double a1 = rint(r0 * 4194304.0) * (1.0 / 4194304.0);
  s.t0 = (float)a0;
  s.t1 = (float)a1;
  s.t2 = (float)(r0 - a1);
  return s;
}

// round-to-nearest-even FP32 -> bfloat16 bits (finite inputs)
PQ_DEV unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
PQ_DEV float bf16_float(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }

// ---------------------------------------------------------------------------
// C (M x N) = P (M x K) Q (K x N), complex, column-major; rs[M] and cs[N]
// are the power-of-two scales of P's rows and Q's columns.  One wave per
// 16 x 16 tile of C, four tiles per workgroup; no barriers.
//
// Operand layouts (CDNA3/4 16 x 16 MFMA): lane l supplies row l % 16 of A and
// column l % 16 of B at K index (l / 16) (x4f32) or 8 (l / 16) + 0..7
// (x32_bf16); the result holds rows 4 (l / 16) + r, r = 0..3, of column
// l % 16.

#define MX_TILE 16

PQ_DEV void mx_acc(double4_t &acc, f4_t v) {
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] += (double)v[r];
}

__global__ __launch_bounds__(256) void k_mx_gemm_fp32(int M, int N, int K, const zd *P, int lda, const zd *Q, int ldb,
                                                      const double *rs, const double *cs, zd *C, int ldc) {
  const int lane = threadIdx.x & 63;
  const int MT = (M + MX_TILE - 1) / MX_TILE, NT = (N + MX_TILE - 1) / MX_TILE;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= MT * NT) return;
  const int tm = tile % MT, tn = tile / MT;
  const int li = lane & 15, lk = lane >> 4;
  const int i = tm * MX_TILE + li, j = tn * MX_TILE + li;
  const double ri = i < M ? 1.0 / rs[i] : 0.0, cj = j < N ? 1.0 / cs[j] : 0.0;
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  const f4_t z = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + lk;
    const zd a = (i < M && k < K) ? P[i + (size_t)lda * k] : zmk(0.0, 0.0);
    const zd b = (j < N && k < K) ? Q[k + (size_t)ldb * j] : zmk(0.0, 0.0);
    const sp3 ar = split_fp32(a.x * ri), ai = split_fp32(a.y * ri);
    const sp3 br = split_fp32(b.x * cj), bi = split_fp32(b.y * cj);
    // order 0, exact: four separate accumulations
    mx_acc(cr, __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t0, br.t0, z, 0, 0, 0));
    mx_acc(cr, __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t0, bi.t0, z, 0, 0, 0));
    mx_acc(ci, __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t0, bi.t0, z, 0, 0, 0));
    mx_acc(ci, __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t0, br.t0, z, 0, 0, 0));
    // order 1
    f4_t e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t0, br.t1, z, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t1, br.t0, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t0, bi.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t1, bi.t0, e, 0, 0, 0);
    mx_acc(cr, e);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t0, bi.t1, z, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t1, bi.t0, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t0, br.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t1, br.t0, e, 0, 0, 0);
    mx_acc(ci, e);
    // order 2
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t0, br.t2, z, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t1, br.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t2, br.t0, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t0, bi.t2, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t1, bi.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t2, bi.t0, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t1, br.t2, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t2, br.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t1, bi.t2, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai.t2, bi.t1, e, 0, 0, 0);
    mx_acc(cr, e);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t0, bi.t2, z, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t1, bi.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t2, bi.t0, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t0, br.t2, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t1, br.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t2, br.t0, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t1, bi.t2, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ar.t2, bi.t1, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t1, br.t2, e, 0, 0, 0);
    e = __builtin_amdgcn_mfma_f32_16x16x4f32(ai.t2, br.t1, e, 0, 0, 0);
    mx_acc(ci, e);
  }
  const int col = tn * MX_TILE + li;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = tm * MX_TILE + 4 * lk + r;
    if (row < M && col < N) {
      const double s = rs[row] * cs[col];
      C[row + (size_t)ldc * col] = zmk(cr[r] * s, ci[r] * s);
    }
  }
}

// three bfloat16 terms of each of 8 values, as bf16 vectors (neg flips the sign bit)
struct bsp3 {
  s8_t b0, b1, b2;
};
PQ_DEV bsp3 split_bf16(const float v[8], bool neg) {
  bsp3 s;
  const unsigned short sg = neg ? 0x8000u : 0u;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const unsigned short h0 = bf16_bits(v[q]);
    const float r0 = v[q] - bf16_float(h0);
    const unsigned short h1 = bf16_bits(r0);
    const float r1 = r0 - bf16_float(h1);
    const unsigned short h2 = bf16_bits(r1);
    s.b0[q] = (short)(h0 ^ sg);
    s.b1[q] = (short)(h1 ^ sg);
    s.b2[q] = (short)(h2 ^ sg);
  }
  return s;
}
PQ_DEV f4_t mfma_bf16(s8_t a, s8_t b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, a), __builtin_bit_cast(bf8_t, b), c, 0,
                                                  0, 0);
}

__global__ __launch_bounds__(256) void k_mx_gemm_bf16(int M, int N, int K, const zd *P, int lda, const zd *Q, int ldb,
                                                      const double *rs, const double *cs, zd *C, int ldc) {
  const int lane = threadIdx.x & 63;
  const int MT = (M + MX_TILE - 1) / MX_TILE, NT = (N + MX_TILE - 1) / MX_TILE;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= MT * NT) return;
  const int tm = tile % MT, tn = tile / MT;
  const int li = lane & 15, lk = lane >> 4;
  const int i = tm * MX_TILE + li, j = tn * MX_TILE + li;
  const double ri = i < M ? 1.0 / rs[i] : 0.0, cj = j < N ? 1.0 / cs[j] : 0.0;
  double4_t cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
  const f4_t z = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    float xr[8], xi[8], yr[8], yi[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + 8 * lk + q;
      const zd a = (i < M && k < K) ? P[i + (size_t)lda * k] : zmk(0.0, 0.0);
      const zd b = (j < N && k < K) ? Q[k + (size_t)ldb * j] : zmk(0.0, 0.0);
      xr[q] = (float)(a.x * ri);
      xi[q] = (float)(a.y * ri);
      yr[q] = (float)(b.x * cj);
      yi[q] = (float)(b.y * cj);
    }
    const bsp3 ar = split_bf16(xr, false), ai = split_bf16(xi, false), nai = split_bf16(xi, true);
    const bsp3 br = split_bf16(yr, false), bi = split_bf16(yi, false);
    // Re: ar br - ai bi;  Im: ar bi + ai br; terms of order <= 2 per class
    f4_t e = mfma_bf16(ar.b0, br.b0, z);
    e = mfma_bf16(nai.b0, bi.b0, e);
    mx_acc(cr, e);
    e = mfma_bf16(ar.b0, br.b1, z);
    e = mfma_bf16(ar.b1, br.b0, e);
    e = mfma_bf16(nai.b0, bi.b1, e);
    e = mfma_bf16(nai.b1, bi.b0, e);
    e = mfma_bf16(ar.b0, br.b2, e);
    e = mfma_bf16(ar.b1, br.b1, e);
    e = mfma_bf16(ar.b2, br.b0, e);
    e = mfma_bf16(nai.b0, bi.b2, e);
    e = mfma_bf16(nai.b1, bi.b1, e);
    e = mfma_bf16(nai.b2, bi.b0, e);
    mx_acc(cr, e);
    e = mfma_bf16(ar.b0, bi.b0, z);
    e = mfma_bf16(ai.b0, br.b0, e);
    mx_acc(ci, e);
    e = mfma_bf16(ar.b0, bi.b1, z);
    e = mfma_bf16(ar.b1, bi.b0, e);
    e = mfma_bf16(ai.b0, br.b1, e);
    e = mfma_bf16(ai.b1, br.b0, e);
    e = mfma_bf16(ar.b0, bi.b2, e);
    e = mfma_bf16(ar.b1, bi.b1, e);
    e = mfma_bf16(ar.b2, bi.b0, e);
    e = mfma_bf16(ai.b0, br.b2, e);
    e = mfma_bf16(ai.b1, br.b1, e);
    e = mfma_bf16(ai.b2, br.b0, e);
    mx_acc(ci, e);
  }
  const int col = tn * MX_TILE + li;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = tm * MX_TILE + 4 * lk + r;
    if (row < M && col < N) {
      const double s = rs[row] * cs[col];
      C[row + (size_t)ldc * col] = zmk(cr[r] * s, ci[r] * s);
    }
  }
}

// C = P Q in mode 1 (FP32 three-term split) or 2 (BF16x3); scales: room for M + N doubles
extern "C" hipError_t pqmc_mx_product(int mode, int M, int N, int K, const zd *P, int lda, const zd *Q, int ldb,
                                      zd *C, int ldc, double *scales, hipStream_t s) {
  if (mode != 1 && mode != 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mx_scales, dim3((M + 3) / 4), dim3(256), 0, s, P, lda, M, K, 1, scales);
  hipLaunchKernelGGL(k_mx_scales, dim3((N + 3) / 4), dim3(256), 0, s, Q, ldb, K, N, 0, scales + M);
  const int tiles = ((M + MX_TILE - 1) / MX_TILE) * ((N + MX_TILE - 1) / MX_TILE);
  if (mode == 1)
    hipLaunchKernelGGL(k_mx_gemm_fp32, dim3((tiles + 3) / 4), dim3(256), 0, s, M, N, K, P, lda, Q, ldb, scales,
                       scales + M, C, ldc);
  else
    hipLaunchKernelGGL(k_mx_gemm_bf16, dim3((tiles + 3) / 4), dim3(256), 0, s, M, N, K, P, lda, Q, ldb, scales,
                       scales + M, C, ldc);
  return hipGetLastError();
}
