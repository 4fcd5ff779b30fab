// Shared device helpers for the gfx950 projector-DQMC kernels.
// Complex numbers are double2 (x = real, y = imaginary), column-major, as
// complex(dp) arrays in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double2 zd;
typedef double double4_t __attribute__((ext_vector_type(4)));

#define PQ_DEV static __device__ __forceinline__

PQ_DEV zd zmk(double r, double i) { return make_double2(r, i); }
PQ_DEV zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
PQ_DEV zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
PQ_DEV zd zmul(zd a, zd b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
PQ_DEV zd zconj(zd a) { return zmk(a.x, -a.y); }
PQ_DEV zd zneg(zd a) { return zmk(-a.x, -a.y); }
PQ_DEV zd zscale(zd a, double s) { return zmk(a.x * s, a.y * s); }
// a * conj(b)
PQ_DEV zd zmulc(zd a, zd b) { return zmk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
// conj(a) * b
PQ_DEV zd zcmul(zd a, zd b) { return zmk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
PQ_DEV zd zdiv(zd a, zd b) {
  double d = b.x * b.x + b.y * b.y;
  return zmk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
PQ_DEV double zabs(zd a) { return hypot(a.x, a.y); }
PQ_DEV zd zlog(zd a) { return zmk(log(hypot(a.x, a.y)), atan2(a.y, a.x)); }

// Compact bond exponential.  For the EPSOCZ bonds
//   E = c I4 + [[0, m],[m^H, 0]],  E^-1 = c I4 - [[0, m],[m^H, 0]]
// with c = cosh(n) and m = h sinh(n)/n (2x2, column-major m[a + 2b]).
struct __align__(16) ekv_t {
  double c, pad;
  zd m[4];
};

// Model constants shared by all kernels of a chain.
struct model_t {
  int Ns, ne, npla, Nb, Lt, Ncell, L1, L2;
  double dtau, M, D, biased, jump, hop, lam;
  int ncopy, tr_weight;
  zd kcoe[4], vcoe[4];
  zd k12[4][4], v12[4][4];  // Kmatrix(1:2,3:4), Vmatrix(1:2,3:4) per family
  zd bc[2];                 // lattice boundary phases
};

// Arguments of k_flush_wrap (the delayed flush fused with a two-way wrap).
struct fw_args {
  model_t md;
  zd *G;
  const zd *U, *VT;     // flush factors (row-major, ld Ns), family-order rows of the flush family
  const int *rank;      // device rank r (null: no flush)
  const int *pos_ffl;   // [Ns] site -> family-order row of the flush family
  const int *tsites;    // [npla*4] sites of the tiling family
  int wrap;             // 0 none, 1 left: E G E^-1, 2 right: E^-1 G E
  const ekv_t *ekv_w;   // [npla] exponentials of the wrap family at the wrap slice
  // accumulate the wrap family's field gradient from G before the wrap
  // (null: none), bond_field_grad of k_wrap.hip
  zd *grad;
  const double *x;
  const int *bfi, *cross;
  const zd *ph;
  int f_wrap, t_wrap;
};

// get_expKV_epsocz (src/mod_phonon_field.f90:534-569) in compact form.
// ph points at the bond's BC_phases(1:2,3:4) (4 values) or is null.
PQ_DEV void expkv_compact(const model_t &md, int f, double x, const zd *ph, ekv_t &e) {
  double ak = hypot(md.kcoe[f].x, md.kcoe[f].y);
  double av = hypot(md.vcoe[f].x, md.vcoe[f].y);
  double t = av * x;
  double norm = md.dtau * sqrt(ak * ak + t * t);
  zd vx = zscale(md.vcoe[f], x);
  double ch = cosh(norm), sh = sinh(norm);
  e.c = ch;
  e.pad = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    zd s = zadd(zmul(md.kcoe[f], md.k12[f][q]), zmul(vx, md.v12[f][q]));
    s = zscale(s, -md.dtau);
    if (ph) s = zmul(s, ph[q]);
    zd t2 = zscale(s, sh);
    e.m[q] = zmk(t2.x / norm, t2.y / norm);
  }
}

// get_expKV_epsoc (src/mod_phonon_field.f90:572-604): the EPSOC model's
// spin-up 2x2 bond, E = cosh(n) I2 + [[0, e12],[conj(e12), 0]] with
// e12 = +-s sinh(n)/n, s = -dtau (K_coe K(1,2) + V_coe x V(1,2)) times the
// bond's BC_phases(1,2) when it crosses the boundary (ph, else null),
// n = dtau sqrt(|K_coe|^2 + (|V_coe| x)^2); inv flips the sign of s.
// out: 2x2 column-major.
PQ_DEV void expkv_epsoc(double dtau, zd kc, zd vc, zd k12, zd v12, double x, const zd *ph, bool inv, zd out[4]) {
  const double ak = hypot(kc.x, kc.y), av = hypot(vc.x, vc.y);
  const double t = av * x;
  const double norm = dtau * sqrt(ak * ak + t * t);
  zd s = zadd(zmul(kc, k12), zmul(zscale(vc, x), v12));
  s = zscale(s, -dtau);
  if (ph) s = zmul(s, *ph);
  const double ch = cosh(norm), sh = sinh(norm);
  if (inv) s = zneg(s);
  const zd t2 = zscale(s, sh);
  const zd e12 = zmk(t2.x / norm, t2.y / norm);
  out[0] = zmk(ch, 0.0);
  out[1] = zconj(e12);
  out[2] = e12;
  out[3] = zmk(ch, 0.0);
}

// y = E v for a 4-vector (inv flips the sign of the off-diagonal blocks).
PQ_DEV void ekv_apply(const ekv_t &e, bool inv, const zd v[4], zd y[4]) {
  zd m00 = e.m[0], m10 = e.m[1], m01 = e.m[2], m11 = e.m[3];
  if (inv) { m00 = zneg(m00); m10 = zneg(m10); m01 = zneg(m01); m11 = zneg(m11); }
  // rows 1,2: c v + m (v3,v4); rows 3,4: c v + m^H (v1,v2)
  y[0] = zadd(zadd(zscale(v[0], e.c), zmul(m00, v[2])), zmul(m01, v[3]));
  y[1] = zadd(zadd(zscale(v[1], e.c), zmul(m10, v[2])), zmul(m11, v[3]));
  y[2] = zadd(zadd(zcmul(m00, v[0]), zcmul(m10, v[1])), zscale(v[2], e.c));
  y[3] = zadd(zadd(zcmul(m01, v[0]), zcmul(m11, v[1])), zscale(v[3], e.c));
}

// y = v E for a row 4-vector.
PQ_DEV void ekv_apply_right(const ekv_t &e, bool inv, const zd v[4], zd y[4]) {
  zd m00 = e.m[0], m10 = e.m[1], m01 = e.m[2], m11 = e.m[3];
  if (inv) { m00 = zneg(m00); m10 = zneg(m10); m01 = zneg(m01); m11 = zneg(m11); }
  // (v E)_a = sum_k v_k E(k,a); E(3,1)=conj(m00), E(4,1)=conj(m01), E(3,2)=conj(m10), E(4,2)=conj(m11)
  y[0] = zadd(zadd(zscale(v[0], e.c), zmulc(v[2], m00)), zmulc(v[3], m01));
  y[1] = zadd(zadd(zscale(v[1], e.c), zmulc(v[2], m10)), zmulc(v[3], m11));
  y[2] = zadd(zadd(zmul(v[0], m00), zmul(v[1], m10)), zscale(v[2], e.c));
  y[3] = zadd(zadd(zmul(v[0], m01), zmul(v[1], m11)), zscale(v[3], e.c));
}

// Dense 4x4 (column-major) from the compact form.
PQ_DEV void ekv_dense(const ekv_t &e, bool inv, zd E[16]) {
#pragma unroll
  for (int q = 0; q < 16; ++q) E[q] = zmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < 4; ++i) E[i + 4 * i] = zmk(e.c, 0.0);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      zd v = inv ? zneg(e.m[a + 2 * b]) : e.m[a + 2 * b];
      E[a + 4 * (2 + b)] = v;
      E[(2 + b) + 4 * a] = zconj(v);
    }
}

// Entry (i, k) of E (inv: E^-1) from the compact form; i and k should be
// compile-time constants so this folds to a select.
PQ_DEV zd ekv_entry(const ekv_t &e, bool inv, int i, int k) {
  if (i == k) return zmk(e.c, 0.0);
  if (i < 2 && k >= 2) {
    const zd v = e.m[i + 2 * (k - 2)];
    return inv ? zneg(v) : v;
  }
  if (i >= 2 && k < 2) {
    const zd v = zconj(e.m[k + 2 * (i - 2)]);
    return inv ? zneg(v) : v;
  }
  return zmk(0.0, 0.0);
}

// Wave-wide sum through DPP lane moves (quad permutes, half-row and row
// mirrors, then the row broadcasts 15 and 31), read back from lane 63 so
// every lane gets the same value.  About 2.5x cheaper than a butterfly of
// ds_bpermute shuffles on gfx950.
template <int CTRL, int ROWMASK>
PQ_DEV double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}
PQ_DEV double wave_sum(double v) {
  v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xF>(v);  // row_mirror
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Three wave-wide sums with their DPP steps interleaved (the same operations
// as wave_sum on each, so the same results).
PQ_DEV void wave_sum3(double &a, double &b, double &c) {
#define PQ_STEP3(CTRL, ROWMASK)          \
  {                                      \
    const double a_ = dpp_move<CTRL, ROWMASK>(a); \
    const double b_ = dpp_move<CTRL, ROWMASK>(b); \
    const double c_ = dpp_move<CTRL, ROWMASK>(c); \
    a += a_;                             \
    b += b_;                             \
    c += c_;                             \
  }
  PQ_STEP3(0xB1, 0xF)
  PQ_STEP3(0x4E, 0xF)
  PQ_STEP3(0x141, 0xF)
  PQ_STEP3(0x140, 0xF)
  PQ_STEP3(0x142, 0xA)
  PQ_STEP3(0x143, 0xC)
#undef PQ_STEP3
  a = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(a), 63), __builtin_amdgcn_readlane(__double2loint(a), 63));
  b = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(b), 63), __builtin_amdgcn_readlane(__double2loint(b), 63));
  c = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(c), 63), __builtin_amdgcn_readlane(__double2loint(c), 63));
}

// Two wave-wide sums with their DPP steps interleaved (the same operations
// as wave_sum on each).
PQ_DEV void wave_sum2(double &a, double &b) {
#define PQ_STEP2(CTRL, ROWMASK)                   \
  {                                               \
    const double a_ = dpp_move<CTRL, ROWMASK>(a); \
    const double b_ = dpp_move<CTRL, ROWMASK>(b); \
    a += a_;                                      \
    b += b_;                                      \
  }
  PQ_STEP2(0xB1, 0xF)
  PQ_STEP2(0x4E, 0xF)
  PQ_STEP2(0x141, 0xF)
  PQ_STEP2(0x140, 0xF)
  PQ_STEP2(0x142, 0xA)
  PQ_STEP2(0x143, 0xC)
#undef PQ_STEP2
  a = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(a), 63), __builtin_amdgcn_readlane(__double2loint(a), 63));
  b = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(b), 63), __builtin_amdgcn_readlane(__double2loint(b), 63));
}

PQ_DEV double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// v_mfma_f64_16x16x4_f64: D(16x16) += A(16x4) B(4x16).
// Lane l supplies A[l%16][l/16] and B[l/16][l%16]; D lane l holds rows
// (l/16) + 4*i (i = 0..3) of column l%16 (measured on gfx950).
PQ_DEV double4_t mfma_f64(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
