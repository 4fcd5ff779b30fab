// Equal-time measurement (ms_take_measurement, src/mod_meas.f90:210-544,
// and ms_k_space_obs :546-566) for the observables registered in
// input/model/EPSOCZ.nml:262-283, as fused reductions over G.
//
// Output layout (total_width complex): BF_KE, BF_PE, BF_X, then for each of
// the 19 field observables in registration order its N_cell displacement
// values followed by its N_cell k-space values.  Jxx_Jyy and Jxy_Jyx are
// registered but never recorded by the reference (only their unregistered
// Hartree-only variants are), so they stay zero.
#include "pqmc_common.h"

enum {
  F_BFxBFx, F_BFxBFy, F_BFyBFy, F_BFxyBFxy, F_JxxJxx, F_JxyJxy, F_JyxJyx, F_JyyJyy, F_JxxJyx, F_JxyJyy,
  F_JxxJyy, F_JxyJyx, F_cJxy, F_cJyx, F_den, F_SC, F_Sz, F_Sx, F_Sy, N_FOBS
};

struct meas_ctx {
  const zd *G;
  int Ns;
  __device__ __forceinline__ zd gh(int i, int j) const { return G[i + (size_t)Ns * j]; }
  __device__ __forceinline__ zd g(int i, int j) const {
    zd v = zneg(G[i + (size_t)Ns * j]);
    if (i == j) v.x += 1.0;
    return v;
  }
};

// sum_i vl_i sum_j mat(i,j) vr_j, mat(i,j) = gh(l2,l1) gh(r2,r1) + gh(r2,l1) g(l2,r1)
// (cal_wick_contraction, src/mod_meas.f90:165-183)
template <int N>
PQ_DEV zd wick_contract(const meas_ctx &c, const int (&ls)[N][2], const int (&rs)[N][2], const zd *vl,
                        const zd *vr) {
  zd tot = zmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    zd t = zmk(0.0, 0.0);
    const zd hl = c.gh(ls[i][1], ls[i][0]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      zd m = zmul(hl, c.gh(rs[j][1], rs[j][0]));
      m = zadd(m, zmul(c.gh(rs[j][1], ls[i][0]), c.g(ls[i][1], rs[j][0])));
      t = zadd(t, zmul(m, vr[j]));
    }
    tot = zadd(tot, zmul(vl[i], t));
  }
  return tot;
}

PQ_DEV void spin_current_lists(int a, int b, int (&l)[4][2]) {
  l[0][0] = 2 * a; l[0][1] = 2 * b + 1;
  l[1][0] = 2 * a + 1; l[1][1] = 2 * b;
  l[2][0] = 2 * b; l[2][1] = 2 * a + 1;
  l[3][0] = 2 * b + 1; l[3][1] = 2 * a;
}
PQ_DEV void spin_momentum_lists(int a, int (&l)[4][2]) {
  l[0][0] = 2 * a; l[0][1] = 2 * a;
  l[1][0] = 2 * a; l[1][1] = 2 * a + 1;
  l[2][0] = 2 * a + 1; l[2][1] = 2 * a;
  l[3][0] = 2 * a + 1; l[3][1] = 2 * a + 1;
}
PQ_DEV void phys_lists(int a, int b, int (&l)[8][2]) {
  l[0][0] = 2 * a; l[0][1] = 2 * b + 1;
  l[1][0] = 2 * a + 1; l[1][1] = 2 * b;
  l[2][0] = 2 * b; l[2][1] = 2 * a + 1;
  l[3][0] = 2 * b + 1; l[3][1] = 2 * a;
  l[4][0] = 2 * a; l[4][1] = 2 * b;
  l[5][0] = 2 * a + 1; l[5][1] = 2 * b + 1;
  l[6][0] = 2 * b; l[6][1] = 2 * a;
  l[7][0] = 2 * b + 1; l[7][1] = 2 * a + 1;
}

// add_boundary_phase (src/mod_meas.f90:199-208)
PQ_DEV void bphase(zd *v, bool cross, zd p) {
  if (!cross) return;
  v[0] = zmul(v[0], p); v[1] = zmul(v[1], p);
  v[2] = zmul(v[2], zconj(p)); v[3] = zmul(v[3], zconj(p));
}

PQ_DEV int cell_of(int x, int y, int L1, int L2) {
  x %= L1; if (x < 0) x += L1;
  y %= L2; if (y < 0) y += L2;
  return x + L1 * y;
}

#define MEAS_THREADS 256

// the launch (GROUP of k_meas_pairs) that computes field observable q
// (Jxx_Jyy and Jxy_Jyx are registered but never recorded: group 0 writes
// their zeros)
PQ_DEV int meas_group(int q) {
  if (q <= F_JxyJxy || q == F_JxxJyy || q == F_JxyJyx) return 0;
  if (q == F_JyxJyx || q == F_JyyJyy) return 4;
  if (q == F_JxxJyx || q == F_JxyJyy) return 5;
  if (q == F_cJyx) return 2;
  if (q == F_cJxy) return 3;
  return 1;
}

// one workgroup per displacement d; the pair (c1, c1 + d) has relative index d
// The observables are split over six launches (GROUP 0: bosons and the
// x-bond spin currents, 4: y-bond, 5: x-bond against y-bond, 1: spin,
// density and pairing, 2 and 3: the two 8 x 8 physical currents), so that no
// launch holds more G values in registers than fit: one launch for all of
// them spilled 4 KB per lane to scratch, which every hardware queue running
// it must back.  Each observable keeps its own per-thread accumulation and
// reduction order.
template <int GROUP>
__global__ __launch_bounds__(MEAS_THREADS) void k_meas_pairs(model_t md, const zd *G, const double *x, int time,
                                                             zd *out) {
  const int L1 = md.L1, L2 = md.L2, Nc = md.Ncell;
  const int d = blockIdx.x;
  const int dx = d % L1, dy = d / L1;
  const meas_ctx mc{G, md.Ns};
  const double *xt = x + (size_t)md.Nb * (time - 1);
  const double hop = md.hop, lam = md.lam, inv = 1.0 / (double)Nc;
  zd acc[N_FOBS];
  for (int q = 0; q < N_FOBS; ++q) acc[q] = zmk(0.0, 0.0);
  const zd vjx[4] = {zmk(0.0, hop), zmk(0.0, hop), zmk(0.0, -hop), zmk(0.0, -hop)};  // -hop*[-i,-i,i,i]
  const zd vjy[4] = {zmk(hop, 0.0), zmk(-hop, 0.0), zmk(-hop, 0.0), zmk(hop, 0.0)};   // -hop*[-1,1,1,-1]
  for (int c1 = threadIdx.x; c1 < Nc; c1 += MEAS_THREADS) {
    const int x1 = c1 % L1, y1 = c1 / L1;
    const int c2 = cell_of(x1 + dx, y1 + dy, L1, L2);
    const int x2 = c2 % L1, y2 = c2 / L1;
    const int c1x = cell_of(x1 + 1, y1, L1, L2), c1y = cell_of(x1, y1 + 1, L1, L2);
    const int c2x = cell_of(x2 + 1, y2, L1, L2), c2y = cell_of(x2, y2 + 1, L1, L2);
    const bool cb1x = (x1 == L1 - 1), cb1y = (y1 == L2 - 1), cb2x = (x2 == L1 - 1), cb2y = (y2 == L2 - 1);
    const double b1x = xt[2 * c1], b1y = xt[2 * c1 + 1], b2x = xt[2 * c2], b2y = xt[2 * c2 + 1];
    int ls[4][2], rs[4][2];
    zd vl[8], vr[8];
    if constexpr (GROUP == 0) {
    acc[F_BFxBFx] = zadd(acc[F_BFxBFx], zmk(b1x * b2x / Nc, 0.0));
    acc[F_BFxBFy] = zadd(acc[F_BFxBFy], zmk(b1x * b2y / Nc, 0.0));
    acc[F_BFyBFy] = zadd(acc[F_BFyBFy], zmk(b1y * b2y / Nc, 0.0));
    acc[F_BFxyBFxy] = zadd(acc[F_BFxyBFxy], zmk(b1x * b1y * b2x * b2y / Nc, 0.0));
    // x-bond spin currents
    spin_current_lists(c1, c1x, ls);
    spin_current_lists(c2, c2x, rs);
    for (int q = 0; q < 4; ++q) { vl[q] = vjx[q]; vr[q] = vjx[q]; }
    bphase(vl, cb1x, md.bc[0]); bphase(vr, cb2x, md.bc[0]);
    acc[F_JxxJxx] = zadd(acc[F_JxxJxx], zscale(wick_contract<4>(mc, ls, rs, vl, vr), inv));
    for (int q = 0; q < 4; ++q) { vl[q] = vjy[q]; vr[q] = vjy[q]; }
    bphase(vl, cb1x, md.bc[0]); bphase(vr, cb2x, md.bc[0]);
    acc[F_JxyJxy] = zadd(acc[F_JxyJxy], zscale(wick_contract<4>(mc, ls, rs, vl, vr), inv));
    }
    if constexpr (GROUP == 4) {
    // y-bond spin currents
    spin_current_lists(c1, c1y, ls);
    spin_current_lists(c2, c2y, rs);
    for (int q = 0; q < 4; ++q) { vl[q] = vjx[q]; vr[q] = vjx[q]; }
    bphase(vl, cb1y, md.bc[1]); bphase(vr, cb2y, md.bc[1]);
    acc[F_JyxJyx] = zadd(acc[F_JyxJyx], zscale(wick_contract<4>(mc, ls, rs, vl, vr), inv));
    for (int q = 0; q < 4; ++q) { vl[q] = vjy[q]; vr[q] = vjy[q]; }
    bphase(vl, cb1y, md.bc[1]); bphase(vr, cb2y, md.bc[1]);
    // Jyy_Jyy contracts with vec_l on both sides (src/mod_meas.f90:353)
    acc[F_JyyJyy] = zadd(acc[F_JyyJyy], zscale(wick_contract<4>(mc, ls, rs, vl, vl), inv));
    }
    if constexpr (GROUP == 5) {
    // x-bond vs y-bond
    spin_current_lists(c1, c1x, ls);
    spin_current_lists(c2, c2y, rs);
    for (int q = 0; q < 4; ++q) { vl[q] = vjx[q]; vr[q] = vjx[q]; }
    bphase(vl, cb1x, md.bc[0]); bphase(vr, cb2y, md.bc[1]);
    acc[F_JxxJyx] = zadd(acc[F_JxxJyx], zscale(wick_contract<4>(mc, ls, rs, vl, vr), inv));
    for (int q = 0; q < 4; ++q) { vl[q] = vjy[q]; vr[q] = vjy[q]; }
    bphase(vl, cb1x, md.bc[0]); bphase(vr, cb2y, md.bc[1]);
    acc[F_JxyJyy] = zadd(acc[F_JxyJyy], zscale(wick_contract<4>(mc, ls, rs, vl, vr), inv));
    }
    if constexpr (GROUP == 1) {
    // spin momentum
    spin_momentum_lists(c1, ls);
    spin_momentum_lists(c2, rs);
    {
      zd v[4] = {zmk(0.5, 0), zmk(0, 0), zmk(0, 0), zmk(0.5, 0)};
      acc[F_den] = zadd(acc[F_den], zscale(wick_contract<4>(mc, ls, rs, v, v), inv));
    }
    {
      const int s1u = 2 * c1, s1d = 2 * c1 + 1, s2u = 2 * c2, s2d = 2 * c2 + 1;
      zd sc = zsub(zmul(mc.gh(s2u, s1u), mc.gh(s2d, s1d)), zmul(mc.gh(s2d, s1u), mc.gh(s2u, s1d)));
      acc[F_SC] = zadd(acc[F_SC], zscale(sc, inv));
    }
    {
      zd v[4] = {zmk(0.5, 0), zmk(0, 0), zmk(0, 0), zmk(-0.5, 0)};
      acc[F_Sz] = zadd(acc[F_Sz], zscale(wick_contract<4>(mc, ls, rs, v, v), inv));
      zd w[4] = {zmk(0, 0), zmk(0.5, 0), zmk(0.5, 0), zmk(0, 0)};
      acc[F_Sx] = zadd(acc[F_Sx], zscale(wick_contract<4>(mc, ls, rs, w, w), inv));
      zd y[4] = {zmk(0, 0), zmk(0, -0.5), zmk(0, 0.5), zmk(0, 0)};
      acc[F_Sy] = zadd(acc[F_Sy], zscale(wick_contract<4>(mc, ls, rs, y, y), inv));
    }
    }
    // physical currents (8x8): y-bond spin-x with kinetic part, x-bond spin-y
    if constexpr (GROUP == 2) {
      int l8[8][2], r8[8][2];
      phys_lists(c1, c1y, l8);
      phys_lists(c2, c2y, r8);
      for (int q = 0; q < 4; ++q) vl[q] = (q < 2) ? zmk(0.0, -hop) : zmk(0.0, hop);
      for (int q = 4; q < 8; ++q) vl[q] = zmk(lam * b1y, 0.0);
      for (int q = 0; q < 4; ++q) vr[q] = vl[q];
      for (int q = 4; q < 8; ++q) vr[q] = zmk(lam * b2y, 0.0);
      bphase(vl, cb1y, md.bc[1]); bphase(vr, cb2y, md.bc[1]);
      bphase(vl + 4, cb1y, md.bc[1]); bphase(vr + 4, cb2y, md.bc[1]);
      acc[F_cJyx] = zadd(acc[F_cJyx], zscale(wick_contract<8>(mc, l8, r8, vl, vr), inv));
    }
    if constexpr (GROUP == 3) {
      int l8[8][2], r8[8][2];
      phys_lists(c1, c1x, l8);
      phys_lists(c2, c2x, r8);
      const double sy[4] = {-1.0, 1.0, 1.0, -1.0};
      for (int q = 0; q < 4; ++q) vl[q] = zmk(hop * sy[q], 0.0);
      for (int q = 4; q < 8; ++q) vl[q] = zmk(lam * b1x, 0.0);
      for (int q = 0; q < 4; ++q) vr[q] = vl[q];
      for (int q = 4; q < 8; ++q) vr[q] = zmk(lam * b2x, 0.0);
      bphase(vl, cb1x, md.bc[0]); bphase(vr, cb2x, md.bc[0]);
      bphase(vl + 4, cb1x, md.bc[0]); bphase(vr + 4, cb2x, md.bc[0]);
      acc[F_cJxy] = zadd(acc[F_cJxy], zscale(wick_contract<8>(mc, l8, r8, vl, vr), inv));
    }
  }
  __shared__ double sr[MEAS_THREADS / 64][N_FOBS], si[MEAS_THREADS / 64][N_FOBS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int q = 0; q < N_FOBS; ++q) {
    const double vr_ = wave_sum(acc[q].x), vi_ = wave_sum(acc[q].y);
    if (lane == 0) { sr[wv][q] = vr_; si[wv][q] = vi_; }
  }
  __syncthreads();
  if (threadIdx.x < N_FOBS && meas_group(threadIdx.x) == GROUP) {
    const int q = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int w = 0; w < MEAS_THREADS / 64; ++w) { a += sr[w][q]; b += si[w][q]; }
    out[3 + (size_t)q * 2 * Nc + d] = zmk(a, b);
  }
}

// BF_KE, BF_PE, BF_X (per-cell sums /N_cell)
__global__ void k_meas_scalars(model_t md, const double *x, int time, zd *out) {
  const int Nc = md.Ncell;
  const double *xt = x + (size_t)md.Nb * (time - 1);
  const double *xt1 = x + (size_t)md.Nb * time;
  __shared__ double s[3][256];
  double ke = 0.0, pe = 0.0, bx = 0.0;
  for (int c = threadIdx.x; c < Nc; c += 256) {
    const double a = xt[2 * c], b = xt[2 * c + 1];
    ke += (0.5 * md.D * (a * a + b * b) - md.biased * (a + b)) / Nc;
    const double da = a - xt1[2 * c], db = b - xt1[2 * c + 1];
    pe += (-0.5 * md.M * (da * da + db * db) / (md.dtau * md.dtau) + 1.0 / (2.0 * md.dtau) * 2) / Nc;
    bx += (a + b) / Nc;
  }
  s[0][threadIdx.x] = ke; s[1][threadIdx.x] = pe; s[2][threadIdx.x] = bx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int q = 0; q < 3; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[threadIdx.x] = zmk(s[threadIdx.x][0], 0.0);
}

// k_data(k) = sum_i exp(i FBZ(k).r_i) r(i) / N_cell for every field observable
__global__ void k_meas_kspace(int Nc, const zd *phase /*Nc x Nc, (k, i)*/, zd *out) {
  const int q = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Nc) return;
  const zd *r = out + 3 + (size_t)q * 2 * Nc;
  zd s = zmk(0.0, 0.0);
  for (int i = 0; i < Nc; ++i) s = zadd(s, zmul(phase[k + (size_t)Nc * i], r[i]));
  out[3 + (size_t)q * 2 * Nc + Nc + k] = zscale(s, 1.0 / (double)Nc);
}

extern "C" hipError_t pqmc_launch_measure(const model_t *md, const zd *G, const double *x, int time,
                                          const zd *phase, zd *out, hipStream_t s) {
  hipLaunchKernelGGL(k_meas_pairs<0>, dim3(md->Ncell), dim3(MEAS_THREADS), 0, s, *md, G, x, time, out);
  hipLaunchKernelGGL(k_meas_pairs<1>, dim3(md->Ncell), dim3(MEAS_THREADS), 0, s, *md, G, x, time, out);
  hipLaunchKernelGGL(k_meas_pairs<2>, dim3(md->Ncell), dim3(MEAS_THREADS), 0, s, *md, G, x, time, out);
  hipLaunchKernelGGL(k_meas_pairs<3>, dim3(md->Ncell), dim3(MEAS_THREADS), 0, s, *md, G, x, time, out);
  hipLaunchKernelGGL(k_meas_pairs<4>, dim3(md->Ncell), dim3(MEAS_THREADS), 0, s, *md, G, x, time, out);
  hipLaunchKernelGGL(k_meas_pairs<5>, dim3(md->Ncell), dim3(MEAS_THREADS), 0, s, *md, G, x, time, out);
  hipLaunchKernelGGL(k_meas_scalars, dim3(1), dim3(256), 0, s, *md, x, time, out);
  hipLaunchKernelGGL(k_meas_kspace, dim3((md->Ncell + 63) / 64, N_FOBS), dim3(64), 0, s, md->Ncell, phase, out);
  return hipGetLastError();
}
