// One-launch blocked Householder QR of A (m x n, n <= m <= 1024) for the
// stabilization's QDR factorizations (qdr, src/mod_matrixlib.f90:323-334 and
// :452-467: ZGEQRF then ZUNGQR).  Outputs as the panel-by-panel form
// (k_qr_panel + k_apply_refl): R in A's upper triangle with beta on the
// diagonal and the reflectors below it, V (unit diagonal, zeros above) in its
// own array, T of each panel's block reflector (16 x 16, ld 16, zero padded),
// tau, beta and 1 + |x|^2 |1/(alpha-beta)|^2 = v^H v per column, and Q
// (m x n) explicit when asked for.  zgeqr2 / zlarfg / zlarft conventions:
// beta = -sign(Re alpha) |(alpha, x)|, tau = ((beta - Re alpha)/beta,
// -Im alpha/beta), v = x / (alpha - beta); H^H = I - conj(tau) v v^H is
// applied from the left.
//
// Workgroup b owns panel b, columns [b nb, b nb + nb), and keeps them in LDS
// for the whole launch:
//  1. for each p < b, it applies H_p^H = I - V_p T_p^H V_p^H to its columns
//     (FP64 MFMA): V_p^H A once panel p's V flag is up, the rest once its T
//     flag is;
//  2. it factors its panel, one wave per column and one barrier per column:
//     every wave forms the column's norm and its own product with the column
//     and derives the reflector itself, so no wave waits on another's scalars;
//  3. it publishes V_b with device-scope stores, waits for them and raises
//     the V flag, then forms T_b and does the same for it;
//  4. it writes its columns of A, then forms Q's columns [b nb, b nb + nb) as
//     H_0 H_1 ... H_b E_b (H_p with p > b leaves them unchanged).
// The critical path is the chain of panels (apply the previous reflector,
// factor, publish); the trailing updates of later panels and the formation of
// Q run beside it.
//
// A workgroup takes its panel from an atomic ticket when it starts, so panel
// b waits only on panels whose workgroups are already running: the launch
// completes whatever else holds the compute units, with no co-residency
// requirement.  Flags hold the launch's epoch, so nothing is cleared between
// launches.  Waits are bounded (0.5 s of wall clock); an expired wait marks
// the launch aborted, counts itself in *timeouts and ends every workgroup.
//
// Cross-XCD visibility follows the family pass: V and T are written only with
// device-scope stores and read only after their flag, each address once per
// launch, and the caches are invalidated at kernel start, so plain L2-cached
// loads serve the readers.
#include <mutex>

#include "k_qrflow.h"

#define QF_THREADS 1024
#define QF_WAVES (QF_THREADS / 64)
#define QF_WAIT_TICKS 50000000ull  // 0.5 s of the 100 MHz wall clock
#define QF_PART 4352               // bytes of one wave's partial 16 x 16 tile (re and im, ld 17)

static __device__ __forceinline__ unsigned long long qf_ld(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void qf_st(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void qf_st_z(zd *p, zd v) {
  double *d = reinterpret_cast<double *>(p);
  __hip_atomic_store(d, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(d + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// v from lane l (uniform) to every lane
static __device__ __forceinline__ zd rdlane_zq(zd v, int l) {
  return zmk(__hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v.x), l),
                              __builtin_amdgcn_readlane(__double2loint(v.x), l)),
             __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v.y), l),
                              __builtin_amdgcn_readlane(__double2loint(v.y), l)));
}

// Wait (all threads) until flag word w holds this launch's epoch.  Returns
// false when the launch is aborted.
static __device__ __forceinline__ bool qf_wait(const qf_args &a, int w, int *sStop) {
  if (threadIdx.x == 0) {
    unsigned long long *f = &a.sync[w];
    const unsigned long long t0 = wall_clock64();
    int st = 0;
    for (unsigned it = 1;; ++it) {
      if (qf_ld(f) >= a.epoch) break;
      if ((it & 31) == 0) {
        if (qf_ld(&a.sync[1]) == a.epoch) {
          st = 1;
          break;
        }
        if (wall_clock64() - t0 > QF_WAIT_TICKS) {
          qf_st(&a.sync[1], a.epoch);
          atomicAdd(a.timeouts, 1ull);
          st = 1;
          break;
        }
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *sStop = st;
  }
  __syncthreads();
  return *sStop == 0;
}

// Wait (all threads) until count word w reaches target.  Returns false when
// the launch is aborted.
static __device__ __forceinline__ bool qf_wait_count(const qf_args &a, int w, unsigned long long target, int *sStop) {
  if (threadIdx.x == 0) {
    unsigned long long *f = &a.sync[w];
    const unsigned long long t0 = wall_clock64();
    int st = 0;
    for (unsigned it = 1;; ++it) {
      if (qf_ld(f) >= target) break;
      if ((it & 31) == 0) {
        if (qf_ld(&a.sync[1]) == a.epoch) {
          st = 1;
          break;
        }
        if (wall_clock64() - t0 > QF_WAIT_TICKS) {
          qf_st(&a.sync[1], a.epoch);
          atomicAdd(a.timeouts, 1ull);
          st = 1;
          break;
        }
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *sStop = st;
  }
  __syncthreads();
  return *sStop == 0;
}

// Rows [r0, m) of one LDS column times s, eight rows per lane at a time
// (one wave).
static __device__ __forceinline__ void qf_scale_col(zd *col, int r0, int m, zd s) {
  const int lane = threadIdx.x & 63;
  for (int base = r0 + lane; base < m; base += 8 * 64) {
    zd v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = base + 64 * i;
      v[i] = r < m ? col[r] : zmk(0.0, 0.0);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = base + 64 * i;
      if (r < m) col[r] = zmul(v[i], s);
    }
  }
}

// sA (m x nb, ld m, the workgroup's ncol columns) <- H_p^H sA (trans = 1) or
// H_p sA (trans = 0) on rows [p nb, m), H_p = I - V_p T_p V_p^H:
//   W = V_p^H sA  (2 npart waves split the rows, four 4-row chunks per wave
//                  in flight; the upper waves' partial tiles are folded into
//                  the lower ones' in sPart, then summed into sW),
//   W2 = op(T_p) W  (into sPart),
//   sA -= V_p W2  (16-row tiles over all waves).
// V_p must be published; T_p is waited for (flag word tflag, none if < 0)
// after the first product and its sums.  Returns false when the launch is
// aborted.
static __device__ __forceinline__ bool qf_apply(const qf_args &a, zd *sA, double *sPart, double (*sWr)[17], double (*sWi)[17],
                                zd *sTp, int p, int ncol, int trans, int tflag, int *sStop) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = a.m, nb = a.nb, lm = m | 1;
  const int r0 = p * nb;
  const int npw = min(nb, a.n - r0);  // width of panel p
  const zd *Vp = a.V + (size_t)m * r0;
  const zd *Tp = a.T + 256 * (size_t)p;
  const bool vk = li < npw, ac = li < ncol;
  const int npart = a.npart, nw = min(2 * npart, QF_WAVES);
  if (wv < nw) {
    // conj(V)^T sA by three real products: P1 = Vr^T Ar, P2 = Vi^T Ai,
    // P3 = (Vr - Vi)^T (Ar + Ai); re = P1 + P2, im = P3 - P1 + P2
    double4_t p1 = {0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
    // chunk q of wave wv at rows r0 + 4 (wv + nw q)
    for (int rb = r0 + 4 * wv; rb < m; rb += 16 * nw) {
      zd vv[4], xx[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = rb + 4 * nw * u + lk;
        vv[u] = (r < m && vk) ? Vp[r + (size_t)m * li] : zmk(0.0, 0.0);
        xx[u] = (r < m && ac) ? sA[r + lm * li] : zmk(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        p1 = mfma_f64(vv[u].x, xx[u].x, p1);
        p2 = mfma_f64(vv[u].y, xx[u].y, p2);
        p3 = mfma_f64(vv[u].x - vv[u].y, xx[u].x + xx[u].y, p3);
      }
    }
    const double4_t cr0 = p1 + p2, ci0 = (p3 - p1) + p2;
    const int slot = wv < npart ? wv : wv - npart;
    double *pr = sPart + (size_t)slot * (QF_PART / 8), *pi = pr + 16 * 17;
    if (wv >= npart) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        pr[(lk + 4 * i) * 17 + li] = cr0[i];
        pi[(lk + 4 * i) * 17 + li] = ci0[i];
      }
    }
    __syncthreads();
    if (wv < npart) {
      const bool fold = wv + npart < nw;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = (lk + 4 * i) * 17 + li;
        pr[e] = fold ? cr0[i] + pr[e] : cr0[i];
        pi[e] = fold ? ci0[i] + pi[e] : ci0[i];
      }
    }
  } else {
    __syncthreads();
  }
  __syncthreads();
  if (tid < 256) {  // W(k, cc): the partial tiles summed, all loads first
    const int k = tid >> 4, cc = tid & 15;
    double xr[8], xi[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      const double *pw = sPart + (size_t)w * (QF_PART / 8);
      xr[w] = w < npart ? pw[k * 17 + cc] : 0.0;
      xi[w] = w < npart ? pw[16 * 17 + k * 17 + cc] : 0.0;
    }
    sWr[k][cc] = ((xr[0] + xr[1]) + (xr[2] + xr[3])) + ((xr[4] + xr[5]) + (xr[6] + xr[7]));
    sWi[k][cc] = ((xi[0] + xi[1]) + (xi[2] + xi[3])) + ((xi[4] + xi[5]) + (xi[6] + xi[7]));
  }
  // T_p (its barrier also closes the sums)
  if (tflag >= 0 && !qf_wait(a, tflag, sStop)) return false;
  if (tid < 256) sTp[tid] = Tp[tid];
  __syncthreads();
  double *w2 = sPart;  // W2, re then im, ld 17 (the partials are consumed)
  if (tid < 256) {
    const int kk = tid >> 4, cc = tid & 15;
    double wr = 0.0, wi = 0.0;
    // T is zero outside the panel's npw x npw block and W outside its
    // npw x ncol block, so the full 16-term sums are the products
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
      zd tk[4];
      double wr_[4], wi_[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        tk[k] = trans ? zconj(sTp[k0 + k + 16 * kk]) : sTp[kk + 16 * (k0 + k)];
        wr_[k] = sWr[k0 + k][cc];
        wi_[k] = sWi[k0 + k][cc];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        wr += tk[k].x * wr_[k] - tk[k].y * wi_[k];
        wi += tk[k].x * wi_[k] + tk[k].y * wr_[k];
      }
    }
    w2[kk * 17 + cc] = (kk < npw && cc < ncol) ? wr : 0.0;
    w2[16 * 17 + kk * 17 + cc] = (kk < npw && cc < ncol) ? wi : 0.0;
  }
  __syncthreads();
  const int nk4 = (npw + 3) >> 2;
  double w2r[4], w2i[4];
#pragma unroll
  for (int k4 = 0; k4 < 4; ++k4) {
    w2r[k4] = w2[(4 * k4 + lk) * 17 + li];
    w2i[k4] = w2[16 * 17 + (4 * k4 + lk) * 17 + li];
  }
  // 16-row tiles t0 = r0 + 16 (wv + 16 q)
  for (int tb = r0 + 16 * wv; tb < m; tb += 16 * QF_WAVES) {
    zd vv[1][4], aa[1][4];
#pragma unroll
    for (int q = 0; q < 1; ++q) {
      const int t0 = tb + 16 * QF_WAVES * q;
#pragma unroll
      for (int k4 = 0; k4 < 4; ++k4) {
        const int r = t0 + li, kc = 4 * k4 + lk;
        vv[q][k4] = (k4 < nk4 && r < m && kc < npw) ? Vp[r + (size_t)m * kc] : zmk(0.0, 0.0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = t0 + lk + 4 * i;
        aa[q][i] = (row < m && ac) ? sA[row + lm * li] : zmk(0.0, 0.0);
      }
    }
#pragma unroll
    for (int q = 0; q < 1; ++q) {
      const int t0 = tb + 16 * QF_WAVES * q;
      if (t0 < m) {  // uniform
        // (Vr + i Vi)(Wr + i Wi) by three real products: Q1 = Vr Wr,
        // Q2 = Vi Wi, Q3 = (Vr + Vi)(Wr + Wi); re = Q1 - Q2, im = Q3 - Q1 - Q2
        double4_t q1 = {0.0, 0.0, 0.0, 0.0}, q2 = q1, q3 = q1;
#pragma unroll
        for (int k4 = 0; k4 < 4; ++k4) {
          if (k4 < nk4) {  // uniform
            q1 = mfma_f64(vv[q][k4].x, w2r[k4], q1);
            q2 = mfma_f64(vv[q][k4].y, w2i[k4], q2);
            q3 = mfma_f64(vv[q][k4].x + vv[q][k4].y, w2r[k4] + w2i[k4], q3);
          }
        }
        const double4_t pr = q1 - q2, pi = (q3 - q1) - q2;
        if (ac) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = t0 + lk + 4 * i;
            if (row < m) sA[row + lm * li] = zsub(aa[q][i], zmk(pr[i], pi[i]));
          }
        }
      }
    }
  }
  __syncthreads();
  return true;
}

// Factor the panel in sA (columns 0..nbp-1, rows [j0, m)), one wave per
// column, every column in LDS (any m).  The pivot wave scales its column
// into the reflector only at the next step, once no wave reads it.  Rows go
// in batches of eight per lane, every load of a batch issued before its
// arithmetic.
static __device__ __forceinline__ void qf_factor_lds(zd *sA, int m, int j0, int nbp, zd *sTau, zd *sBeta, double *sVn2) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lm = m | 1;
  zd sck = zmk(0.0, 0.0);  // the pivot wave's 1 / (alpha - beta), kept for that store
  for (int j = 0; j < nbp; ++j) {
    const int dg = j0 + j;
    if (wv == j - 1) qf_scale_col(sA + (size_t)lm * (j - 1), dg, m, sck);
    if (wv >= j && wv < nbp) {  // uniform per wave
      const int c = wv;
      const zd *xc = sA + (size_t)lm * j;
      zd *yc = sA + (size_t)lm * c;
      double nx0 = 0.0, nx1 = 0.0, dr0 = 0.0, dr1 = 0.0, di0 = 0.0, di1 = 0.0;
      for (int base = dg + 1 + lane; base < m; base += 8 * 64) {
        zd xs[8], ys[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int r = base + 64 * i;
          xs[i] = r < m ? xc[r] : zmk(0.0, 0.0);
          ys[i] = r < m ? yc[r] : zmk(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
          nx0 += xs[i].x * xs[i].x + xs[i].y * xs[i].y;
          nx1 += xs[i + 1].x * xs[i + 1].x + xs[i + 1].y * xs[i + 1].y;
          dr0 += xs[i].x * ys[i].x + xs[i].y * ys[i].y;  // conj(x) y
          dr1 += xs[i + 1].x * ys[i + 1].x + xs[i + 1].y * ys[i + 1].y;
          di0 += xs[i].x * ys[i].y - xs[i].y * ys[i].x;
          di1 += xs[i + 1].x * ys[i + 1].y - xs[i + 1].y * ys[i + 1].x;
        }
      }
      double nx = nx0 + nx1, dr = dr0 + dr1, di = di0 + di1;
      wave_sum3(nx, dr, di);
      const zd alpha = xc[dg];
      zd t, sc, bt;
      double v2;
      if (nx == 0.0 && alpha.y == 0.0) {
        t = zmk(0.0, 0.0);
        sc = zmk(0.0, 0.0);
        bt = alpha;
        v2 = 1.0;
      } else {
        const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + nx), alpha.x);
        const double rb = 1.0 / beta;
        t = zmk((beta - alpha.x) * rb, -alpha.y * rb);
        // sc = 1 / (alpha - beta) = conj(alpha - beta) / |alpha - beta|^2
        const double er = alpha.x - beta, rd = 1.0 / (er * er + alpha.y * alpha.y);
        sc = zmk(er * rd, -alpha.y * rd);
        bt = zmk(beta, 0.0);
        v2 = 1.0 + nx * rd;
      }
      if (c == j) {
        sck = sc;
        if (lane == 0) {
          sTau[j] = t;
          sBeta[j] = bt;
          sVn2[j] = v2;
        }
      } else {
        // w_c = tau^* (a_{dg,c} + sc^* x^H a_c); a_c -= v w_c with v = (1, x sc)
        const zd acd = yc[dg];
        const zd wc = zmul(zconj(t), zadd(acd, zmul(zconj(sc), zmk(dr, di))));
        const zd f = zmul(sc, wc);
        if (lane == 0) yc[dg] = zsub(acd, wc);
        for (int base = dg + 1 + lane; base < m; base += 8 * 64) {
          zd xs[8], ys[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int r = base + 64 * i;
            xs[i] = r < m ? xc[r] : zmk(0.0, 0.0);
            ys[i] = r < m ? yc[r] : zmk(0.0, 0.0);
          }
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int r = base + 64 * i;
            if (r < m) yc[r] = zsub(ys[i], zmul(xs[i], f));
          }
        }
      }
    }
    __syncthreads();
  }
  if (wv == nbp - 1) qf_scale_col(sA + (size_t)lm * (nbp - 1), j0 + nbp, m, sck);
  __syncthreads();
}

// Reflector of the column held in registers (rows lane + 64 i) at diagonal
// row dg (one wave): zlarfg's scalars from |x|^2 (rows > dg) and alpha; the
// rows below dg scaled into v; v stored to the LDS column and tau, beta and
// v^H v to the panel's scalars (slot j).
static __device__ __forceinline__ void qf_reflector(zd (&col)[8], zd *scol, int m, int dg, int j, zd *sTau,
                                                    zd *sBeta, double *sVn2) {
  const int lane = threadIdx.x & 63;
  double n0 = 0.0, n1 = 0.0;
  zd adg = zmk(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < 8; i += 2) {
    const int r = lane + 64 * i, r1 = r + 64;
    if (r > dg && r < m) n0 += col[i].x * col[i].x + col[i].y * col[i].y;
    if (r1 > dg && r1 < m) n1 += col[i + 1].x * col[i + 1].x + col[i + 1].y * col[i + 1].y;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (i == (dg >> 6)) adg = col[i];
  const double nx = wave_sum(n0 + n1);
  const zd alpha = rdlane_zq(adg, dg & 63);
  zd t, sc, bt;
  double v2;
  if (nx == 0.0 && alpha.y == 0.0) {
    t = zmk(0.0, 0.0);
    sc = zmk(0.0, 0.0);
    bt = alpha;
    v2 = 1.0;
  } else {
    const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + nx), alpha.x);
    const double rb = 1.0 / beta;
    t = zmk((beta - alpha.x) * rb, -alpha.y * rb);
    // sc = 1 / (alpha - beta) = conj(alpha - beta) / |alpha - beta|^2
    const double er = alpha.x - beta, rd = 1.0 / (er * er + alpha.y * alpha.y);
    sc = zmk(er * rd, -alpha.y * rd);
    bt = zmk(beta, 0.0);
    v2 = 1.0 + nx * rd;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = lane + 64 * i;
    if (r > dg && r < m) {
      col[i] = zmul(col[i], sc);
      scol[r] = col[i];
    }
  }
  if (lane == 0) {
    sTau[j] = t;
    sBeta[j] = bt;
    sVn2[j] = v2;
  }
}

// The same with each wave's column in registers, eight rows per lane
// (m <= 512).  The next pivot forms its reflector as soon as its own update
// is done and leaves v (already scaled) in its LDS column with tau in the
// panel's scalars; at a step the other waves then only form v^H a_c (one
// two-value wave sum) and update their registers.
static __device__ __forceinline__ void qf_factor_regs(zd *sA, int m, int j0, int nbp, zd *sTau, zd *sBeta, double *sVn2) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lm = m | 1;
  const bool own = wv < nbp;
  zd col[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = lane + 64 * i;
    col[i] = (own && r < m) ? sA[r + (size_t)lm * wv] : zmk(0.0, 0.0);
  }
  if (wv == 0) qf_reflector(col, sA, m, j0, 0, sTau, sBeta, sVn2);
  __syncthreads();
  for (int j = 0; j < nbp; ++j) {
    const int dg = j0 + j;
    if (own && wv > j) {  // uniform per wave
      // the next pivot's wave carries the step's critical chain: its
      // instructions issue ahead of the other updates on its SIMD
      if (wv == j + 1) __builtin_amdgcn_s_setprio(3);
      const zd *vc = sA + (size_t)lm * j;
      zd vs[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int r = lane + 64 * i;
        vs[i] = (r > dg && r < m) ? vc[r] : zmk(0.0, 0.0);
      }
      double dr0 = 0.0, dr1 = 0.0, di0 = 0.0, di1 = 0.0;
#pragma unroll
      for (int i = 0; i < 8; i += 2) {
        dr0 += vs[i].x * col[i].x + vs[i].y * col[i].y;  // conj(v) a
        dr1 += vs[i + 1].x * col[i + 1].x + vs[i + 1].y * col[i + 1].y;
        di0 += vs[i].x * col[i].y - vs[i].y * col[i].x;
        di1 += vs[i + 1].x * col[i + 1].y - vs[i + 1].y * col[i + 1].x;
      }
      double dr = dr0 + dr1, di = di0 + di1;
      wave_sum2(dr, di);
      zd adg = zmk(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (i == (dg >> 6)) adg = col[i];
      const zd acd = rdlane_zq(adg, dg & 63);
      // w_c = tau^* v^H a_c with v_dg = 1; a_c -= v w_c
      const zd wc = zmul(zconj(sTau[j]), zadd(acd, zmk(dr, di)));
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int r = lane + 64 * i;
        if (r == dg) col[i] = zsub(col[i], wc);
        else if (r > dg && r < m) col[i] = zsub(col[i], zmul(vs[i], wc));
      }
      if (wv == j + 1) {
        qf_reflector(col, sA + (size_t)lm * wv, m, dg + 1, j + 1, sTau, sBeta, sVn2);
        __builtin_amdgcn_s_setprio(0);
      }
    }
    __syncthreads();
  }
  if (own) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = lane + 64 * i;
      if (r < m) sA[r + (size_t)lm * wv] = col[i];
    }
  }
  __syncthreads();
}

// Shared memory of one panel workgroup (the kernels declare it once).
struct qf_smem {
  double (*sW2r)[17], (*sW2i)[17];
  zd *sTau, *sBeta;
  double *sVn2;
  zd *sTp;
  int *sStop;
};

// Panel b of the QR a: steps 1-4 of the one-launch form.  With ycnt >= 0 the
// panel's columns are waited for first: count word ycnt + b of a.sync must
// reach ytarget (the passive workgroups of a fused launch).
static __device__ __forceinline__ void qf_panel(const qf_args &a, int b, zd *qf_lds, const qf_smem &sm, int ycnt,
                                                unsigned long long ytarget) {
  double(*sW2r)[17] = sm.sW2r;
  double(*sW2i)[17] = sm.sW2i;
  zd *sTau = sm.sTau, *sBeta = sm.sBeta, *sTp = sm.sTp;
  double *sVn2 = sm.sVn2;
  int &sStop = *sm.sStop;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m = a.m, nb = a.nb;
  const int lm = m | 1;  // LDS column stride: odd, so the 16 columns of an MFMA operand hit 16 bank groups
  zd *sA = qf_lds;
  double *sPart = reinterpret_cast<double *>(qf_lds + (size_t)lm * nb);
  const int j0 = b * nb, nbp = min(nb, a.n - j0);
  if (ycnt >= 0 && !qf_wait_count(a, ycnt + b, ytarget, &sStop)) return;
  // the panel's columns, rows [0, m); columns past the matrix are zero
  for (int e = tid; e < m * nb; e += QF_THREADS) {
    const int r = e % m, c = e / m;
    sA[r + lm * c] = c < nbp ? a.A[r + (size_t)a.lda * (j0 + c)] : zmk(0.0, 0.0);
  }
  __syncthreads();
  // 1. the reflectors of the panels to the left
  for (int p = 0; p < b; ++p) {
    if (!qf_wait(a, 2 + p, &sStop)) return;
    if (!qf_apply(a, sA, sPart, sW2r, sW2i, sTp, p, nbp, 1, QF_TFLAG + p, &sStop)) return;
  }
  // 2. factor the panel, one wave per column
  if (m <= 512)
    qf_factor_regs(sA, m, j0, nbp, sTau, sBeta, sVn2);
  else
    qf_factor_lds(sA, m, j0, nbp, sTau, sBeta, sVn2);
  // 3. publish V_b (0 above row j0 + c, 1 on it, the reflector below it)
  // and raise its flag: the next panels' first products V_b^H A overlap the
  // forming of T_b, which starts while the stores drain
  for (int e = tid; e < (m - j0) * nbp; e += QF_THREADS) {
    const int r = j0 + e % (m - j0), c = e / (m - j0);
    const zd v = (r < j0 + c) ? zmk(0.0, 0.0) : (r == j0 + c ? zmk(1.0, 0.0) : sA[r + lm * c]);
    qf_st_z(a.V + r + (size_t)m * (j0 + c), v);
  }
  // T of zlarft from the Gram matrix Z = V^H V (rows [j0, m)), its FP64
  // MFMA products while the stores of V_b drain: Re Z = Vr^T Vr + Vi^T Vi and
  // X = Vr^T Vi, Im Z = X - X^T
  const int npart = a.npart;
  if (wv < npart) {
    const int li = lane & 15, lk = lane >> 4;
    double4_t zr = {0.0, 0.0, 0.0, 0.0}, zx = zr;
    for (int r0 = j0 + 4 * wv; r0 < m; r0 += 4 * npart) {
      const int r = r0 + lk;
      zd v = zmk(0.0, 0.0);
      if (r < m && li < nbp) v = (r < j0 + li) ? zmk(0.0, 0.0) : (r == j0 + li ? zmk(1.0, 0.0) : sA[r + lm * li]);
      zr = mfma_f64(v.x, v.x, zr);
      zx = mfma_f64(v.x, v.y, zx);
      zr = mfma_f64(v.y, v.y, zr);
    }
    double *pr = sPart + (size_t)wv * (QF_PART / 8), *px = pr + 16 * 17;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      pr[(lk + 4 * i) * 17 + li] = zr[i];
      px[(lk + 4 * i) * 17 + li] = zx[i];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's stores are complete
  __syncthreads();
  if (tid == 0) qf_st(&a.sync[2 + b], a.epoch);
  if (tid < 256) {  // Z(l, i) into sW2r / sW2i, all loads first
    const int l = tid >> 4, i = tid & 15;
    double zr[8], zx[8], zt[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      const double *pw = sPart + (size_t)w * (QF_PART / 8);
      zr[w] = w < npart ? pw[l * 17 + i] : 0.0;
      zx[w] = w < npart ? pw[16 * 17 + l * 17 + i] : 0.0;
      zt[w] = w < npart ? pw[16 * 17 + i * 17 + l] : 0.0;
    }
    sW2r[l][i] = ((zr[0] + zr[1]) + (zr[2] + zr[3])) + ((zr[4] + zr[5]) + (zr[6] + zr[7]));
    sW2i[l][i] = (((zx[0] + zx[1]) + (zx[2] + zx[3])) + ((zx[4] + zx[5]) + (zx[6] + zx[7]))) -
                 (((zt[0] + zt[1]) + (zt[2] + zt[3])) + ((zt[4] + zt[5]) + (zt[6] + zt[7])));
  }
  __syncthreads();
  // T^-1 = diag(1/tau) + striu(Z), so T = (I + N)^-1 diag(tau) with
  // N = diag(tau) striu(Z) strictly upper (N^16 = 0) and
  // (I + N)^-1 = (I - N)(I + N^2)(I + N^4)(I + N^8): the squarings (threads
  // 0..255) and the running product W (threads 256..511) side by side, one
  // entry per thread, in four rounds over ping-pong buffers (no 1/tau, so
  // tau = 0 needs no care).
  {
    zd *sPa = reinterpret_cast<zd *>(sPart), *sPb = sPa + 256;  // powers of N
    zd *sWa = sPa + 512, *sWb = sPa + 768;                      // W
    const int i = tid & 15, k = (tid >> 4) & 15, g = tid >> 8;
    auto mm = [&](const zd *A, const zd *B) {  // (A B)(i, k), ld 16
      zd pp = zmk(0.0, 0.0);
#pragma unroll
      for (int l0 = 0; l0 < 16; l0 += 8) {
        zd x[8], y[8];
#pragma unroll
        for (int l = 0; l < 8; ++l) {
          x[l] = A[i + 16 * (l0 + l)];
          y[l] = B[l0 + l + 16 * k];
        }
#pragma unroll
        for (int l = 0; l < 8; ++l) pp = zadd(pp, zmul(x[l], y[l]));
      }
      return pp;
    };
    if (g == 0) {
      const zd ti = i < nbp ? sTau[i] : zmk(0.0, 0.0);
      const zd nik = (i < k && k < nbp) ? zmul(ti, zmk(sW2r[i][k], sW2i[i][k])) : zmk(0.0, 0.0);
      sPa[i + 16 * k] = nik;
      sWa[i + 16 * k] = zsub(zmk(i == k ? 1.0 : 0.0, 0.0), nik);
    }
    __syncthreads();
    if (g == 0) sPb[i + 16 * k] = mm(sPa, sPa);  // N^2
    __syncthreads();
    if (g == 0) sPa[i + 16 * k] = mm(sPb, sPb);  // N^4
    if (g == 1) sWb[i + 16 * k] = zadd(sWa[i + 16 * k], mm(sWa, sPb));
    __syncthreads();
    if (g == 0) sPb[i + 16 * k] = mm(sPa, sPa);  // N^8
    if (g == 1) sWa[i + 16 * k] = zadd(sWb[i + 16 * k], mm(sWb, sPa));
    __syncthreads();
    if (g == 1) sWb[i + 16 * k] = zadd(sWa[i + 16 * k], mm(sWa, sPb));
    __syncthreads();
    const bool own = tid < 256;
    zd *sWk = sWb;
    // T_b, then its flag once every store of it is complete
    if (own) qf_st_z(a.T + 256 * (size_t)b + i + 16 * k, (k < nbp && i < nbp) ? zmul(sWk[i + 16 * k], sTau[k]) : zmk(0.0, 0.0));
  }
  if (tid < nbp) {
    a.tau[j0 + tid] = sTau[tid];
    a.beta[j0 + tid] = sBeta[tid];
    a.vn2[j0 + tid] = sVn2[tid];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's stores are complete
  __syncthreads();
  if (tid == 0) qf_st(&a.sync[QF_TFLAG + b], a.epoch);
  // 4. the panel's columns of A: R above the diagonal, beta on it, v below
  for (int e = tid; e < m * nbp; e += QF_THREADS) {
    const int r = e % m, c = e / m;
    a.A[r + (size_t)a.lda * (j0 + c)] = (r == j0 + c) ? sBeta[c] : sA[r + lm * c];
  }
  if (!a.Q) return;
  __syncthreads();
  // Q(:, j0 + c) = H_0 ... H_b e_{j0 + c}
  for (int e = tid; e < m * nb; e += QF_THREADS) {
    const int r = e % m, c = e / m;
    sA[r + lm * c] = zmk((c < nbp && r == j0 + c) ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int p = b; p >= 0; --p) qf_apply(a, sA, sPart, sW2r, sW2i, sTp, p, nbp, 0, -1, &sStop);
  for (int e = tid; e < m * nbp; e += QF_THREADS) {
    const int r = e % m, c = e / m;
    a.Q[r + (size_t)m * (j0 + c)] = sA[r + lm * c];
  }
}

// Ticket of a launch's workgroup: 0..count-1, or -1 (launch aborted) when
// the ticket counter is out of step with the host.
static __device__ __forceinline__ int qf_ticket(const qf_args &a, int count, int *sIdx, int *sStop) {
  if (threadIdx.x == 0) {
    const unsigned long long t =
        __hip_atomic_fetch_add(&a.sync[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - a.ticket_base;
    *sIdx = t < (unsigned long long)count ? (int)t : -1;
    *sStop = 0;
    if (*sIdx < 0) {
      qf_st(&a.sync[1], a.epoch);
      atomicAdd(a.timeouts, 1ull);
    }
  }
  __syncthreads();
  return *sIdx;
}

__global__ __launch_bounds__(QF_THREADS) void k_qr_flow(qf_args a) {
  extern __shared__ zd qf_lds[];  // zd: 16-byte aligned after the static arrays
  __shared__ double sW2r[16][17], sW2i[16][17];
  __shared__ zd sTau[16], sBeta[16];
  __shared__ double sVn2[16];
  __shared__ zd sTp[256];
  __shared__ int sIdx, sStop;
  const qf_smem sm{sW2r, sW2i, sTau, sBeta, sVn2, sTp, &sStop};
  const int b = qf_ticket(a, a.np, &sIdx, &sStop);
  if (b < 0) return;
  qf_panel(a, b, qf_lds, sm, -1, 0ull);
}

// Passive workgroup c of a fused launch: columns [c nb, c nb + nb) of S get
// H_0^H .. H_{np-1}^H of the first QR; after H_p^H, rows [p nb, p nb + nb) of
// Q_full^H S are final and go, conjugated, into Y's columns [p nb, p nb + nb)
// at rows [c nb, c nb + nb); then count word QF_YCNT + p of the second QR.
static __device__ __forceinline__ void qf_passive(const qf_pair_args &P, int c, zd *qf_lds, const qf_smem &sm) {
  const qf_args &a = P.a1;
  double(*sW2r)[17] = sm.sW2r;
  double(*sW2i)[17] = sm.sW2i;
  const int tid = threadIdx.x;
  const int m = a.m, n = a.n, nb = a.nb, lm = m | 1;
  zd *sA = qf_lds;
  double *sPart = reinterpret_cast<double *>(qf_lds + (size_t)lm * nb);
  const int c0 = c * nb, ncol = min(nb, n - c0);
  for (int e = tid; e < m * nb; e += QF_THREADS) {
    const int r = e % m, cc = e / m;
    sA[r + lm * cc] = cc < ncol ? P.S[r + (size_t)m * (c0 + cc)] : zmk(0.0, 0.0);
  }
  __syncthreads();
  zd *Y = P.a2.A;
  for (int p = 0; p < a.np; ++p) {
    if (!qf_wait(a, 2 + p, sm.sStop)) return;
    if (!qf_apply(a, sA, sPart, sW2r, sW2i, sm.sTp, p, ncol, 1, QF_TFLAG + p, sm.sStop)) return;
    const int r0 = p * nb, npw = min(nb, n - r0);
    for (int e = tid; e < npw * ncol; e += QF_THREADS) {
      const int r = e % npw, cc = e / npw;
      qf_st_z(Y + (c0 + cc) + (size_t)P.a2.lda * (r0 + r), zconj(sA[r0 + r + lm * cc]));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's stores are complete
    __syncthreads();
    if (tid == 0) atomicAdd(&P.a2.sync[QF_YCNT + p], 1ull);
  }
}

__global__ __launch_bounds__(QF_THREADS) void k_qr_pair(qf_pair_args P) {
  extern __shared__ zd qf_lds[];
  __shared__ double sW2r[16][17], sW2i[16][17];
  __shared__ zd sTau[16], sBeta[16];
  __shared__ double sVn2[16];
  __shared__ zd sTp[256];
  __shared__ int sIdx, sStop;
  const qf_smem sm{sW2r, sW2i, sTau, sBeta, sVn2, sTp, &sStop};
  // tickets: the first QR's panels, the passive workgroups, the second QR's
  // panels (each waits only on lower tickets)
  const int np1 = P.a1.np, np2 = P.a2.np;
  const int t = qf_ticket(P.a1, 2 * np1 + np2, &sIdx, &sStop);
  if (t < 0) return;
  if (t < np1)
    qf_panel(P.a1, t, qf_lds, sm, -1, 0ull);
  else if (t < 2 * np1)
    qf_passive(P, t - np1, qf_lds, sm);
  else
    qf_panel(P.a2, t - 2 * np1, qf_lds, sm, QF_YCNT, P.ytarget);
}

extern "C" size_t pqmc_qr_flow_lds(int m, int nb, int *npart) {
  const size_t base = sizeof(zd) * (size_t)(m | 1) * nb;
  const size_t budget = 160 * 1024 - 10 * 1024;  // less the kernel's static LDS
  const int np = (base + 8 * QF_PART <= budget) ? 8 : 4;
  if (npart) *npart = np;
  return base + (size_t)np * QF_PART;
}

// Launch: a->np = ceil(n / nb) workgroups; the caller sets epoch (one more
// than the last launch on this sync array) and ticket_base (the tickets
// issued so far, np per launch).
extern "C" hipError_t pqmc_qr_flow(qf_args *a, hipStream_t s) {
  if (a->m > 1024 || a->n > a->m || a->n <= 0 || (a->nb != 8 && a->nb != 16) || a->np != (a->n + a->nb - 1) / a->nb ||
      a->np > 128)
    return hipErrorInvalidValue;
  int npart = 0;
  const size_t lds = pqmc_qr_flow_lds(a->m, a->nb, &npart);
  if (lds > 160 * 1024 - 10 * 1024) return hipErrorInvalidValue;
  a->npart = npart;
  static std::mutex attr_mu;  // launches may come from several host threads
  static size_t attr = 0;
  {
    std::lock_guard<std::mutex> lk(attr_mu);
    if (lds > attr) {
      hipError_t e = hipFuncSetAttribute((const void *)k_qr_flow, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = lds;
    }
  }
  hipLaunchKernelGGL(k_qr_flow, dim3(a->np), dim3(QF_THREADS), lds, s, *a);
  return hipGetLastError();
}

// Launch of the fused pair: 2 np1 + np2 workgroups; the tickets come from
// a1's counter (ticket_base there), the epochs are each QR's own.
extern "C" hipError_t pqmc_qr_pair(qf_pair_args *P, hipStream_t s) {
  qf_args &a1 = P->a1, &a2 = P->a2;
  if (a1.m > 1024 || a1.n > a1.m || a1.n <= 0 || (a1.nb != 8 && a1.nb != 16) || a1.np != (a1.n + a1.nb - 1) / a1.nb ||
      a1.np > 128 || a2.m != a1.n || a2.n != a1.n || a2.nb != a1.nb || a2.np != a1.np || !P->S || a2.lda != a1.n)
    return hipErrorInvalidValue;
  int npart1 = 0, npart2 = 0;
  const size_t lds1 = pqmc_qr_flow_lds(a1.m, a1.nb, &npart1), lds2 = pqmc_qr_flow_lds(a2.m, a2.nb, &npart2);
  const size_t lds = lds1 > lds2 ? lds1 : lds2;
  if (lds > 160 * 1024 - 10 * 1024) return hipErrorInvalidValue;
  a1.npart = npart1;
  a2.npart = npart2;
  static std::mutex attr_mu;
  static size_t attr = 0;
  {
    std::lock_guard<std::mutex> lk(attr_mu);
    if (lds > attr) {
      hipError_t e = hipFuncSetAttribute((const void *)k_qr_pair, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      attr = lds;
    }
  }
  hipLaunchKernelGGL(k_qr_pair, dim3(2 * a1.np + a2.np), dim3(QF_THREADS), lds, s, *P);
  return hipGetLastError();
}
