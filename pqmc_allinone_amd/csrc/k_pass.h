// Shared declarations of the fused family-pass kernel (k_pass.hip) and its
// host launcher.
#pragma once
#include "pqmc_common.h"

#define PS_THREADS 512
#define PS_NB 8             // bonds per decision block (maximum)
#define PS_BS (4 * PS_NB)   // sites per block
#define PS_WAIT_TICKS 50000000ull  // 0.5 s of the 100 MHz wall clock

struct ps_rec {
  int a;            // accepted bonds in the block
  int li[PS_NB];    // their block-local indices, ascending
  int pad[7];
  zd X[PS_NB][16];  // X_m, column-major
  // for accepted m and a later block-local bond j:
  //   ub[m][j][r*4 + c] = u_m[S_j(r)][c],  vb[m][j][c*4 + s] = v_m[c][S_j(s)]
  zd ub[PS_NB][PS_NB][16];
  zd vb[PS_NB][PS_NB][16];
};

struct ps_args {
  model_t md;
  int f, t, wrap;    // family (0-based), slice (1-based); wrap 0 none, 1 E G E^-1 after, 2 E^-1 G E before
  int nb, nblk, bph, H;
  const int *fsites;  // [npla*4] this family's sites in family order
  const int *fbfi, *fcross;
  const zd *fph;
  double *x;          // boson_field(Nb, Lt)
  ekv_t *ekv;         // entry ((t-1)*4 + f)*npla + b
  const float *u;     // 2*npla draws: (ran_sysm source, rands) per bond
  zd *G;              // Ns x Ns, column-major, in/out
  zd *slot;           // [2][2][PS_BS][Ns] columns of two blocks after a step
  ps_rec *rec;        // [nblk]
  unsigned long long *sync;  // [0] blocks published, [1] done, [2] abort, [8 + k + 1] arrivals after step k
  uint8_t *acc;       // [npla]
  double *lncw_acc;
  unsigned long long *stats;  // [0] accepted, [1] rejected, [2] wait timeouts
  unsigned long long *prof;   // optional [gridDim][8] wall-clock ticks
};

extern "C" {
size_t pqmc_pass_lds(int Ns, int bph);
size_t pqmc_pass_rec_bytes(void);
hipError_t pqmc_launch_pass(const ps_args *a, hipStream_t s);
hipError_t pqmc_pass_max_groups(int Ns, int bph, int *out);
}
