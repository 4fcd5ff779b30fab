// Shared declarations of the fused family-pass kernel (k_pass.hip) and its
// host launcher.
#pragma once
#include <cstddef>

#include "pqmc_common.h"

#define PS_THREADS 512
#define PS_NB 8             // bonds per decision block (maximum)
#define PS_BS (4 * PS_NB)   // sites per block
#define PS_WAIT_TICKS 50000000ull  // 0.5 s of the 100 MHz wall clock

// What a helper needs of a decided block: the a accepted bonds and, for the
// rank-4 updates G <- G + u_m v_m of accepts m = 0..a-1 (u_m = G[:, S_m] X_m,
// v_m = (I - G)[S_m, :], G current before accept m), the 4 x 4 blocks
//     X_m                           on the diagonal block (m, m),
//     T[m][m'] = u_m'[S_m]          below it (m' < m),
//     X_m' V[m'][m] = X_m' v_m'[S_m] above it (m' < m),
// row-major with stride PS_BS.  With them a helper forms the folded update
// of its own columns q from the pivot columns at the block's start:
//     v_m(q) = w_m(q) - sum_{m'<m} T[m][m'] v_m'(q),    w_m(q) = (I - G)[S_m, q],
//     y_m = X_m v_m(q) + sum_{m'>m} (X_m V[m][m']) y_m'   (m = a-1 .. 0),
//     G_after[:, q] = G[:, q] + sum_m G[:, S_m] y_m.
// The accept count itself travels in the block's flag word.
// Helpers read a record with plain, L2-cached loads after its flag, so no
// 128-byte line of one record may also hold part of another: a line cached
// while reading record k would otherwise serve stale indices of record k+1.
// The header fills one whole line and the record is a whole number of lines
// (the array itself comes from hipMalloc, 256-byte aligned).
struct __align__(128) ps_rec {
  int li[PS_NB];          // block-local indices of the accepted bonds, ascending
  int pad[32 - PS_NB];
  zd AT[PS_BS * PS_BS];   // X on the diagonal blocks, V above, T below; row-major, stride PS_BS
};
static_assert(sizeof(ps_rec) % 128 == 0, "a decision record must be a whole number of 128-byte lines");
static_assert(offsetof(ps_rec, AT) == 128, "the record header must fill exactly one 128-byte line");

// Proposal of one bond for the leader's decisions (ps_propose).
struct ps_prop {
  zd D[16];  // Delta = E(old)^-1 E(new) - I, column-major
  ekv_t en;
  double thr, xnew;
  int fld, pad[3];
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
  zd *slot;           // [nblk + 1][2][PS_BS][lds] columns of block k+1 after step k = -1..nblk-1 (the
                      // second half of each step's slot is not used)
  int lds;            // column stride of the slot: Ns rounded up to whole 128-byte lines
  zd *tslot;          // [nblk + 1][2][PS_BS][2 PS_BS] the same columns at the rows of blocks k+1, k+2
  ps_rec *rec;        // [nblk]
  unsigned long long *sync;  // [1] done, [2] abort, [8 + 2(k + 1) + mb] owners of block k+1+mb that have
                             // published their tile-slot columns after step k (k = -1..nblk-1),
                             // [10 + 2 nblk + k] block k decided: accepts + 1, [10 + 3 nblk + k + 1]
                             // owners of block k+1 that have published their slot columns after step k
  // block 0's proposals formed by the previous pass (or null: formed here);
  // this pass's wave 7 forms those of the next pass (family nx_f, slice nx_t,
  // its draws and tables) into nxt, unless null.  The next pass reads no
  // field or exponential that this one changes (another family, or another
  // slice with another family).
  const ps_prop *pre;
  ps_prop *nxt;
  int nx_f, nx_t;
  const float *nx_u;
  const int *nx_fbfi, *nx_fcross;
  const zd *nx_fph;
  uint8_t *acc;       // [npla]
  double *lncw_acc;
  unsigned long long *stats;  // [0] accepted, [1] rejected, [2] wait timeouts
  unsigned long long *prof;   // optional [gridDim][8] wall-clock ticks
};

// column stride and size of the slot: whole 128-byte lines (8 complex
// entries), like the tile-slot columns of 2 PS_BS entries, so that every
// column read with plain loads starts and ends on a line boundary
static_assert((2 * PS_BS * sizeof(zd)) % 128 == 0, "tile-slot columns must be whole 128-byte lines");
static inline int pqmc_pass_slot_ld(int Ns) { return (Ns + 7) & ~7; }
static inline size_t pqmc_pass_slot_elems(int Ns, int nblk) {
  return (size_t)(nblk + 1) * 2 * PS_BS * pqmc_pass_slot_ld(Ns);
}
static inline size_t pqmc_pass_tslot_elems(int nblk) { return (size_t)(nblk + 1) * 2 * PS_BS * 2 * PS_BS; }
// words of the sync array for nblk decision blocks
static inline size_t pqmc_pass_sync_words(int nblk) { return 16 + 4 * (size_t)nblk; }

extern "C" {
size_t pqmc_pass_lds(int Ns, int bph);
size_t pqmc_pass_rec_bytes(void);
hipError_t pqmc_launch_pass(const ps_args *a, hipStream_t s);
hipError_t pqmc_pass_max_groups(int Ns, int bph, int *out);
}
