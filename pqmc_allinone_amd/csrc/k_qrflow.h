// Arguments of the one-launch blocked Householder QR (k_qrflow.hip).
#pragma once
#include "pqmc_common.h"

struct qf_args {
  zd *A;
  int lda, m, n, nb, np, npart;
  zd *Q;       // m x n, ld m, or null
  zd *V;       // m x n, ld m
  zd *T;       // 256 per panel (16 x 16, ld 16)
  zd *tau, *beta;
  double *vn2;
  unsigned long long *sync;      // [0] tickets, [1] aborted epoch, [2 + p] panel p's V flag,
                                 // [2 + 128 + p] its T flag (epochs)
  unsigned long long *timeouts;  // counts expired waits
  unsigned long long epoch, ticket_base;
};

#define QF_SYNC_WORDS 400  // 2 + two flags for each of at most 128 panels, the column counts, rounded up
#define QF_TFLAG 130       // offset of the T flags
#define QF_YCNT 258        // offset of the counts of the fused launch's second QR (at most 128 panels)

// The stabilization's two QRs in one launch (k_qr_pair): the first QR (a1,
// m x n) and, beside its panels, np1 "passive" workgroups that apply its
// block reflectors to the columns of S (m x n, ld m) and hand the first n
// rows of Q_full^H S over as the columns of Y = (Q_full^H S)[0:n, :]^H (n x n,
// ld n, a2.A) as soon as each row block is final; the second QR (a2) factors
// Y panel by panel as its columns arrive.
struct qf_pair_args {
  qf_args a1, a2;
  const zd *S;
  unsigned long long ytarget;  // the second QR's panel p starts once count word QF_YCNT + p reaches it
};

extern "C" {
hipError_t pqmc_qr_flow(qf_args *a, hipStream_t s);
hipError_t pqmc_qr_pair(qf_pair_args *p, hipStream_t s);
size_t pqmc_qr_flow_lds(int m, int nb, int *npart);
}
