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

#define QF_SYNC_WORDS 320  // 2 + two flags for each of at most 128 panels, rounded up
#define QF_TFLAG 130       // offset of the T flags

extern "C" {
hipError_t pqmc_qr_flow(qf_args *a, hipStream_t s);
size_t pqmc_qr_flow_lds(int m, int nb, int *npart);
}
