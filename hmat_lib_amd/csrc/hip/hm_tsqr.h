// Batched tall-skinny QR with explicit Q (internal interface).
#pragma once

#include "hm_common.h"

namespace hm {

// One panel A (m x ncol, ld lda, factored in place: R and reflectors).
// Outputs (each optional):
//   Q   (m x nq, ld ldq)  first nq columns of the orthogonal factor
//   R   (ncol x ncol, ld ldr) upper triangular factor
//   Mt  (ncol x ncol, ld ncol) R^T
//   est |R[ell:t+1, t]| maximised over columns t >= ell (device scalar)
// Requires m >= ncol.
struct TsqrPanel {
  double *A;
  long long lda;
  int m, ncol, nq;
  double *Q;
  long long ldq;
  double *R;
  long long ldr;
  double *Mt;
  double *est;
  int ell;
};

// Factors all panels (one launch per tree level for all panels at once);
// synchronizes the stream before returning.
int tsqr_run(const std::vector<TsqrPanel> &panels, hipStream_t s);

}  // namespace hm
