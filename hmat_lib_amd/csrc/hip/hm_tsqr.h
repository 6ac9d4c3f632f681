// Batched tall-skinny QR with explicit Q (internal interface).
#pragma once

#include "hm_prog.h"

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

// Appends the factorisation of all panels to a program (one launch per
// reduction level for all panels at once, then the top-down Q build).
int tsqr_add(Program &prog, Arena &ar, const std::vector<TsqrPanel> &panels);

// Builds a program twice (counting, then real), audits the counting pass
// against `inputs` plus the scratch, runs it and synchronizes the stream.
int program_exec(const std::function<int(Program &, Arena &)> &build,
                 const std::vector<MemRange> &inputs, hipStream_t s);

// Householder QR of wide panels: block Gram-Schmidt run twice over panels
// of `bw` columns (BCGS2), each panel orthonormalised by TSQR.  Y (m x L)
// is destroyed; Q (m x L) receives the orthonormal factor and, when R is
// set, R (L x L, upper triangular, ld ldr) the triangular one, Y = Q R.
// Requires L <= m.
struct QrJob {
  double *Y;
  long long ldy;
  int m, L;
  double *Q;
  long long ldq;
  double *R;
  long long ldr;
  // Continuing a QR whose first c_start columns (a multiple of the panel
  // width) are done: their orthonormal columns are copied from Qprev (ld
  // ldqp) into Q and, with R, R's leading c_start x c_start block from the
  // transpose held in Rprev_t (ld ldrp); Y's first c_start columns are not
  // read and the panel sweep starts at column c_start.
  int c_start = 0;
  const double *Qprev = nullptr;
  long long ldqp = 0;
  const double *Rprev_t = nullptr;
  long long ldrp = 0;
};
int qr_add(Program &prog, Arena &ar, const std::vector<QrJob> &jobs, int bw = 64);


// Free the process-wide buffers kept between calls: the launch programs'
// scratch arena, blob and pinned staging (up to 1 GiB + 256 MiB) and the
// compression rounds' work buffer (up to 8 GiB).  INPUT_ERROR when one of
// them is in use by a running call (it is then kept).
int release_exec_buffers();
int release_round_buffer();

}  // namespace hm
