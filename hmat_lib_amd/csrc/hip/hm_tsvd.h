// Batched truncated SVD of dense blocks (internal interface).
#pragma once

#include <functional>

#include "hm_common.h"

namespace hm {

enum TsvdRule { RULE_RELATIVE = 0, RULE_ABSOLUTE = 1 };

// One dense block to factor as U V^T (U carries sigma).
struct TsvdBlock {
  const double *A;
  long long lda;
  int m, n;
  int ell;       // sketch width
  int ntest;     // posterior test vectors (0 = exact width)
  int rule;
  int status;    // out: 1 accepted
  int k;         // out: rank
  int need;      // out: singular values of the core >= accept/20 * thr (* sigma_0)
  double thr;
  double err;    // out: residual estimate (Frobenius-type bound)
  double perr;   // residual estimate after one power step (spectral; +inf until computed)
  double terr;   // max_t |(I - Q Q^T) A w_t| over the test columns (input of the decision)
  double sigma0; // out: sigma_0 lower bound, then exact top singular value
  long long yoff, ooff, noff, qoff, boff, moff, soff;
  unsigned long long uid;
  double *uout;  // out: U (m x k, ld m) and V (n x k, ld n)
  double *vout;
};

// Sorted singular values of the first round's core kept per block.
static constexpr int FIRST_SIGMA = 16;

struct TsvdOptions {
  int sketch_max = 512;
  int n_test = 6;
  double accept = 0.1;
  unsigned long long seed = 0x9e3779b97f4a7c15ull;
  // After the first round: sig (nb x FIRST_SIGMA, zero padded) holds the
  // first round's singular values of every block.  `widen`, when set,
  // chooses the next sketch width of the blocks in `retry` (instead of
  // doubling); later rounds double.
  std::function<int(std::vector<TsvdBlock> &, const std::vector<int> &retry,
                    const std::vector<double> &sig)>
      widen;
  std::vector<double> *first_sigma = nullptr;
};

struct TsvdStats {
  int rounds = 0;
  int retried = 0;
  int max_sketch = 0;
  int exact_blocks = 0;
  int capped = 0;      // blocks that hit sketch_max without passing the test
  int widened = 0;     // blocks whose second width came from `widen`
  int widened_failed = 0;  // ... and failed the residual test at that width
  int continued = 0;   // block rounds that continued the previous round's range finder
};

// Factor every block.  On entry blocks[i].{A,lda,m,n,ell,rule,thr,uid} are
// set.  On return .k is set and U/V of block i live at .uout/.vout inside
// pools allocated here and returned in *u_pool / *v_pool (caller frees).
// sigma (optional, host) receives the sorted singular values of the final
// core of each block, up to `sigma_cap` per block.
int tsvd_run(std::vector<TsvdBlock> &blocks, const TsvdOptions &opt,
             double **u_pool, double **v_pool, TsvdStats *stats,
             std::vector<double> *sigma, int sigma_cap, hipStream_t s);

}  // namespace hm
