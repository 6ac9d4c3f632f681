// Launch programs: a host-built sequence of kernel launches whose work-item
// tables live in one device blob, uploaded once and then launched with no
// host synchronisation in between (internal interface).
//
// Builders run twice: first on a counting Arena (device pointers are fake
// offsets, nothing is launched) to size the scratch, then on a real Arena.
// Every item that touches memory records the address ranges it reads and
// writes, so a program can be checked on the host against the buffers it
// was given before anything runs on the device (Program::audit).
#pragma once

#include <cstring>
#include <functional>
#include <memory>
#include <string>

#include "hm_common.h"

namespace hm {

// Scratch carving.  In counting mode take() returns fake, never dereferenced
// pointers (offsets from a fixed fake base) and only the size is tracked.
struct Arena {
  char *base = nullptr;
  size_t off = 0, cap = 0;
  bool counting = true;
  template <class T>
  T *take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    char *p = (counting ? fake_base() : base) + off;
    off += count * sizeof(T) + 16;
    return reinterpret_cast<T *>(p);
  }
  static char *fake_base() { return reinterpret_cast<char *>((uintptr_t)1 << 44); }
};

struct MemRange {
  const void *lo;
  size_t bytes;
};

// An op with several launch shapes (tile variants); the autotuner times
// each on the live operands and keeps the fastest in *choice.
struct Tunable {
  size_t op;                  // index into Program::ops
  int nvariants;
  std::shared_ptr<int> choice;
};

struct Program {
  std::vector<unsigned char> blob;
  std::vector<std::function<void(hipStream_t, const unsigned char *)>> ops;
  std::vector<std::string> names;
  std::vector<Tunable> tunables;
  unsigned char *d_blob = nullptr;
  bool own_blob = true;  // false: d_blob belongs to the caller (not freed here)
  // audit data: address ranges touched by items / buffers they may touch
  bool record = false;
  std::vector<std::pair<MemRange, const char *>> touched;  // labels: string literals
  std::vector<MemRange> allowed;

  ~Program() {
    if (own_blob) dfree(d_blob);
  }
  template <class T>
  size_t push(const std::vector<T> &v) {
    size_t off = (blob.size() + 63) & ~(size_t)63;
    blob.resize(off + v.size() * sizeof(T));
    if (!v.empty()) memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
  void add(const char *name, std::function<void(hipStream_t, const unsigned char *)> f) {
    names.push_back(name);
    ops.push_back(std::move(f));
  }
  // what: a string literal (kept by pointer; recording runs for every item
  // of a program, hundreds of thousands for a batched compression)
  void touch(const void *lo, size_t bytes, const char *what) {
    if (record && bytes) touched.push_back({{lo, bytes}, what});
  }
  void allow(const void *lo, size_t bytes) { allowed.push_back({lo, bytes}); }
  // Every touched range inside one allowed range; message names the first
  // violation.  Returns SUCCESS or INPUT_ERROR.
  int audit(std::string *msg) const;
  int upload(hipStream_t s);
  // launches every op in order (no sync); timed: bracket each op with the
  // kernel-timing events under its name (off while a graph is captured)
  int run(hipStream_t s, bool timed = true);
};

// ---------------------------------------------------------- batched GEMM
// C = alpha op(A) op(B) + beta C, op(A): M x K (A^T when ta), op(B): K x N
// (B^T when tb), column-major.  beta == 0 never reads C.  Long K with few
// output tiles is split into segments whose partial products go to arena
// scratch and are summed in segment order (deterministic).
struct GemmJob {
  const double *A, *B;
  double *C;
  double *part;
  long long lda, ldb, ldc;
  int M, N, K, ksplit;
  int kseg, pad;
  double alpha, beta;
};

// Tile variants (TM x TN): 0 = 64 x 64, 1 = 64 x 16 (narrow RHS),
// 2 = 16 x 64 (few rows, e.g. rank-1 factors).  variant -1 registers the
// op as tunable (starts at 0); GEMM_BY_SHAPE picks from the largest M and N
// of the jobs (N <= 16: 64 x 16, or 16 x 64 when M <= 16 too; M <= 16:
// 16 x 64; else 64 x 64).  The kernel is persistent: at most GEMM_MAX_GRID
// workgroups loop over the tiles.
static constexpr int GEMM_VARIANTS = 3;
static constexpr int GEMM_BY_SHAPE = -2;
static constexpr int GEMM_MAX_GRID = 4096;
int gemm_add(Program &p, Arena &ar, std::vector<GemmJob> jobs, bool ta, bool tb,
             const char *name, int variant = 0);

// Chained GEMM: C (M x N) = sum over its links of op(A_i) op(B_i), the links
// concatenated along K (column-major; op(A) = A^T when ta, op(B) = B^T when
// tb, B then stored N x K; with tc, C is stored transposed, C(i, j) at
// C[j + i * ldc]).  One pass over C; same tile variants as gemm_add.
struct GemmLink {
  const double *A, *B;
  long long lda, ldb;
  int K, ta;
  int tb = 0, pad = 0;
};
struct GemmChain {
  double *C;
  long long ldc;
  int M, N, link0, nlink;
  int tc = 0, pad = 0;
};
int gemm_chain_add(Program &p, const std::vector<GemmChain> &chains,
                   const std::vector<GemmLink> &links, const char *name, int variant = 0);

// Times every variant of every tunable op (3 runs each after a warm-up)
// and keeps the fastest.  Runs ops out of order: outputs are garbage
// until the program is run again from the start.
int program_tune(Program &p, hipStream_t s);

// dst(rows x cols, ldd) = src(rows x cols, lds), or src^T when trans
// (then dst is cols x rows); src == NULL fills dst with zeros.
struct Copy2D {
  const double *src;
  double *dst;
  long long lds, ldd;
  int rows, cols;
  int trans, pad;
};
int copy2d_add(Program &p, const std::vector<Copy2D> &items);

// *out = max over columns of the 2-norm of a (rows x cols, ld) panel
// *out = max over columns j of |a_j| (or of |a_j| / |b_j| when b is set,
// b having b_rows rows and leading dimension ldb)
struct ColNorm {
  const double *a;
  long long ld;
  int rows, cols;
  double *out;
  const double *b;
  long long ldb;
  int b_rows, pad;
};
int colnorm_add(Program &p, const std::vector<ColNorm> &items);

// ------------------------------------------------ tree products as items
// Y = op(H_qq) X for the diagonal subtree of H at node (l, q) (l = 0: the
// whole matrix; l = height: one diagonal leaf).  X and Y (R columns,
// leading dims ldx / ldy) start at the node's first row.  Appends one
// W_b = F_in^T X[rows of b's input node] job (for gemm_add with ta = 1,
// W in `ar` scratch) per off-diagonal block of the subtree, and one chained
// GEMM per leaf c: Y_c = op(D_c) X_c + sum over the levels below l of
// F_out(rows of c) W_b, the diagonal leaf and the blocks concatenated
// along K.  F_in / F_out are V / U for H X and U / V for H^T X.
// rowmajor: X and Y are given as R x n column-major (leading dims ldx,
// ldy, starting at the node's first column) and the product is taken on
// their transposes, Y^T = op(H_qq) X^T — the W jobs then need gemm_add
// with tb = true.
void subtree_product(Arena &ar, const hm_dtree *H, int l, int q, int trans, const double *X,
                     long long ldx, int R, double *Y, long long ldy, std::vector<GemmJob> &wj,
                     std::vector<GemmChain> &chains, std::vector<GemmLink> &links,
                     bool rowmajor = false);

// ------------------------------------------ sliced (mixed-precision) items
// One product op(H) X (op = transpose when trans) of a tree whose factors
// are sliced (hm_hip_tree_set_precision > 0, mode sliced with
// mixed_ensure_mode), X / Y element (row, col) at [row * rs + col * cs]:
// only the off-diagonal part, Y += sum of the blocks; the FP64 diagonal is
// added separately (and must run first).
struct MixedJob {
  hm_dtree *t;
  int trans, R;
  const double *X;
  long x_rs, x_cs;
  double *Y;
  long y_rs, y_cs;
};
// Appends the sliced stages of all jobs (one launch per stage and slice
// count, whatever the number of trees).
int mixed_add(Program &p, Arena &ar, const std::vector<MixedJob> &jobs);

// ------------------------------------ batched tree products (FP64 trees)
// One product op(H) X (op = transpose when trans), X / Y element (row, col)
// at [row * rs + col * cs], run by the tree-product kernels of a single
// product (chunk partials with X read once, fixed-order reductions, apply
// tiles over the leaves) with every tree of the batch in the same launches.
struct TreeProdJob {
  hm_dtree *t;
  int trans, R;
  const double *X;
  long x_rs, x_cs;
  double *Y;
  long y_rs, y_cs;
  bool diag_only;  // Y = op(D) X only (the FP64 diagonal of a sliced tree)
};
// diag_only: any tree with leaves; else FP64 factors and the chunk W path
// (rank sum per chunk within the partial kernel's tiles) in mode trans.
bool tree_prod_batchable(const hm_dtree *t, int trans, bool diag_only = false);
// Appends partials, the two reduction stages and one apply launch per RHS
// width group (<= 16, <= 32, wider), each apply tunable over its tile shape;
// product scratch of every tree is taken from `ar`.  nfast: X is row-major
// (X H as H^T X^T), the same for every job.
int tree_prod_add(Program &p, Arena &ar, const std::vector<TreeProdJob> &jobs, bool nfast);

}  // namespace hm
