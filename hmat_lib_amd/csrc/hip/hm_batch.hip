// Batched products over many independent HODLR trees of uneven size, rank
// and RHS count (FP64).
//
// The reference multiplies one tree at a time, two BLAS calls per block
// (src/dense_algebra.c:82-429).  A batch here is one launch program, built
// once for fixed operands:
//
//   W       W_b = F_in^T X_src  for every block of every tree (split-K, then
//           a fixed-order sum of the segments where K is long)
//   apply   Y_c = op(D_c) X_c + sum over the levels of F_out(rows of c) W_b
//           for every leaf of every tree: one chained GEMM per leaf whose K
//           concatenates the leaf and the (level, rank) pairs
//
// so a product over the whole batch is 2-3 launches of the persistent
// batched GEMM kernels, whatever the number of trees or levels, and Y is
// written once.  Trees whose W goes through the chunk path of a single
// product (rank sums within its partial tiles) take that path instead,
// all of them in the same launches (tree_prod_add): partials of every level
// per row chunk with X read once, the fixed-order reductions, and apply
// tiles per leaf, one apply launch per RHS width group.  (A launch per
// level was measured first: Y is then read and written once per level,
// 2.25 ms for the 512-tree workload.)  Tile shapes of every launch are
// chosen by timing (program_tune), and the program can be captured once
// into a HIP graph and replayed.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "hm_prog.h"

struct hm_batch {
  std::vector<hm_dtree *> trees;
  std::vector<const MixedFactors *> mixed;  // each tree's sliced factors at creation
  int trans = 0;
  hm::Program prog;
  char *arena = nullptr;
  bool tuned = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t graph_stream = nullptr;
  std::vector<hm::MemRange> inputs;
};

namespace hm {

struct BatchOperands {
  int R;
  const double *X;
  long ldx;
  double *Y;
  long ldy;
};

void subtree_product(Arena &ar, const hm_dtree *H, int l, int q, int trans, const double *X,
                     long long ldx, int R, double *Y, long long ldy, std::vector<GemmJob> &wj,
                     std::vector<GemmChain> &chains, std::vector<GemmLink> &links,
                     bool rowmajor) {
  // element (row r, column j) of X^T / Y^T when rowmajor, of X / Y otherwise
  auto at = [&](auto *base, long long ld, long long r) {
    return rowmajor ? base + r * ld : base + r;
  };
  if (R <= 0) return;
  const int h = H->height, span = 1 << (h - l);
  const int r0 = H->leaf_off[q * span];
  // W_b for the blocks of the subtree (levels l + 1 .. h)
  std::vector<double *> W(H->nblock, nullptr);
  for (int l2 = l + 1; l2 <= h; l2++) {
    for (int j = q << (l2 - l); j < (q + 1) << (l2 - l); j++) {
      const int g = blk_index(l2, j);
      const BlockMeta &b = H->blk[g];
      if (b.s <= 0) continue;
      W[g] = ar.take<double>((size_t)b.s * R);
      const double *F = trans ? H->d_U + b.uoff : H->d_V + b.voff;
      const int K = trans ? b.m : b.n;
      const int src = (trans ? b.r0 : b.c0) - r0;
      // W_b (s x R, ld s) = F^T X[src : src + K, :]
      wj.push_back({F, at(X, ldx, src), W[g], nullptr, std::max(K, 1), ldx, b.s, b.s, R, K, 1,
                    0, 0, 1.0, 0.0});
    }
  }
  // Y_c = op(D_c) X_c + sum over levels of the output factor's rows of
  // leaf c times W_b, one chain per leaf
  for (int c = q * span; c < (q + 1) * span; c++) {
    const int lo = H->leaf_off[c], mc = H->leaf_off[c + 1] - lo;
    if (mc <= 0) continue;
    GemmChain ch;
    ch.C = at(Y, ldy, lo - r0);
    ch.ldc = ldy;
    ch.M = mc;
    ch.N = R;
    ch.tc = rowmajor ? 1 : 0;
    ch.link0 = (int)links.size();
    GemmLink dl{H->d_diag + H->diag_off[c], at(X, ldx, lo - r0), mc, ldx, mc, trans ? 1 : 0};
    dl.tb = rowmajor ? 1 : 0;
    links.push_back(dl);
    for (int l2 = l + 1; l2 <= h; l2++) {
      const int q2 = c >> (h - l2);
      // H X: block (l2, q2) has the leaf's rows (U); H^T X: block
      // (l2, q2^1) has them as columns (V)
      const int g = blk_index(l2, trans ? (q2 ^ 1) : q2);
      const BlockMeta &b = H->blk[g];
      if (b.s <= 0) continue;
      const double *G = trans ? H->d_V + b.voff + (lo - b.c0) : H->d_U + b.uoff + (lo - b.r0);
      const long long ld = trans ? b.n : b.m;
      links.push_back({G, W[g], ld, b.s, b.s, 0});
    }
    ch.nlink = (int)links.size() - ch.link0;
    chains.push_back(ch);
  }
}

// Y_c = op(D_c) X_c for every leaf of H (chains of one link): the FP64
// diagonal part of a tree whose off-diagonal products are sliced.
static void diag_chains(const hm_dtree *H, int trans, const double *X, long long ldx, int R,
                        double *Y, long long ldy, std::vector<GemmChain> &chains,
                        std::vector<GemmLink> &links, bool rowmajor) {
  if (R <= 0) return;
  for (int c = 0; c < H->nleaf; c++) {
    const int lo = H->leaf_off[c], mc = H->leaf_off[c + 1] - lo;
    if (mc <= 0) continue;
    GemmChain ch;
    ch.C = rowmajor ? Y + (long long)lo * ldy : Y + lo;
    ch.ldc = ldy;
    ch.M = mc;
    ch.N = R;
    ch.tc = rowmajor ? 1 : 0;
    ch.link0 = (int)links.size();
    GemmLink dl{H->d_diag + H->diag_off[c], rowmajor ? X + (long long)lo * ldx : X + lo, mc,
                ldx, mc, trans ? 1 : 0};
    dl.tb = rowmajor ? 1 : 0;
    links.push_back(dl);
    ch.nlink = 1;
    chains.push_back(ch);
  }
}

// HM_BATCH_CHAINS=1 sends every FP64 tree through the chained GEMMs (for
// tests and timings that compare the two routes).
static bool chain_route_forced() {
  const char *e = getenv("HM_BATCH_CHAINS");
  return e && e[0] == '1';
}

// trans: 0 Y = H X, 1 Y = H^T X, 2 Y = X H (X, Y: R x n), taken as
// Y^T = H^T X^T.  Trees with sliced factors (hm_hip_tree_set_precision)
// take their FP64 diagonal from the batched apply tiles (with no levels; the
// chained GEMMs when HM_BATCH_CHAINS=1) and their off-diagonal blocks from
// the sliced stages, which run after it, one launch per stage for all of
// them.
static int build_batch(Program &p, Arena &ar, const std::vector<hm_dtree *> &trees, int trans,
                       const std::vector<BatchOperands> &ops) {
  std::vector<GemmJob> wj;
  std::vector<GemmChain> chains;
  std::vector<GemmLink> links;
  std::vector<MixedJob> mj;
  std::vector<TreeProdJob> tj;
  const bool rowmajor = trans == 2;
  const bool chains_only = chain_route_forced();
  for (size_t i = 0; i < trees.size(); i++) {
    const BatchOperands &o = ops[i];
    const bool sliced = trees[i]->slices > 0;
    if (!chains_only && o.R > 0 && tree_prod_batchable(trees[i], trans ? 1 : 0, sliced)) {
      // X H = (H^T X^T)^T: X^T's element (row r, column j) is X[j + r ldx]
      if (rowmajor)
        tj.push_back({trees[i], 1, o.R, o.X, o.ldx, 1, o.Y, o.ldy, 1, sliced});
      else
        tj.push_back({trees[i], trans ? 1 : 0, o.R, o.X, 1, o.ldx, o.Y, 1, o.ldy, sliced});
    } else if (sliced) {
      diag_chains(trees[i], trans ? 1 : 0, o.X, o.ldx, o.R, o.Y, o.ldy, chains, links,
                  rowmajor);
    } else {
      subtree_product(ar, trees[i], 0, 0, trans ? 1 : 0, o.X, o.ldx, o.R, o.Y, o.ldy, wj,
                      chains, links, rowmajor);
    }
    if (sliced && o.R > 0)
      mj.push_back({trees[i], trans ? 1 : 0, o.R, o.X, rowmajor ? o.ldx : 1,
                    rowmajor ? 1 : o.ldx, o.Y, rowmajor ? o.ldy : 1, rowmajor ? 1 : o.ldy});
  }
  HM_OK(gemm_add(p, ar, wj, true, rowmajor, "batch W", -1));
  HM_OK(gemm_chain_add(p, chains, links, "batch apply", -1));
  HM_OK(tree_prod_add(p, ar, tj, rowmajor));
  if (!mj.empty()) HM_OK(mixed_add(p, ar, mj));
  return SUCCESS;
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_hip_batch_create(int ntrees, hm_dtree *const *trees, int trans, const int *R,
                        const double *const *d_X, const long *ldx, double *const *d_Y,
                        const long *ldy, hm_batch **out) {
  HM_OK(ensure_init());
  if (!out || ntrees < 0 || (ntrees > 0 && (!trees || !R || !d_X || !ldx || !d_Y || !ldy))) {
    set_error("hm_hip_batch_create: NULL argument");
    return INPUT_ERROR;
  }
  *out = nullptr;
  std::vector<hm_dtree *> tv(trees, trees + ntrees);
  std::vector<BatchOperands> ops(ntrees);
  std::vector<MemRange> inputs;
  for (int i = 0; i < ntrees; i++) {
    const hm_dtree *t = tv[i];
    // H X, H^T X: X, Y are n x R; X H: R x n
    const long rows = trans == 2 ? R[i] : (t ? t->n : 0);
    const long cols = trans == 2 ? (t ? t->n : 0) : R[i];
    if (!t || R[i] < 0 || (R[i] > 0 && (!d_X[i] || !d_Y[i] || ldx[i] < std::max(1L, rows) ||
                                        ldy[i] < std::max(1L, rows)))) {
      set_error("hm_hip_batch_create: tree %d: bad operands", i);
      return INPUT_ERROR;
    }
    ops[i] = {R[i], d_X[i], ldx[i], d_Y[i], ldy[i]};
    if (R[i] > 0 && cols > 0) {
      inputs.push_back({d_X[i], ((size_t)(cols - 1) * ldx[i] + rows) * sizeof(double)});
      inputs.push_back({d_Y[i], ((size_t)(cols - 1) * ldy[i] + rows) * sizeof(double)});
    }
    long long u = 0, v = 0;
    for (const auto &b : t->blk) {
      u += (long long)b.m * b.s;
      v += (long long)b.n * b.s;
    }
    inputs.push_back({t->d_U, (size_t)u * sizeof(double)});
    inputs.push_back({t->d_V, (size_t)v * sizeof(double)});
    inputs.push_back({t->d_diag, (size_t)t->diag_off[t->nleaf] * sizeof(double)});
    if (t->slices > 0) HM_OK(mixed_ensure_mode(tv[i], trans ? 1 : 0, stream_of(nullptr)));
  }
  // counting pass + audit, then the real program on its own arena
  Program dry;
  Arena count;
  dry.record = true;
  HM_OK(build_batch(dry, count, tv, trans, ops));
  for (const auto &r : inputs) dry.allow(r.lo, r.bytes);
  dry.allow(Arena::fake_base(), count.off + 256);
  std::string msg;
  if (dry.audit(&msg) != SUCCESS) {
    set_error("hm_hip_batch_create: audit: %s", msg.c_str());
    return INPUT_ERROR;
  }
  hm_batch *b = new hm_batch();
  b->trees = tv;
  for (hm_dtree *t : tv) b->mixed.push_back(t->mixed);
  b->trans = trans;
  Arena ar;
  ar.counting = false;
  ar.cap = count.off + 256;
  int rc = dalloc(&ar.base, ar.cap);
  if (rc == SUCCESS) {
    b->arena = ar.base;
    rc = build_batch(b->prog, ar, tv, trans, ops);
  }
  if (rc == SUCCESS) rc = b->prog.upload(stream_of(nullptr));
  if (rc == SUCCESS) {
    hipError_t e = hipStreamSynchronize(stream_of(nullptr));
    if (e != hipSuccess) {
      set_error("hm_hip_batch_create: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  if (rc != SUCCESS) {
    dfree(b->arena);
    delete b;
    return rc;
  }
  *out = b;
  return SUCCESS;
}

int hm_hip_batch_tune(hm_batch *b, void *stream) {
  if (!b) return INPUT_ERROR;
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  HM_OK(program_tune(b->prog, s));
  b->tuned = true;
  // a captured graph holds the old shapes
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  HM_HIP(hipStreamSynchronize(s));
  return SUCCESS;
}

int hm_hip_batch_run(hm_batch *b, int use_graph, void *stream) {
  if (!b) return INPUT_ERROR;
  HM_OK(ensure_init());
  for (size_t i = 0; i < b->trees.size(); i++)
    if (b->trees[i]->mixed != b->mixed[i]) {
      set_error("hm_hip_batch_run: tree %zu changed precision after the batch was created "
                "(create a new batch)", i);
      return INPUT_ERROR;
    }
  hipStream_t s = stream_of(stream);
  if (!use_graph) return b->prog.run(s);
  if (!b->gexec) {
    HM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = b->prog.run(s, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != SUCCESS || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc == SUCCESS) set_error("hm_hip_batch_run: capture: %s", hipGetErrorString(e));
      return rc != SUCCESS ? rc : ALLOCATION_FAILURE;
    }
    b->graph = g;
    HM_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  }
  HM_HIP(hipGraphLaunch(b->gexec, s));
  return SUCCESS;
}

int hm_hip_batch_variants(const hm_batch *b, int *out, int cap) {
  if (!b) return -1;
  const int n = (int)b->prog.tunables.size();
  for (int i = 0; i < n && i < cap; i++) out[i] = *b->prog.tunables[i].choice;
  return n;
}

int hm_hip_batch_destroy(hm_batch *b) {
  if (!b) return SUCCESS;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipStreamSynchronize(stream_of(nullptr));
  dfree(b->arena);
  delete b;
  return SUCCESS;
}

}  // extern "C"
