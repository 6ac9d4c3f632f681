// HODLR x HODLR product C = A B on the device (FP64).
//
// The reference (src/hodlr_algebra.c:1452-1553) walks the output tree and
// builds every block from BLAS calls on the two inputs, one block after
// the other: diagonal leaves first (compute_diagonal, 1267-1352), then the
// off-diagonal pairs of each internal node from the innermost level up
// (compute_inner_off_diagonal 780-869, compute_other_off_diagonal
// 1057-1127), each pair recompressed as soon as it is formed.  No block of
// C depends on another block of C, so here the whole product is one launch
// program of four batched launches and one batched recompression:
//
//   T + W   T_b = V_A(b)^T U_B(b~) for every block b = (l, q) of A and its
//           partner b~ = (l, q^1) of B: the s_A x s_B core of
//           A_{q,q^1} B_{q^1,q}, shared by everything below node q; and the
//           W = F^T X pieces of the subtree products below  (gemm, A^T B)
//   M + V   M_{l,c} = T_b V_B(b~)[rows of leaf c]^T for the diagonal leaves,
//           and the ancestor terms V_B(a~)[columns of the block] T_a^T of
//           the off-diagonal blocks                           (gemm, A B^T)
//   chains  C_cc = A_cc B_cc + sum_l U_A(l, q)[rows of c] M_{l,c}, written
//           straight into C's diagonal pool; A_qq U_B and B_pp^T V_A of
//           every off-diagonal block (subtree products)   (chained gemm)
//   copies  U_A rows of the ancestor terms, V_B and U_A of the block itself
//
// then every off-diagonal block (its factor pair of rank s_tot = sum of the
// A ranks above it + s_A + s_B, in the reference's column order) goes
// through the batched recompression (QR method when s_tot < min(m, n),
// dense SVD otherwise, absolute threshold: src/hodlr_algebra.c:835-845) and
// is packed into C's pools.
#include <algorithm>
#include <cstring>

#include "hm_prog.h"
#include "hm_tsqr.h"

namespace hm {

struct HhLayout {
  std::vector<int> stot;          // pre-recompression rank per block
  std::vector<long long> uo, vo;  // offsets of the m x stot / n x stot panels
  long long utot = 0, vtot = 0;
};

static HhLayout hh_layout(const hm_dtree *A, const hm_dtree *B) {
  HhLayout L;
  const int nb = A->nblock;
  L.stot.assign(nb, 0);
  L.uo.assign(nb, 0);
  L.vo.assign(nb, 0);
  for (int g = 0; g < nb; g++) {
    const BlockMeta &a = A->blk[g];
    int s = a.s + B->blk[g].s;
    for (int l2 = a.level - 1; l2 >= 1; l2--) s += A->blk[blk_index(l2, a.j >> (a.level - l2))].s;
    L.stot[g] = s;
    L.uo[g] = L.utot;
    L.vo[g] = L.vtot;
    L.utot += (long long)a.m * s;
    L.vtot += (long long)a.n * s;
  }
  return L;
}

static int build_hh(Program &p, Arena &ar, const hm_dtree *A, const hm_dtree *B,
                    const HhLayout &L, double *Cdiag, const std::vector<long long> &cdiag_off,
                    double *Upre, double *Vpre) {
  const int h = A->height, nb = A->nblock;
  std::vector<GemmJob> tn, nt;
  std::vector<GemmChain> chains;
  std::vector<GemmLink> links;
  std::vector<Copy2D> copies;
  // T_b = V_A(b)^T U_B(b~): (s_A x s_B~), K = the size of node q^1
  std::vector<double *> T(nb, nullptr);
  for (int g = 0; g < nb; g++) {
    const BlockMeta &a = A->blk[g];
    const BlockMeta &bp = B->blk[blk_index(a.level, a.j ^ 1)];
    if (a.s <= 0 || bp.s <= 0) continue;
    T[g] = ar.take<double>((size_t)a.s * bp.s);
    tn.push_back({A->d_V + a.voff, B->d_U + bp.uoff, T[g], nullptr, std::max(a.n, 1),
                  std::max(bp.m, 1), a.s, a.s, bp.s, a.n, 1, 0, 0, 1.0, 0.0});
  }
  // diagonal leaves
  for (int c = 0; c < A->nleaf; c++) {
    const int lo = A->leaf_off[c], mc = A->leaf_off[c + 1] - lo;
    if (mc <= 0) continue;
    GemmChain ch;
    ch.C = Cdiag + cdiag_off[c];
    ch.ldc = mc;
    ch.M = mc;
    ch.N = mc;
    ch.link0 = (int)links.size();
    links.push_back({A->d_diag + A->diag_off[c], B->d_diag + B->diag_off[c], mc, mc, mc, 0});
    for (int l = h; l >= 1; l--) {
      const int q = c >> (h - l);
      const int g = blk_index(l, q);
      const BlockMeta &a = A->blk[g];
      const BlockMeta &bp = B->blk[blk_index(l, q ^ 1)];
      if (!T[g]) continue;
      // M (s_A x m_c) = T V_B(b~)[rows of c]^T; V_B(b~) has the rows of node q
      double *M = ar.take<double>((size_t)a.s * mc);
      nt.push_back({T[g], B->d_V + bp.voff + (lo - bp.c0), M, nullptr, a.s, std::max(bp.n, 1),
                    a.s, a.s, mc, bp.s, 1, 0, 0, 1.0, 0.0});
      links.push_back({A->d_U + a.uoff + (lo - a.r0), M, a.m, a.s, a.s, 0});
    }
    ch.nlink = (int)links.size() - ch.link0;
    chains.push_back(ch);
  }
  // off-diagonal blocks: [ancestor terms, level l-1 .. 1 | A_qq U_B | U_A]
  for (int g = 0; g < nb; g++) {
    const BlockMeta &a = A->blk[g];
    const BlockMeta &bg = B->blk[g];
    const int l = a.level, q = a.j, m = a.m, n = a.n;
    if (L.stot[g] <= 0) continue;
    double *U = Upre + L.uo[g], *V = Vpre + L.vo[g];
    int col = 0;
    for (int l2 = l - 1; l2 >= 1; l2--) {
      const int an = q >> (l - l2);
      const int ga = blk_index(l2, an);
      const BlockMeta &aa = A->blk[ga];
      const BlockMeta &bb = B->blk[blk_index(l2, an ^ 1)];
      if (aa.s <= 0) continue;
      // U: rows of node q inside A's block (l2, an)
      if (m > 0)
        copies.push_back({A->d_U + aa.uoff + (a.r0 - aa.r0), U + (long long)col * m, aa.m, m, m,
                          aa.s, 0, 0});
      // V: rows of node q^1 inside V_B(l2, an^1) (whose rows are node an)
      // times T_a^T
      if (n > 0) {
        if (T[ga])
          nt.push_back({B->d_V + bb.voff + (a.c0 - bb.c0), T[ga], V + (long long)col * n, nullptr,
                        std::max(bb.n, 1), aa.s, n, n, aa.s, bb.s, 1, 0, 0, 1.0, 0.0});
        else  // B's block has rank 0: the term is zero
          copies.push_back({nullptr, V + (long long)col * n, n, n, n, aa.s, 0, 0});
      }
      col += aa.s;
    }
    // A_qq U_B(b), V_B(b)
    if (bg.s > 0) {
      if (m > 0)
        subtree_product(ar, A, l, q, 0, B->d_U + bg.uoff, std::max(bg.m, 1), bg.s,
                        U + (long long)col * m, m, tn, chains, links);
      if (n > 0) copies.push_back({B->d_V + bg.voff, V + (long long)col * n, bg.n, n, n, bg.s, 0, 0});
      col += bg.s;
    }
    // U_A(b), B_pp^T V_A(b) with p = q^1
    if (a.s > 0) {
      if (m > 0) copies.push_back({A->d_U + a.uoff, U + (long long)col * m, a.m, m, m, a.s, 0, 0});
      if (n > 0)
        subtree_product(ar, B, l, q ^ 1, 1, A->d_V + a.voff, std::max(a.n, 1), a.s,
                        V + (long long)col * n, n, tn, chains, links);
      col += a.s;
    }
    if (col != L.stot[g]) {
      set_error("multiply_hodlr_hodlr: block %d: %d columns formed, %d expected", g, col,
                L.stot[g]);
      return INPUT_ERROR;
    }
  }
  HM_OK(gemm_add(p, ar, tn, true, false, "hh T and W", 0));
  HM_OK(gemm_add(p, ar, nt, false, true, "hh M and ancestors", 0));
  HM_OK(gemm_chain_add(p, chains, links, "hh chains", 0));
  HM_OK(copy2d_add(p, copies));
  return SUCCESS;
}

static bool same_shape(const hm_dtree *A, const hm_dtree *B) {
  return A->height == B->height && A->leaf_off == B->leaf_off;
}

static void tree_bytes(const hm_dtree *t, std::vector<MemRange> &r) {
  long long u = 0, v = 0;
  for (const auto &b : t->blk) {
    u += (long long)b.m * b.s;
    v += (long long)b.n * b.s;
  }
  r.push_back({t->d_U, (size_t)u * sizeof(double)});
  r.push_back({t->d_V, (size_t)v * sizeof(double)});
  r.push_back({t->d_diag, (size_t)t->diag_off[t->nleaf] * sizeof(double)});
}

}  // namespace hm

using namespace hm;

extern "C" int hm_hip_multiply_hodlr_hodlr(const hm_dtree *a, const hm_dtree *b, double thr,
                                           hm_dtree **out, void *stream) {
  HM_OK(ensure_init());
  if (!a || !b || !out) {
    set_error("hm_hip_multiply_hodlr_hodlr: NULL argument");
    return INPUT_ERROR;
  }
  *out = nullptr;
  if (!same_shape(a, b)) {
    set_error("hm_hip_multiply_hodlr_hodlr: trees differ in height or leaf sizes");
    return INPUT_ERROR;
  }
  hipStream_t s = stream_of(stream);
  Timed t0 = timing_begin("hh total", s);
  const HhLayout L = hh_layout(a, b);
  hm_dtree *c = new hm_dtree();
  c->height = a->height;
  c->nleaf = a->nleaf;
  c->nblock = a->nblock;
  c->leaf_off = a->leaf_off;
  c->blk = a->blk;
  c->diag_off.assign(c->nleaf + 1, 0);
  for (int i = 0; i < c->nleaf; i++) {
    const long long m = c->leaf_off[i + 1] - c->leaf_off[i];
    c->diag_off[i + 1] = c->diag_off[i] + m * m;
  }
  double *Upre = nullptr, *Vpre = nullptr;
  int rc = dalloc(&c->d_diag, (size_t)c->diag_off[c->nleaf] + 1);
  if (rc == SUCCESS) rc = dalloc(&Upre, (size_t)L.utot + 1);
  if (rc == SUCCESS) rc = dalloc(&Vpre, (size_t)L.vtot + 1);
  if (rc == SUCCESS) {
    std::vector<MemRange> inputs;
    tree_bytes(a, inputs);
    tree_bytes(b, inputs);
    inputs.push_back({c->d_diag, (size_t)c->diag_off[c->nleaf] * sizeof(double)});
    inputs.push_back({Upre, (size_t)L.utot * sizeof(double)});
    inputs.push_back({Vpre, (size_t)L.vtot * sizeof(double)});
    rc = program_exec([&](Program &p, Arena &ar) {
      return build_hh(p, ar, a, b, L, c->d_diag, c->diag_off, Upre, Vpre);
    }, inputs, s);
  }
  // recompression of every off-diagonal block
  std::vector<int> mm(c->nblock), nn(c->nblock), k(c->nblock, 0);
  std::vector<double *> du(c->nblock), dv(c->nblock);
  for (int g = 0; g < c->nblock; g++) {
    mm[g] = c->blk[g].m;
    nn[g] = c->blk[g].n;
    du[g] = Upre + L.uo[g];
    dv[g] = Vpre + L.vo[g];
  }
  if (rc == SUCCESS) {
    // blocks with an empty side keep rank 0 (nothing to recompress)
    std::vector<int> idx;
    for (int g = 0; g < c->nblock; g++)
      if (mm[g] > 0 && nn[g] > 0 && L.stot[g] > 0) idx.push_back(g);
    std::vector<int> m2, n2, s2, k2(idx.size());
    std::vector<double *> u2, v2;
    for (int g : idx) {
      m2.push_back(mm[g]);
      n2.push_back(nn[g]);
      s2.push_back(L.stot[g]);
      u2.push_back(du[g]);
      v2.push_back(dv[g]);
    }
    rc = hm_hip_recompress_batch((int)idx.size(), m2.data(), n2.data(), s2.data(), u2.data(),
                                 v2.data(), thr, -1, k2.data(), stream);
    for (size_t t = 0; rc == SUCCESS && t < idx.size(); t++) k[idx[t]] = k2[t];
  }
  if (rc == SUCCESS) {
    for (int g = 0; g < c->nblock; g++) c->blk[g].s = k[g];
    rc = tree_finalize_meta(c, s);
  }
  if (rc == SUCCESS) {
    long long ut = 0, vt = 0;
    for (const auto &bm : c->blk) {
      ut += (long long)bm.m * bm.s;
      vt += (long long)bm.n * bm.s;
    }
    rc = dalloc(&c->d_U, (size_t)ut + 1);
    if (rc == SUCCESS) rc = dalloc(&c->d_V, (size_t)vt + 1);
    if (rc == SUCCESS) {
      std::vector<Copy2D> pack;
      for (int g = 0; g < c->nblock; g++) {
        const BlockMeta &bm = c->blk[g];
        if (bm.s <= 0) continue;
        pack.push_back({du[g], c->d_U + bm.uoff, bm.m, bm.m, bm.m, bm.s, 0, 0});
        pack.push_back({dv[g], c->d_V + bm.voff, bm.n, bm.n, bm.n, bm.s, 0, 0});
      }
      std::vector<MemRange> inputs = {{Upre, (size_t)L.utot * sizeof(double)},
                                      {Vpre, (size_t)L.vtot * sizeof(double)},
                                      {c->d_U, (size_t)ut * sizeof(double)},
                                      {c->d_V, (size_t)vt * sizeof(double)}};
      rc = program_exec([&](Program &p, Arena &) { return copy2d_add(p, pack); }, inputs, s);
    }
  }
  dfree(Upre);
  dfree(Vpre);
  timing_end(t0, s);
  if (rc != SUCCESS) {
    tree_free(c);
    return rc;
  }
  *out = c;
  return SUCCESS;
}

namespace hm {

// A tree skeleton with fake device addresses (never dereferenced) for the
// host-side audit of the product program.
static hm_dtree *fake_tree(int height, const int *leaves, const int *ranks, char *base,
                           std::vector<MemRange> &bufs) {
  hm_dtree *t = new hm_dtree();
  t->height = height;
  t->nleaf = 1 << height;
  t->nblock = (1 << (height + 1)) - 2;
  t->leaf_off.assign(t->nleaf + 1, 0);
  for (int c = 0; c < t->nleaf; c++) t->leaf_off[c + 1] = t->leaf_off[c] + leaves[c];
  t->diag_off.assign(t->nleaf + 1, 0);
  for (int c = 0; c < t->nleaf; c++)
    t->diag_off[c + 1] = t->diag_off[c] + (long long)leaves[c] * leaves[c];
  t->blk.resize(t->nblock);
  long long uo = 0, vo = 0;
  for (int l = 1; l <= height; l++) {
    const int span = 1 << (height - l);
    for (int j = 0; j < (1 << l); j++) {
      BlockMeta &b = t->blk[blk_index(l, j)];
      b.r0 = t->leaf_off[j * span];
      b.m = t->leaf_off[(j + 1) * span] - b.r0;
      b.c0 = t->leaf_off[(j ^ 1) * span];
      b.n = t->leaf_off[((j ^ 1) + 1) * span] - b.c0;
      b.level = l;
      b.j = j;
      b.s = ranks[blk_index(l, j)];
      b.uoff = uo;
      b.voff = vo;
      uo += (long long)b.m * b.s;
      vo += (long long)b.n * b.s;
    }
  }
  auto take = [&](long long n) {
    double *p = reinterpret_cast<double *>(base);
    bufs.push_back({p, (size_t)n * sizeof(double)});
    return p;
  };
  t->d_U = take(uo);
  t->d_V = reinterpret_cast<double *>(base + ((size_t)1 << 36));
  bufs.push_back({t->d_V, (size_t)vo * sizeof(double)});
  t->d_diag = reinterpret_cast<double *>(base + ((size_t)2 << 36));
  bufs.push_back({t->d_diag, (size_t)t->diag_off[t->nleaf] * sizeof(double)});
  return t;
}

}  // namespace hm

// Builds the product program for trees of the given leaf sizes and block
// ranks (level-major, index 2^l - 2 + j) on fake addresses and checks every
// range it touches against the operands and its scratch, on the host.
extern "C" int hm_hip_audit_hh(int height, const int *leaves, const int *ranks_a,
                               const int *ranks_b, char *msg, int msglen) {
  std::vector<MemRange> bufs;
  char *fa = reinterpret_cast<char *>((uintptr_t)1 << 40);
  char *fb = reinterpret_cast<char *>((uintptr_t)2 << 40);
  char *fc = reinterpret_cast<char *>((uintptr_t)3 << 40);
  hm_dtree *A = fake_tree(height, leaves, ranks_a, fa, bufs);
  hm_dtree *B = fake_tree(height, leaves, ranks_b, fb, bufs);
  const HhLayout L = hh_layout(A, B);
  double *Cdiag = reinterpret_cast<double *>(fc);
  double *Upre = reinterpret_cast<double *>(fc + ((size_t)1 << 36));
  double *Vpre = reinterpret_cast<double *>(fc + ((size_t)2 << 36));
  bufs.push_back({Cdiag, (size_t)A->diag_off[A->nleaf] * sizeof(double)});
  bufs.push_back({Upre, (size_t)L.utot * sizeof(double)});
  bufs.push_back({Vpre, (size_t)L.vtot * sizeof(double)});
  Program p;
  Arena ar;
  p.record = true;
  int rc = build_hh(p, ar, A, B, L, Cdiag, A->diag_off, Upre, Vpre);
  std::string text = rc == SUCCESS ? "" : hm_hip_last_error();
  if (rc == SUCCESS) {
    for (const auto &b : bufs) p.allow(b.lo, b.bytes);
    p.allow(Arena::fake_base(), ar.off + 256);
    rc = p.audit(&text);
    if (rc == SUCCESS) {
      char buf[160];
      snprintf(buf, sizeof(buf), "%zu ops, %zu item bytes, %zu scratch bytes, %lld+%lld factor",
               p.ops.size(), p.blob.size(), ar.off, L.utot, L.vtot);
      text = buf;
    }
  }
  delete A;
  delete B;
  if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", text.c_str());
  return rc;
}
