// Device-resident HODLR trees: metadata, product plan, upload from and
// download to the host pointer tree of the C API.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "hm_common.h"
#include "hmat_lib/hodlr.h"

namespace hm {

void tree_free(hm_dtree *t) {
  if (!t) return;
  mixed_free(t->mixed);
  t->mixed = nullptr;
  if (t->shared) {
    t->shared.reset();  // the pools go with the last tree holding them
  } else {
    dfree(t->d_diag);
    dfree(t->d_U);
    dfree(t->d_V);
    dfree(t->d_meta);  // every metadata and plan array
  }
  dfree(t->d_scratch);
  delete t;
}

// The tree's metadata and plan arrays packed into one device allocation
// and one host -> device copy (each array at a 256-byte boundary): a tree
// has about twenty of them, and one allocation and copy each cost more host
// time than the compression's last round.
struct MetaPack {
  struct Part {
    void **dst;
    const void *src;
    size_t off, bytes;
  };
  std::vector<Part> parts;
  size_t total = 0;
  template <class T>
  void add(T **dst, const std::vector<T> &h) {
    total = (total + 255) & ~(size_t)255;
    parts.push_back({reinterpret_cast<void **>(dst), h.data(), total, h.size() * sizeof(T)});
    total += h.size() * sizeof(T);
  }
  int upload(char **blob, hipStream_t s) {
    dfree(*blob);
    *blob = nullptr;
    HM_OK(dalloc(blob, total + 256));
    std::vector<char> host(total + 256, 0);
    for (const Part &p : parts)
      if (p.bytes) memcpy(host.data() + p.off, p.src, p.bytes);
    HM_HIP(hipMemcpyAsync(*blob, host.data(), host.size(), hipMemcpyHostToDevice, s));
    for (const Part &p : parts) *p.dst = *blob + p.off;
    // the staging copy must finish before host goes out of scope
    HM_HIP(hipStreamSynchronize(s));
    return SUCCESS;
  }
};

SharedPools::~SharedPools() {
  for (void *p : ptrs) dfree(static_cast<char *>(p));
}

// Sizes, offsets and the product plan of a tree; its metadata and plan
// arrays are added to mp (ch: storage for the chunk table, which must live
// until mp is uploaded).  Requires leaf_off (nleaf+1) and
// blk[g].{r0,m,c0,n,s,level,j}.
static void meta_plan(hm_dtree *t, MetaPack &mp, std::vector<int4> &ch) {
  const int h = t->height, nl = t->nleaf;
  t->n = t->leaf_off[nl];
  t->diag_off.assign(nl + 1, 0);
  t->max_leaf = 0;
  for (int c = 0; c < nl; c++) {
    long long m = t->leaf_off[c + 1] - t->leaf_off[c];
    t->diag_off[c + 1] = t->diag_off[c] + m * m;
    t->max_leaf = std::max<int>(t->max_leaf, (int)m);
  }
  long long uo = 0, vo = 0;
  t->max_rank = 0;
  for (auto &b : t->blk) {
    b.uoff = uo;
    b.voff = vo;
    uo += (long long)b.m * b.s;
    vo += (long long)b.n * b.s;
    t->max_rank = std::max(t->max_rank, b.s);
  }
  ProductPlan &p = t->plan;
  p.ch_leaf.clear();
  p.ch_r0.clear();
  p.ch_rows.clear();
  p.leaf_chunk0.assign(nl + 1, 0);
  for (int c = 0; c < nl; c++) {
    p.leaf_chunk0[c] = (int)p.ch_leaf.size();
    const int m = t->leaf_off[c + 1] - t->leaf_off[c];
    for (int r = 0; r < m; r += CHUNK_ROWS) {
      p.ch_leaf.push_back(c);
      p.ch_r0.push_back(r);
      p.ch_rows.push_back(std::min(CHUNK_ROWS, m - r));
    }
  }
  p.leaf_chunk0[nl] = (int)p.ch_leaf.size();
  p.nchunk = (int)p.ch_leaf.size();
  for (int tr = 0; tr < 2; tr++) {
    p.pofs[tr].assign((size_t)h * p.nchunk + 1, 0);
    p.s_in_max[tr] = 0;
    p.s_out_max[tr] = 0;
    long long run = 0;
    for (int l = 1; l <= h; l++) {
      for (int k = 0; k < p.nchunk; k++) {
        const int q = p.ch_leaf[k] >> (h - l);
        const int bin = tr ? q : (q ^ 1);
        p.pofs[tr][(size_t)(l - 1) * p.nchunk + k] = run;
        run += t->blk[blk_index(l, bin)].s;
      }
    }
    p.pofs[tr][(size_t)h * p.nchunk] = run;
    for (int c = 0; c < nl; c++) {
      int sin = 0, sout = 0;
      for (int l = 1; l <= h; l++) {
        const int q = c >> (h - l);
        sin += t->blk[blk_index(l, tr ? q : (q ^ 1))].s;
        sout += t->blk[blk_index(l, tr ? (q ^ 1) : q)].s;
      }
      p.s_in_max[tr] = std::max(p.s_in_max[tr], sin);
      p.s_out_max[tr] = std::max(p.s_out_max[tr], sout);
    }
  }
  p.wofs.assign(t->nblock + 1, 0);
  for (int g = 0; g < t->nblock; g++) p.wofs[g + 1] = p.wofs[g] + t->blk[g].s;
  p.seg_list.clear();
  p.part_ofs.assign(t->nblock + 1, 0);
  std::vector<int> nseg_of(t->nblock, 0);
  long long pr = 0;
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    const int span = 1 << (h - b.level);
    const int q0 = (b.j ^ 1) * span, q1 = b.j * span;  // source nodes (either mode)
    const int nch = std::max(p.leaf_chunk0[q0 + span] - p.leaf_chunk0[q0],
                             p.leaf_chunk0[q1 + span] - p.leaf_chunk0[q1]);
    const int nseg = (nch + REDUCE_SEG - 1) / REDUCE_SEG;
    p.part_ofs[g] = pr;
    if (nseg > 1 && b.s > 0) {
      nseg_of[g] = nseg;
      for (int sgi = 0; sgi < nseg; sgi++) p.seg_list.push_back(make_int2(g, sgi));
      pr += (long long)nseg * b.s;
    }
  }
  p.part_ofs[t->nblock] = pr;
  for (int tr = 0; tr < 2; tr++) {
    // source node of block g: its columns' node (H X) or its rows' (H^T X)
    auto src_chunks = [&](const BlockMeta &b, int &cb, int &ce) {
      const int span = 1 << (h - b.level);
      const int src = tr ? b.j : (b.j ^ 1);
      cb = p.leaf_chunk0[src * span];
      ce = p.leaf_chunk0[(src + 1) * span];
    };
    p.red_seg[tr].clear();
    p.red_w[tr].clear();
    for (const int2 &it : p.seg_list) {
      const BlockMeta &b = t->blk[it.x];
      int cb, ce;
      src_chunks(b, cb, ce);
      const int c0 = cb + it.y * REDUCE_SEG;
      RedItem r;
      r.src = c0 < ce ? p.pofs[tr][(size_t)(b.level - 1) * p.nchunk + c0] : 0;
      r.dst = p.part_ofs[it.x] + (long long)it.y * b.s;
      r.n = std::max(0, std::min(ce, c0 + REDUCE_SEG) - c0);
      r.s = b.s;
      r.from_seg = 0;
      r.pad = 0;
      p.red_seg[tr].push_back(r);
    }
    for (int g = 0; g < t->nblock; g++) {
      const BlockMeta &b = t->blk[g];
      if (b.s <= 0) continue;
      int cb, ce;
      src_chunks(b, cb, ce);
      RedItem r;
      if (nseg_of[g] > 1) {
        r.src = p.part_ofs[g];
        r.n = nseg_of[g];
        r.from_seg = 1;
      } else {
        r.src = ce > cb ? p.pofs[tr][(size_t)(b.level - 1) * p.nchunk + cb] : 0;
        r.n = ce - cb;  // 0 for an empty source node: W_b = 0
        r.from_seg = 0;
      }
      r.dst = p.wofs[g];
      r.s = b.s;
      r.pad = 0;
      p.red_w[tr].push_back(r);
    }
  }
  // H x with the reduction in the apply kernel (small per-leaf partial sets)
  p.fsum.assign(t->nblock, RedItem{0, 0, 0, 0, 0, 0});
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    const int span = 1 << (h - b.level);
    const int cb = p.leaf_chunk0[(b.j ^ 1) * span], ce = p.leaf_chunk0[((b.j ^ 1) + 1) * span];
    RedItem &r = p.fsum[g];
    r.src = ce > cb ? p.pofs[0][(size_t)(b.level - 1) * p.nchunk + cb] : 0;
    r.n = b.s > 0 ? ce - cb : 0;
    r.s = b.s;
  }
  p.mv_fused = true;
  for (int c = 0; c < nl && p.mv_fused; c++) {
    long long vals = 0, ranks = 0;
    for (int l = 1; l <= h; l++) {
      const RedItem &r = p.fsum[blk_index(l, c >> (h - l))];
      vals += (long long)r.n * r.s;
      ranks += r.s;
    }
    p.mv_fused = vals <= MV_FUSE_VALUES && ranks <= MV_FUSE_RANKS;
  }
  // block path plan
  for (int tr = 0; tr < 2; tr++) {
    p.witems[tr].clear();
    p.wmulti[tr].clear();
    p.wmulti_smax[tr] = 0;
    p.wpofs[tr].assign(t->nblock + 1, 0);
    long long run = 0;
    for (int g = 0; g < t->nblock; g++) {
      const BlockMeta &b = t->blk[g];
      const int K = tr ? b.m : b.n;
      const int nseg = (K + WSEG_ROWS - 1) / WSEG_ROWS;
      p.wpofs[tr][g] = run;
      if (b.s <= 0) continue;
      if (K <= 0) {  // empty source node: W_b = 0 (no segments to sum)
        p.wmulti[tr].push_back(g);
        p.wmulti_smax[tr] = std::max(p.wmulti_smax[tr], b.s);
        continue;
      }
      for (int sg = 0; sg < nseg; sg++)
        for (int st = 0; st * WTILE < b.s; st++)
          p.witems[tr].push_back(
              make_int4(g, st, sg, std::min(WSEG_ROWS, K - sg * WSEG_ROWS)));
      if (nseg > 1) {
        p.wmulti[tr].push_back(g);
        p.wmulti_smax[tr] = std::max(p.wmulti_smax[tr], b.s);
        run += (long long)nseg * b.s;
      }
    }
    p.wpofs[tr][t->nblock] = run;
    // longest K first (load balance)
    std::stable_sort(p.witems[tr].begin(), p.witems[tr].end(),
                     [](const int4 &x, const int4 &y) { return x.w > y.w; });
  }
  ch.assign(p.nchunk, make_int4(0, 0, 0, 0));
  for (int k = 0; k < p.nchunk; k++)
    ch[k] = make_int4(p.ch_leaf[k], p.ch_r0[k], p.ch_rows[k],
                      t->leaf_off[p.ch_leaf[k]] + p.ch_r0[k]);
  for (int tr = 0; tr < 2; tr++) {
    const int mtiles = (p.s_in_max[tr] + 15) / 16;
    p.part_mtiles[tr] = mtiles;
    p.prow[tr].assign((size_t)p.nchunk * mtiles * 16, PartRow{0, 0, 0, 1, 0, 0});
    for (int k = 0; k < p.nchunk; k++) {
      const int c = p.ch_leaf[k];
      int rho = 0;
      for (int l = 1; l <= h; l++) {
        const int q = c >> (h - l);
        // input block of the chunk's rows at level l: (l, q^1) for H X (its
        // columns are the chunk's node), (l, q) for H^T X (its rows are)
        const int g = blk_index(l, tr ? q : (q ^ 1));
        const BlockMeta &b = t->blk[g];
        const int off = t->leaf_off[c] - t->leaf_off[q << (h - l)] + p.ch_r0[k];
        for (int kk = 0; kk < b.s; kk++, rho++) {
          PartRow &r = p.prow[tr][(size_t)k * mtiles * 16 + rho];
          r.ld = tr ? b.m : b.n;
          r.f_off = (tr ? b.uoff : b.voff) + off + (long long)kk * r.ld;
          r.p_base = p.pofs[tr][(size_t)(l - 1) * p.nchunk + k];
          r.kk = kk;
          r.s = b.s;
        }
      }
    }
  }

  mp.add(&t->d_leaf_off, t->leaf_off);
  mp.add(&t->d_diag_off, t->diag_off);
  mp.add(&t->d_blk, t->blk);
  mp.add(&p.d_ch, ch);
  for (int tr = 0; tr < 2; tr++) {
    mp.add(&p.d_prow[tr], p.prow[tr]);
    mp.add(&p.d_red_seg[tr], p.red_seg[tr]);
    mp.add(&p.d_red_w[tr], p.red_w[tr]);
  }
  mp.add(&p.d_fsum, p.fsum);
  for (int tr = 0; tr < 2; tr++) mp.add(&p.d_pofs[tr], p.pofs[tr]);
  mp.add(&p.d_wofs, p.wofs);
  for (int tr = 0; tr < 2; tr++) {
    mp.add(&p.d_witems[tr], p.witems[tr]);
    mp.add(&p.d_wpofs[tr], p.wpofs[tr]);
    mp.add(&p.d_wmulti[tr], p.wmulti[tr]);
  }
}

int tree_finalize_meta(hm_dtree *t, hipStream_t s) {
  MetaPack mp;
  std::vector<int4> ch;
  meta_plan(t, mp, ch);
  return mp.upload(&t->d_meta, s);
}

int trees_finalize_meta(const std::vector<hm_dtree *> &ts, char **blob, hipStream_t s) {
  MetaPack mp;
  std::vector<std::vector<int4>> ch(ts.size());
  for (size_t i = 0; i < ts.size(); i++) meta_plan(ts[i], mp, ch[i]);
  *blob = nullptr;
  return mp.upload(blob, s);
}

int tree_alloc_pools(hm_dtree *t) {
  long long u = 0, v = 0;
  for (auto &b : t->blk) {
    u += (long long)b.m * b.s;
    v += (long long)b.n * b.s;
  }
  HM_OK(dalloc(&t->d_diag, (size_t)t->diag_off[t->nleaf]));
  HM_OK(dalloc(&t->d_U, (size_t)u));
  HM_OK(dalloc(&t->d_V, (size_t)v));
  return SUCCESS;
}

long long scratch_need(const hm_dtree *t, int R) {
  const ProductPlan &p = t->plan;
  long long need_p = std::max(std::max(p.pofs[0].back(), p.pofs[1].back()),
                              std::max(p.wpofs[0].back(), p.wpofs[1].back()));
  return (need_p + p.wofs[t->nblock] + p.part_ofs[t->nblock]) * (long long)R + 64;
}

int ensure_scratch(hm_dtree *t, int R) {
  const long long need = scratch_need(t, R);
  if ((size_t)need <= t->scratch_elems) return SUCCESS;
  dfree(t->d_scratch);
  t->d_scratch = nullptr;
  t->scratch_elems = 0;
  HM_OK(dalloc(&t->d_scratch, (size_t)need));
  t->scratch_elems = (size_t)need;
  return SUCCESS;
}

// Internal nodes of a host subtree per depth, left to right.
static std::vector<const HODLRInternalNode *> host_levels(
    const HODLRInternalNode *root, int height) {
  std::vector<const HODLRInternalNode *> all((1u << height) - 1);
  all[0] = root;
  for (int d = 0; d + 1 < height; d++) {
    size_t base = (1u << d) - 1, nbase = (1u << (d + 1)) - 1;
    for (size_t i = 0; i < (1u << d); i++) {
      all[nbase + 2 * i] = all[base + i]->children[0].internal;
      all[nbase + 2 * i + 1] = all[base + i]->children[3].internal;
    }
  }
  return all;
}

}  // namespace hm

using namespace hm;

extern "C" {

int hm_hip_tree_upload(const HODLRInternalNode *node, int height,
                       hm_dtree **out, void *stream) {
  *out = nullptr;
  if (!node || height < 1 || height > 24) {
    set_error("hm_hip_tree_upload: bad node/height (%d)", height);
    return INPUT_ERROR;
  }
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  auto lev = host_levels(node, height);
  hm_dtree *t = new hm_dtree();
  t->height = height;
  t->nleaf = 1 << height;
  t->nblock = (1 << (height + 1)) - 2;
  t->leaf_off.assign(t->nleaf + 1, 0);
  const HODLRInternalNode *const *bottom = &lev[(1u << (height - 1)) - 1];
  std::vector<const double *> leaf_data(t->nleaf);
  for (int q = 0; q < t->nleaf / 2; q++) {
    for (int c = 0; c < 2; c++) {
      const NodeDiagonal &dg = bottom[q]->children[c ? 3 : 0].leaf->data.diagonal;
      if (dg.m < 0 || (dg.m > 0 && !dg.data)) {
        set_error("hm_hip_tree_upload: diagonal leaf %d has no data", 2 * q + c);
        tree_free(t);
        return INPUT_ERROR;
      }
      t->leaf_off[2 * q + c + 1] = dg.m;
      leaf_data[2 * q + c] = dg.data;
    }
  }
  for (int c = 0; c < t->nleaf; c++) t->leaf_off[c + 1] += t->leaf_off[c];
  t->blk.resize(t->nblock);
  std::vector<const double *> hu(t->nblock), hv(t->nblock);
  for (int l = 1; l <= height; l++) {
    const HODLRInternalNode *const *parents = &lev[(1u << (l - 1)) - 1];
    int span = 1 << (height - l);
    for (int j = 0; j < (1 << l); j++) {
      const NodeOffDiagonal &od =
          parents[j >> 1]->children[(j & 1) ? 2 : 1].leaf->data.off_diagonal;
      BlockMeta &b = t->blk[blk_index(l, j)];
      b.r0 = t->leaf_off[j * span];
      b.m = t->leaf_off[(j + 1) * span] - b.r0;
      int jc = j ^ 1;
      b.c0 = t->leaf_off[jc * span];
      b.n = t->leaf_off[(jc + 1) * span] - b.c0;
      b.level = l;
      b.j = j;
      b.pad = 0;
      b.s = od.s < 0 ? 0 : od.s;
      if (od.m != b.m || od.n != b.n || (b.s > 0 && (!od.u || !od.v))) {
        set_error("hm_hip_tree_upload: block level %d index %d inconsistent "
                  "(m=%d/%d n=%d/%d s=%d)", l, j, od.m, b.m, od.n, b.n, od.s);
        tree_free(t);
        return INPUT_ERROR;
      }
      hu[blk_index(l, j)] = od.u;
      hv[blk_index(l, j)] = od.v;
    }
  }
  int rc = tree_finalize_meta(t, s);
  if (rc == SUCCESS) rc = tree_alloc_pools(t);
  if (rc != SUCCESS) {
    tree_free(t);
    return rc;
  }
  // pack on the host, one copy per pool
  std::vector<double> stage((size_t)t->diag_off[t->nleaf] + 1);
  for (int c = 0; c < t->nleaf; c++) {
    long long cnt = t->diag_off[c + 1] - t->diag_off[c];
    if (cnt) memcpy(stage.data() + t->diag_off[c], leaf_data[c], cnt * sizeof(double));
  }
  hipError_t e = hipMemcpy(t->d_diag, stage.data(),
                           t->diag_off[t->nleaf] * sizeof(double),
                           hipMemcpyHostToDevice);
  long long utot = 0, vtot = 0;
  for (auto &b : t->blk) {
    utot += (long long)b.m * b.s;
    vtot += (long long)b.n * b.s;
  }
  std::vector<double> su(utot + 1), sv(vtot + 1);
  for (int g = 0; g < t->nblock; g++) {
    const BlockMeta &b = t->blk[g];
    if (b.s == 0) continue;
    memcpy(su.data() + b.uoff, hu[g], (size_t)b.m * b.s * sizeof(double));
    memcpy(sv.data() + b.voff, hv[g], (size_t)b.n * b.s * sizeof(double));
  }
  if (e == hipSuccess && utot)
    e = hipMemcpy(t->d_U, su.data(), utot * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && vtot)
    e = hipMemcpy(t->d_V, sv.data(), vtot * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("hm_hip_tree_upload: copy failed: %s", hipGetErrorString(e));
    tree_free(t);
    return ALLOCATION_FAILURE;
  }
  *out = t;
  return SUCCESS;
}

int hm_hip_tree_download(const hm_dtree *t, TreeHODLR *host,
                         void *(*alloc)(size_t), void *stream) {
  if (!t || !host || host->height != t->height || !host->root) {
    set_error("hm_hip_tree_download: tree/height mismatch");
    return INPUT_ERROR;
  }
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  if (!alloc) alloc = malloc;
  const int h = t->height;
  auto lev = host_levels(host->root, h);
  // internal node sizes
  for (int d = 0; d < h; d++) {
    HODLRInternalNode *const *nodes =
        const_cast<HODLRInternalNode *const *>(&lev[(1u << d) - 1]);
    int span = 1 << (h - d);
    for (int i = 0; i < (1 << d); i++)
      nodes[i]->m = t->leaf_off[(i + 1) * span] - t->leaf_off[i * span];
  }
  std::vector<double> stage((size_t)t->diag_off[t->nleaf] + 1);
  HM_HIP(hipMemcpyAsync(stage.data(), t->d_diag,
                        t->diag_off[t->nleaf] * sizeof(double),
                        hipMemcpyDeviceToHost, s));
  long long utot = 0, vtot = 0;
  for (auto &b : t->blk) {
    utot += (long long)b.m * b.s;
    vtot += (long long)b.n * b.s;
  }
  std::vector<double> su(utot + 1), sv(vtot + 1);
  if (utot)
    HM_HIP(hipMemcpyAsync(su.data(), t->d_U, utot * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  if (vtot)
    HM_HIP(hipMemcpyAsync(sv.data(), t->d_V, vtot * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  HM_HIP(hipStreamSynchronize(s));
  HODLRInternalNode *const *bottom =
      const_cast<HODLRInternalNode *const *>(&lev[(1u << (h - 1)) - 1]);
  for (int q = 0; q < t->nleaf / 2; q++) {
    for (int c = 0; c < 2; c++) {
      int leaf = 2 * q + c;
      NodeDiagonal &dg = bottom[q]->children[c ? 3 : 0].leaf->data.diagonal;
      int m = t->leaf_off[leaf + 1] - t->leaf_off[leaf];
      dg.m = m;
      size_t bytes = (size_t)m * m * sizeof(double);
      dg.data = static_cast<double *>(alloc(bytes ? bytes : 8));
      if (!dg.data) {
        set_error("hm_hip_tree_download: host allocation failed");
        return ALLOCATION_FAILURE;
      }
      if (bytes) memcpy(dg.data, stage.data() + t->diag_off[leaf], bytes);
    }
  }
  for (int l = 1; l <= h; l++) {
    HODLRInternalNode *const *parents =
        const_cast<HODLRInternalNode *const *>(&lev[(1u << (l - 1)) - 1]);
    for (int j = 0; j < (1 << l); j++) {
      NodeOffDiagonal &od =
          parents[j >> 1]->children[(j & 1) ? 2 : 1].leaf->data.off_diagonal;
      const BlockMeta &b = t->blk[blk_index(l, j)];
      od.m = b.m;
      od.n = b.n;
      od.s = b.s;
      size_t ub = (size_t)b.m * b.s * sizeof(double);
      size_t vb = (size_t)b.n * b.s * sizeof(double);
      od.u = static_cast<double *>(alloc(ub ? ub : 8));
      od.v = static_cast<double *>(alloc(vb ? vb : 8));
      if (!od.u || !od.v) {
        set_error("hm_hip_tree_download: host allocation failed");
        return ALLOCATION_FAILURE;
      }
      if (ub) memcpy(od.u, su.data() + b.uoff, ub);
      if (vb) memcpy(od.v, sv.data() + b.voff, vb);
    }
  }
  return SUCCESS;
}

int hm_hip_tree_destroy(hm_dtree *t) {
  tree_free(t);
  return SUCCESS;
}

int hm_hip_tree_info(const hm_dtree *t, hm_tree_info *info) {
  if (!t || !info) return INPUT_ERROR;
  info->height = t->height;
  info->n = t->n;
  info->nleaf = t->nleaf;
  info->nblock = t->nblock;
  info->max_rank = t->max_rank;
  long long srd = 0;
  for (auto &b : t->blk) srd += (long long)(b.m + b.n) * b.s;
  info->sum_rank_dims = srd;
  info->diag_elems = t->diag_off[t->nleaf];
  return SUCCESS;
}

int hm_hip_tree_ranks(const hm_dtree *t, int *ranks) {
  if (!t || !ranks) return INPUT_ERROR;
  for (int g = 0; g < t->nblock; g++) ranks[g] = t->blk[g].s;
  return SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------- binary files
// The layout of hmat_amd_save_tree (csrc/host/serialize.c): header, leaf
// sizes, ranks in block order, padding to 8 bytes, then the diagonal, U and
// V pools exactly as they lie on the device.
namespace hm {

static const char kMagic[8] = {'H', 'O', 'D', 'L', 'R', 'B', 'I', 'N'};
static constexpr uint32_t kVersion = 1, kEndian = 0x01020304u;
static constexpr int kMaxHeight = 28;

static long long file_header_bytes(int h) {
  const long long b = 32 + 4LL * ((1LL << h) + (1LL << (h + 1)) - 2);
  return (b + 7) & ~7LL;
}

static long long pool_elems(const hm_dtree *t, int which) {
  long long e = 0;
  for (const auto &b : t->blk) e += (long long)(which ? b.n : b.m) * b.s;
  return e;
}

}  // namespace hm

extern "C" int hm_hip_tree_save(const hm_dtree *t, const char *path, void *stream) {
  if (!t || !path) {
    set_error("hm_hip_tree_save: NULL argument");
    return INPUT_ERROR;
  }
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  const int h = t->height;
  std::vector<int32_t> ints;
  for (int c = 0; c < t->nleaf; c++) ints.push_back(t->leaf_off[c + 1] - t->leaf_off[c]);
  for (const auto &b : t->blk) ints.push_back(b.s);
  const long long nd = t->diag_off[t->nleaf], nu = pool_elems(t, 0), nv = pool_elems(t, 1);
  std::vector<double> pool((size_t)std::max(std::max(nd, nu), nv) + 1);
  FILE *f = fopen(path, "wb");
  if (!f) {
    set_error("hm_hip_tree_save: cannot open %s", path);
    return INPUT_ERROR;
  }
  unsigned char head[32];
  memset(head, 0, sizeof(head));
  memcpy(head, kMagic, 8);
  const int32_t hh = h;
  const int64_t n = t->leaf_off[t->nleaf];
  memcpy(head + 8, &kVersion, 4);
  memcpy(head + 12, &kEndian, 4);
  memcpy(head + 16, &hh, 4);
  memcpy(head + 24, &n, 8);
  const char zeros[8] = {0};
  const size_t pad = (size_t)(file_header_bytes(h) - 32 - 4 * (long long)ints.size());
  bool ok = fwrite(head, 1, 32, f) == 32 &&
            fwrite(ints.data(), 4, ints.size(), f) == ints.size() &&
            (pad == 0 || fwrite(zeros, 1, pad, f) == pad);
  const double *src[3] = {t->d_diag, t->d_U, t->d_V};
  const long long cnt[3] = {nd, nu, nv};
  int rc = SUCCESS;
  for (int k = 0; k < 3 && ok; k++) {
    if (cnt[k] == 0) continue;
    hipError_t e = hipMemcpyAsync(pool.data(), src[k], (size_t)cnt[k] * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("hm_hip_tree_save: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
      ok = false;
      break;
    }
    ok = fwrite(pool.data(), 8, (size_t)cnt[k], f) == (size_t)cnt[k];
  }
  if (fclose(f) != 0) ok = false;
  if (rc == SUCCESS && !ok) {
    set_error("hm_hip_tree_save: write to %s failed", path);
    rc = INPUT_ERROR;
  }
  return rc;
}

extern "C" int hm_hip_tree_load(const char *path, hm_dtree **out, void *stream) {
  if (!path || !out) {
    set_error("hm_hip_tree_load: NULL argument");
    return INPUT_ERROR;
  }
  *out = nullptr;
  HM_OK(ensure_init());
  hipStream_t s = stream_of(stream);
  FILE *f = fopen(path, "rb");
  if (!f) {
    set_error("hm_hip_tree_load: cannot open %s", path);
    return INPUT_ERROR;
  }
  auto fail = [&](const char *why) {
    set_error("hm_hip_tree_load: %s: %s", path, why);
    fclose(f);
    return INPUT_ERROR;
  };
  unsigned char head[32];
  if (fread(head, 1, 32, f) != 32 || memcmp(head, kMagic, 8) != 0) return fail("not a HODLR file");
  uint32_t ver, end;
  int32_t h;
  int64_t n;
  memcpy(&ver, head + 8, 4);
  memcpy(&end, head + 12, 4);
  memcpy(&h, head + 16, 4);
  memcpy(&n, head + 24, 8);
  if (ver != kVersion || end != kEndian || h < 1 || h > kMaxHeight || n < 0)
    return fail("unsupported header");
  hm_dtree *t = new hm_dtree();
  t->height = h;
  t->nleaf = 1 << h;
  t->nblock = (1 << (h + 1)) - 2;
  std::vector<int32_t> ints((size_t)t->nleaf + t->nblock);
  if (fread(ints.data(), 4, ints.size(), f) != ints.size()) {
    delete t;
    return fail("truncated header");
  }
  t->leaf_off.assign(t->nleaf + 1, 0);
  bool ok = true;
  for (int c = 0; c < t->nleaf; c++) {
    ok &= ints[c] >= 0;
    t->leaf_off[c + 1] = t->leaf_off[c] + ints[c];
  }
  ok &= t->leaf_off[t->nleaf] == n;
  t->blk.resize(t->nblock);
  for (int l = 1; l <= h && ok; l++) {
    const int span = 1 << (h - l);
    for (int j = 0; j < (1 << l); j++) {
      BlockMeta &b = t->blk[blk_index(l, j)];
      b.r0 = t->leaf_off[j * span];
      b.m = t->leaf_off[(j + 1) * span] - b.r0;
      b.c0 = t->leaf_off[(j ^ 1) * span];
      b.n = t->leaf_off[((j ^ 1) + 1) * span] - b.c0;
      b.level = l;
      b.j = j;
      b.pad = 0;
      b.s = ints[t->nleaf + blk_index(l, j)];
      ok &= b.s >= 0;
    }
  }
  if (!ok) {
    delete t;
    return fail("inconsistent sizes");
  }
  long long nd = 0;
  for (int c = 0; c < t->nleaf; c++) nd += (long long)ints[c] * ints[c];
  const long long nu = pool_elems(t, 0), nv = pool_elems(t, 1);
  long long fsize = -1;
  if (fseek(f, 0, SEEK_END) == 0) fsize = ftell(f);
  if (fsize != file_header_bytes(h) + 8 * (nd + nu + nv) ||
      fseek(f, (long)file_header_bytes(h), SEEK_SET) != 0) {
    delete t;
    return fail("length does not match the header");
  }
  int rc = tree_finalize_meta(t, s);
  if (rc == SUCCESS) rc = tree_alloc_pools(t);
  std::vector<double> pool((size_t)std::max(std::max(nd, nu), nv) + 1);
  double *dst[3] = {t->d_diag, t->d_U, t->d_V};
  const long long cnt[3] = {nd, nu, nv};
  for (int k = 0; k < 3 && rc == SUCCESS; k++) {
    if (cnt[k] == 0) continue;
    if (fread(pool.data(), 8, (size_t)cnt[k], f) != (size_t)cnt[k]) {
      set_error("hm_hip_tree_load: %s: read failed", path);
      rc = INPUT_ERROR;
      break;
    }
    hipError_t e = hipMemcpyAsync(dst[k], pool.data(), (size_t)cnt[k] * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error("hm_hip_tree_load: %s", hipGetErrorString(e));
      rc = ALLOCATION_FAILURE;
    }
  }
  fclose(f);
  if (rc != SUCCESS) {
    tree_free(t);
    return rc;
  }
  *out = t;
  return SUCCESS;
}
