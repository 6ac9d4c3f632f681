/* Binary HODLR files and the reference's text dense-matrix format.
 *
 * Not in the reference, which reads only whitespace text dense matrices in
 * its tests (tests/utils/io.c:15-81).  One binary layout serves host trees
 * (here) and device trees (hm_hip_tree_save / hm_hip_tree_load), so a file
 * written by either side loads on both; its payload is the device tree's
 * three packed pools in order, so the device side moves it with one copy
 * per pool.
 *
 *   0   char[8]  "HODLRBIN"
 *   8   uint32   version (1)
 *   12  uint32   0x01020304 (byte order of the writer)
 *   16  int32    height h
 *   20  int32    0
 *   24  int64    n (sum of the leaf sizes)
 *   32  int32    leaf sizes [2^h], left to right
 *       int32    ranks [2^(h+1) - 2], block (level l, index j) at 2^l - 2 + j
 *                (j even: top-right child of parent j/2, odd: bottom-left)
 *       zero padding to a multiple of 8 bytes
 *       float64  diagonal leaves, leaf order, each m x m column-major
 *       float64  U of every block in the order above, m x s column-major
 *       float64  V of every block likewise, n x s
 *
 * Block (l, j) spans the rows of node (l, j) and the columns of node
 * (l, j^1), so its sizes follow from the leaf sizes.  A file whose length
 * differs from what its header implies is rejected (INPUT_ERROR). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hm_host.h"
#include "hmat_lib/allocators.h"

#define HM_MAGIC "HODLRBIN"
#define HM_VERSION 1u
#define HM_ENDIAN 0x01020304u
#define HM_MAX_HEIGHT 28

/* node (level, q) of a tree with these leaf sizes: first row and size */
static void node_range(const long *off, int h, int level, long q, long *r0, long *m) {
  const long span = 1L << (h - level);
  *r0 = off[q * span];
  *m = off[(q + 1) * span] - *r0;
}

static long header_bytes(int h) {
  const long nleaf = 1L << h, nblock = (1L << (h + 1)) - 2;
  const long b = 32 + 4 * (nleaf + nblock);
  return (b + 7) & ~7L;
}

static struct NodeOffDiagonal *block_of(struct HODLRInternalNode **lev, int level, long j) {
  struct HODLRInternalNode *p = lev[(1L << (level - 1)) - 1 + (j >> 1)];
  return &p->children[(j & 1) ? 2 : 1].leaf->data.off_diagonal;
}

static int write_all(FILE *f, const void *p, size_t bytes) {
  return bytes == 0 || fwrite(p, 1, bytes, f) == bytes;
}

HM_EXPORT int hmat_amd_save_tree(const struct TreeHODLR *tree, const char *path) {
  if (!tree || !tree->root || !path || tree->height < 1 || tree->height > HM_MAX_HEIGHT)
    return INPUT_ERROR;
  const int h = tree->height;
  const long nleaf = 1L << h, nblock = (1L << (h + 1)) - 2;
  struct HODLRInternalNode **lev = hm_levels(tree->root, h);
  int32_t *ints = malloc((size_t)(nleaf + nblock) * sizeof(int32_t));
  if (!lev || !ints) {
    free(lev);
    free(ints);
    return ALLOCATION_FAILURE;
  }
  long n = 0;
  for (long c = 0; c < nleaf; c++) {
    ints[c] = tree->innermost_leaves[c]->data.diagonal.m;
    n += ints[c];
  }
  int rc = SUCCESS;
  for (int l = 1; l <= h && rc == SUCCESS; l++)
    for (long j = 0; j < (1L << l); j++) {
      const struct NodeOffDiagonal *od = block_of(lev, l, j);
      if (od->s < 0 || (od->s > 0 && (!od->u || !od->v))) rc = INPUT_ERROR;
      ints[nleaf + (1L << l) - 2 + j] = od->s;
    }
  FILE *f = rc == SUCCESS ? fopen(path, "wb") : NULL;
  if (rc == SUCCESS && !f) rc = INPUT_ERROR;
  if (rc == SUCCESS) {
    unsigned char head[32];
    memset(head, 0, sizeof(head));
    memcpy(head, HM_MAGIC, 8);
    const uint32_t ver = HM_VERSION, end = HM_ENDIAN;
    const int32_t hh = h;
    const int64_t nn = n;
    memcpy(head + 8, &ver, 4);
    memcpy(head + 12, &end, 4);
    memcpy(head + 16, &hh, 4);
    memcpy(head + 24, &nn, 8);
    const long pad = header_bytes(h) - 32 - 4 * (nleaf + nblock);
    const char zeros[8] = {0};
    int ok = write_all(f, head, 32) &&
             write_all(f, ints, (size_t)(nleaf + nblock) * sizeof(int32_t)) &&
             write_all(f, zeros, (size_t)pad);
    for (long c = 0; ok && c < nleaf; c++) {
      const struct NodeDiagonal *d = &tree->innermost_leaves[c]->data.diagonal;
      const size_t cnt = (size_t)d->m * (size_t)d->m;
      ok = (cnt == 0 || d->data) && write_all(f, d->data, cnt * sizeof(double));
    }
    for (int side = 0; side < 2 && ok; side++)
      for (int l = 1; l <= h && ok; l++)
        for (long j = 0; j < (1L << l) && ok; j++) {
          const struct NodeOffDiagonal *od = block_of(lev, l, j);
          const size_t cnt = (size_t)(side ? od->n : od->m) * (size_t)od->s;
          ok = write_all(f, side ? od->v : od->u, cnt * sizeof(double));
        }
    if (fclose(f) != 0) ok = 0;
    if (!ok) rc = INPUT_ERROR;
  }
  free(ints);
  free(lev);
  return rc;
}

static void drop_tree(struct TreeHODLR **t) {
#ifdef _TEST_HODLR
  free_tree_hodlr(t, free);
#else
  free_tree_hodlr(t);
#endif
}

static int read_all(FILE *f, void *p, size_t bytes) {
  return bytes == 0 || fread(p, 1, bytes, f) == bytes;
}

HM_EXPORT struct TreeHODLR *hmat_amd_load_tree(const char *path, int *ierr) {
  *ierr = INPUT_ERROR;
  FILE *f = path ? fopen(path, "rb") : NULL;
  if (!f) return NULL;
  unsigned char head[32];
  uint32_t ver = 0, end = 0;
  int32_t h = 0;
  int64_t n = 0;
  if (!read_all(f, head, 32) || memcmp(head, HM_MAGIC, 8) != 0) {
    fclose(f);
    return NULL;
  }
  memcpy(&ver, head + 8, 4);
  memcpy(&end, head + 12, 4);
  memcpy(&h, head + 16, 4);
  memcpy(&n, head + 24, 8);
  if (ver != HM_VERSION || end != HM_ENDIAN || h < 1 || h > HM_MAX_HEIGHT || n < 0) {
    fclose(f);
    return NULL;
  }
  const long nleaf = 1L << h, nblock = (1L << (h + 1)) - 2;
  int32_t *ints = malloc((size_t)(nleaf + nblock) * sizeof(int32_t));
  long *off = malloc((size_t)(nleaf + 1) * sizeof(long));
  if (!ints || !off || !read_all(f, ints, (size_t)(nleaf + nblock) * sizeof(int32_t))) {
    free(ints);
    free(off);
    fclose(f);
    if (!ints || !off) *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  /* sizes: leaves sum to n, ranks non-negative, payload length exact */
  int ok = 1;
  off[0] = 0;
  for (long c = 0; c < nleaf; c++) {
    ok &= ints[c] >= 0;
    off[c + 1] = off[c] + ints[c];
  }
  ok &= off[nleaf] == n;
  long long payload = 0;
  for (long c = 0; ok && c < nleaf; c++) payload += (long long)ints[c] * ints[c];
  for (int l = 1; ok && l <= h; l++)
    for (long j = 0; j < (1L << l); j++) {
      const int32_t s = ints[nleaf + (1L << l) - 2 + j];
      long r0, m, c0, nn;
      node_range(off, h, l, j, &r0, &m);
      node_range(off, h, l, j ^ 1, &c0, &nn);
      ok &= s >= 0;
      payload += (long long)(m + nn) * s;
    }
  long long fsize = -1;
  if (ok && fseek(f, 0, SEEK_END) == 0) fsize = ftell(f);
  ok &= fsize == header_bytes(h) + 8 * payload;
  ok &= fseek(f, header_bytes(h), SEEK_SET) == 0;
  if (!ok) {
    free(ints);
    free(off);
    fclose(f);
    return NULL;
  }
  int rc = SUCCESS;
  struct TreeHODLR *t = allocate_tree(h, &rc);
  struct HODLRInternalNode **lev = t ? hm_levels(t->root, h) : NULL;
  if (!t || !lev) {
    if (t) drop_tree(&t);
    free(ints);
    free(off);
    fclose(f);
    *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  for (int d = 0; d < h; d++)
    for (long q = 0; q < (1L << d); q++) {
      long r0, m;
      node_range(off, h, d, q, &r0, &m);
      lev[(1L << d) - 1 + q]->m = (int)m;
    }
  for (long c = 0; ok && c < nleaf; c++) {
    struct NodeDiagonal *dg = &t->innermost_leaves[c]->data.diagonal;
    const size_t cnt = (size_t)ints[c] * (size_t)ints[c];
    dg->m = ints[c];
    dg->data = malloc(cnt ? cnt * sizeof(double) : 8);
    ok = dg->data && read_all(f, dg->data, cnt * sizeof(double));
  }
  for (int l = 1; ok && l <= h; l++)
    for (long j = 0; ok && j < (1L << l); j++) {
      struct NodeOffDiagonal *od = block_of(lev, l, j);
      long r0, m, c0, nn;
      node_range(off, h, l, j, &r0, &m);
      node_range(off, h, l, j ^ 1, &c0, &nn);
      od->m = (int)m;
      od->n = (int)nn;
      od->s = ints[nleaf + (1L << l) - 2 + j];
      const size_t us = (size_t)m * od->s, vs = (size_t)nn * od->s;
      od->u = malloc(us ? us * sizeof(double) : 8);
      od->v = malloc(vs ? vs * sizeof(double) : 8);
      ok = od->u && od->v;
    }
  for (int side = 0; side < 2 && ok; side++)
    for (int l = 1; l <= h && ok; l++)
      for (long j = 0; j < (1L << l) && ok; j++) {
        struct NodeOffDiagonal *od = block_of(lev, l, j);
        const size_t cnt = (size_t)(side ? od->n : od->m) * (size_t)od->s;
        ok = read_all(f, side ? od->v : od->u, cnt * sizeof(double));
      }
  fclose(f);
  free(lev);
  free(ints);
  free(off);
  if (!ok) {
    drop_tree(&t);
    *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  *ierr = SUCCESS;
  return t;
}

/* The reference's test format (tests/utils/io.c:15-81): one matrix row per
 * line, entries separated by spaces or tabs; the first line's entry count
 * is n and n x n values are read row by row into column-major storage. */
HM_EXPORT double *hmat_amd_read_dense_text(const char *path, int *n_out, int *ierr) {
  *ierr = INPUT_ERROR;
  FILE *f = path ? fopen(path, "r") : NULL;
  if (!f) return NULL;
  int n = 0, in_token = 0, ch;
  while ((ch = fgetc(f)) != '\n') {
    if (ch == EOF) {
      fclose(f);
      return NULL;
    }
    if (ch == ' ' || ch == '\t') {
      if (in_token) n++;
      in_token = 0;
    } else if (ch != '\r') {
      in_token = 1;
    }
  }
  if (in_token) n++;
  rewind(f);
  double *a = n > 0 ? malloc((size_t)n * (size_t)n * sizeof(double)) : NULL;
  if (!a) {
    fclose(f);
    if (n > 0) *ierr = ALLOCATION_FAILURE;
    return NULL;
  }
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++)
      if (fscanf(f, "%lf", &a[i + j * (long)n]) != 1) {
        free(a);
        fclose(f);
        return NULL;
      }
  fclose(f);
  *n_out = n;
  *ierr = SUCCESS;
  return a;
}
