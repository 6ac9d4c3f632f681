/* Side table of device-resident copies of host trees, plus shared helpers
 * of the host layer.  Not in the reference (which has no device); see
 * include/hmat_lib_amd.h for the contract. */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hm_host.h"
#include "hmat_lib_amd.h"

struct side_entry {
  const struct HODLRInternalNode *node;
  int height;
  uint64_t fp;
  hm_dtree *dt;
  struct side_entry *next;
};

static struct side_entry *g_head = NULL;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

struct HODLRInternalNode **hm_levels(const struct HODLRInternalNode *root,
                                     int height) {
  long total = (1L << height) - 1;
  struct HODLRInternalNode **all = malloc((size_t)total * sizeof(*all));
  if (!all) return NULL;
  all[0] = (struct HODLRInternalNode *)root;
  for (int d = 0; d + 1 < height; d++) {
    long base = (1L << d) - 1, nbase = (1L << (d + 1)) - 1;
    for (long i = 0; i < (1L << d); i++) {
      all[nbase + 2 * i] = all[base + i]->children[0].internal;
      all[nbase + 2 * i + 1] = all[base + i]->children[3].internal;
    }
  }
  return all;
}

static uint64_t fnv(uint64_t h, uint64_t v) {
  for (int i = 0; i < 8; i++) {
    h ^= (v >> (8 * i)) & 0xff;
    h *= 0x100000001b3ULL;
  }
  return h;
}

/* Eight entries of an array of `count` doubles, spread over it (first and
 * last included): a refilled array at a reused address changes them. */
static uint64_t fnv_samples(uint64_t h, const double *p, long count) {
  if (!p || count <= 0) return h;
  for (int k = 0; k < 8; k++) {
    const long i = (long)((double)(count - 1) * k / 7.0);
    uint64_t bits;
    memcpy(&bits, p + i, sizeof(bits));
    h = fnv(h, bits);
  }
  return h;
}

uint64_t hm_fingerprint(const struct HODLRInternalNode *node, int height) {
  struct HODLRInternalNode **all = hm_levels(node, height);
  if (!all) return 0;
  uint64_t h = 0xcbf29ce484222325ULL;
  long total = (1L << height) - 1;
  for (long k = 0; k < total; k++) {
    h = fnv(h, (uint64_t)all[k]->m);
    for (int c = 1; c < 3; c++) {
      const struct NodeOffDiagonal *o = &all[k]->children[c].leaf->data.off_diagonal;
      h = fnv(h, (uint64_t)(uintptr_t)o->u);
      h = fnv(h, (uint64_t)(uintptr_t)o->v);
      h = fnv(h, ((uint64_t)(uint32_t)o->m << 32) | (uint32_t)o->n);
      h = fnv(h, (uint64_t)(uint32_t)o->s);
      h = fnv_samples(h, o->u, (long)o->m * o->s);
      h = fnv_samples(h, o->v, (long)o->n * o->s);
    }
  }
  long first_bottom = (1L << (height - 1)) - 1;
  for (long k = first_bottom; k < total; k++) {
    for (int c = 0; c < 4; c += 3) {
      const struct NodeDiagonal *d = &all[k]->children[c].leaf->data.diagonal;
      h = fnv(h, (uint64_t)(uintptr_t)d->data);
      h = fnv(h, (uint64_t)(uint32_t)d->m);
      h = fnv_samples(h, d->data, (long)d->m * d->m);
    }
  }
  free(all);
  return h;
}

static struct side_entry **find_locked(const struct HODLRInternalNode *node,
                                       int height) {
  struct side_entry **pp = &g_head;
  while (*pp) {
    if ((*pp)->node == node && (*pp)->height == height) return pp;
    pp = &(*pp)->next;
  }
  return NULL;
}

hm_dtree *hm_side_get(const struct HODLRInternalNode *node, int height,
                      int *ierr) {
  uint64_t fp = hm_fingerprint(node, height);
  pthread_mutex_lock(&g_mu);
  struct side_entry **pp = find_locked(node, height);
  if (pp && (*pp)->fp == fp) {
    hm_dtree *dt = (*pp)->dt;
    pthread_mutex_unlock(&g_mu);
    *ierr = SUCCESS;
    return dt;
  }
  if (pp) {
    struct side_entry *e = *pp;
    *pp = e->next;
    hm_hip_tree_destroy(e->dt);
    free(e);
  }
  pthread_mutex_unlock(&g_mu);
  hm_dtree *dt = NULL;
  int rc = hm_hip_tree_upload(node, height, &dt, NULL);
  if (rc != SUCCESS) {
    *ierr = hm_report(rc, "tree upload");
    return NULL;
  }
  hm_side_put(node, height, dt);
  *ierr = SUCCESS;
  return dt;
}

void hm_side_put(const struct HODLRInternalNode *node, int height,
                 hm_dtree *dt) {
  uint64_t fp = hm_fingerprint(node, height);
  pthread_mutex_lock(&g_mu);
  struct side_entry **pp = find_locked(node, height);
  if (pp) {
    if ((*pp)->dt != dt) hm_hip_tree_destroy((*pp)->dt);
    (*pp)->dt = dt;
    (*pp)->fp = fp;
  } else {
    struct side_entry *e = malloc(sizeof(*e));
    if (e) {
      e->node = node;
      e->height = height;
      e->fp = fp;
      e->dt = dt;
      e->next = g_head;
      g_head = e;
    } else {
      hm_hip_tree_destroy(dt);
    }
  }
  pthread_mutex_unlock(&g_mu);
}

void hm_side_drop_tree(const struct TreeHODLR *tree) {
  if (!tree || !tree->root || tree->height < 1) return;
  struct HODLRInternalNode **all = hm_levels(tree->root, tree->height);
  if (!all) return;
  long total = (1L << tree->height) - 1;
  pthread_mutex_lock(&g_mu);
  struct side_entry **pp = &g_head;
  while (*pp) {
    int hit = 0;
    for (long k = 0; k < total && !hit; k++) hit = ((*pp)->node == all[k]);
    if (hit) {
      struct side_entry *e = *pp;
      *pp = e->next;
      hm_hip_tree_destroy(e->dt);
      free(e);
    } else {
      pp = &(*pp)->next;
    }
  }
  pthread_mutex_unlock(&g_mu);
  free(all);
}

int hm_report(int rc, const char *where) {
  if (rc != SUCCESS)
    fprintf(stderr, "hmat_lib (MI355X): %s failed: %s\n", where,
            hm_hip_last_error());
  return rc;
}

static pthread_mutex_t g_stage_mu = PTHREAD_MUTEX_INITIALIZER;
static double *g_sa = NULL, *g_sb = NULL;
static size_t g_na = 0, g_nb = 0;

void hm_staging_lock(void) { pthread_mutex_lock(&g_stage_mu); }
void hm_staging_unlock(void) { pthread_mutex_unlock(&g_stage_mu); }

/* Caller holds the staging lock. */
int hm_staging(size_t bytes_a, size_t bytes_b, double **a, double **b) {
  if (bytes_a > g_na) {
    hm_hip_free(g_sa);
    g_sa = NULL;
    g_na = 0;
    int rc = hm_hip_malloc((void **)&g_sa, bytes_a);
    if (rc != SUCCESS) return rc;
    g_na = bytes_a;
  }
  if (bytes_b > g_nb) {
    hm_hip_free(g_sb);
    g_sb = NULL;
    g_nb = 0;
    int rc = hm_hip_malloc((void **)&g_sb, bytes_b);
    if (rc != SUCCESS) return rc;
    g_nb = bytes_b;
  }
  *a = g_sa;
  *b = g_sb;
  return SUCCESS;
}

HM_EXPORT void hmat_amd_invalidate(const struct TreeHODLR *tree) {
  hm_side_drop_tree(tree);
}

HM_EXPORT hm_dtree *hmat_amd_device_tree(const struct TreeHODLR *tree,
                                         int *ierr) {
  if (!tree || !tree->root) {
    *ierr = INPUT_ERROR;
    return NULL;
  }
  return hm_side_get(tree->root, tree->height, ierr);
}

HM_EXPORT void hmat_amd_attach(struct TreeHODLR *tree, hm_dtree *dtree,
                               int *ierr) {
  if (!tree || !dtree) {
    *ierr = INPUT_ERROR;
    return;
  }
  int rc = hm_hip_tree_download(dtree, tree, NULL, NULL);
  if (rc != SUCCESS) {
    *ierr = hm_report(rc, "tree download");
    return;
  }
  hm_side_put(tree->root, tree->height, dtree);
  *ierr = SUCCESS;
}

HM_EXPORT int hmat_amd_cache_size(void) {
  pthread_mutex_lock(&g_mu);
  int n = 0;
  for (struct side_entry *e = g_head; e; e = e->next) n++;
  pthread_mutex_unlock(&g_mu);
  return n;
}
