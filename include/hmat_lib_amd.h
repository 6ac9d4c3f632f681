/*
 * Extensions of libhmat_lib (MI355X build) beyond the reference API.
 *
 * The reference has no device: every call of its API hands over host
 * arrays.  This build keeps that API unchanged (the hmat_lib headers) and
 * keeps, per tree, a device copy in a side table keyed by the node pointer
 * and height.  An entry is reused while the tree's fingerprint is unchanged
 * and is dropped by free_tree_data / free_tree_hodlr.  The fingerprint
 * covers node sizes, ranks, the u/v/data pointer values and eight sampled
 * entries of every u/v/data array (first and last included), so replacing
 * an array (free + malloc, even at the same address) and refilling it is
 * detected.  Callers that overwrite entries of u/v/data in place must call
 * hmat_amd_invalidate: an edit that misses the sampled entries is not seen
 * and the products would use the old device copy.
 *
 * No counterpart in the reference; these are for callers who want the
 * device-resident path (inputs already in HBM).
 */
#ifndef HMAT_LIB_AMD_H
#define HMAT_LIB_AMD_H

#include "hm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

struct TreeHODLR;

/* Drop the cached device copy of a tree (after in-place edits). */
void hmat_amd_invalidate(const struct TreeHODLR *tree);

/* The device copy of a host tree (uploaded on first use, then cached and
 * owned by the side table). */
hm_dtree *hmat_amd_device_tree(const struct TreeHODLR *tree, int *ierr);

/* Fill an empty host tree of the same height from a device tree (malloc'd
 * u/v/data as dense_to_tree_hodlr would) and cache the device tree for it;
 * ownership of `dtree` passes to the side table. */
void hmat_amd_attach(struct TreeHODLR *tree, hm_dtree *dtree, int *ierr);

/* Number of cached device trees (diagnostics). */
int hmat_amd_cache_size(void);

/* Binary HODLR files (one layout for host and device trees, see
 * hm_hip_tree_save): header (magic "HODLRBIN", version 1, byte-order mark,
 * height, n, leaf sizes, block ranks in level-major order), then the
 * diagonal leaves, all U panels and all V panels, column-major FP64.
 * Save returns an ErrorCode (INPUT_ERROR on a bad tree or an I/O error);
 * load returns a tree allocated as allocate_tree does with malloc'd
 * arrays (free with free_tree_hodlr), or NULL with *ierr set (INPUT_ERROR
 * for a missing, malformed or truncated file).  No counterpart in the
 * reference. */
int hmat_amd_save_tree(const struct TreeHODLR *tree, const char *path);
struct TreeHODLR *hmat_amd_load_tree(const char *path, int *ierr);

/* Dense n x n matrix in the reference's text test format (one row per
 * line, whitespace-separated; n = the first line's entry count), returned
 * malloc'd column-major; NULL with *ierr = INPUT_ERROR on a missing or
 * malformed file.  Replaces tests/utils/io.c:15-81 (read_dense_matrix). */
double *hmat_amd_read_dense_text(const char *path, int *n, int *ierr);

#ifdef __cplusplus
}
#endif

#endif
