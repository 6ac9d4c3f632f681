/* Internal header of the host C layer of libhmat_lib. */
#ifndef HM_HOST_H
#define HM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "hm_hip.h"
#include "hmat_lib/error.h"
#include "hmat_lib/hodlr.h"

#define HM_EXPORT __attribute__((visibility("default")))

/* Side table of device copies, keyed by (internal node, height). */
hm_dtree *hm_side_get(const struct HODLRInternalNode *node, int height,
                      int *ierr);
void hm_side_put(const struct HODLRInternalNode *node, int height,
                 hm_dtree *dt);
void hm_side_drop_tree(const struct TreeHODLR *tree);
uint64_t hm_fingerprint(const struct HODLRInternalNode *node, int height);

/* Internal nodes of a subtree per depth, left to right (malloc'd, 2^h - 1
 * pointers; depth d starts at index 2^d - 1). */
struct HODLRInternalNode **hm_levels(const struct HODLRInternalNode *root,
                                     int height);

/* Report a device-layer failure on stderr and return its code. */
int hm_report(int rc, const char *where);

/* Process-wide device staging buffers for host-array entry points. */
int hm_staging(size_t bytes_a, size_t bytes_b, double **a, double **b);
void hm_staging_lock(void);
void hm_staging_unlock(void);

#endif
