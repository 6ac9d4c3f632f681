/*
 * Tree skeleton allocation (host only; no device memory is touched here
 * except that freeing a tree also drops its device copy).
 *
 * Replaces include/hmat_lib/allocators.h:5-49 of the reference, same
 * signatures.  With -D_TEST_HODLR the allocating/freeing calls take the
 * injected malloc/free pointers exactly as the reference's test ABI does;
 * that variant is built as libhmat_lib_test.so.
 */
#ifndef HMAT_LIB_ALLOCATORS_H
#define HMAT_LIB_ALLOCATORS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct TreeHODLR;
struct HODLRInternalNode;
struct HODLRLeafNode;

/* Reference: src/allocators.c:229-344 (nodes malloc'd one by one). */
struct TreeHODLR *allocate_tree(const int height, int *ierr);

/* Reference: src/allocators.c:404-464 (nodes in two arrays). */
#ifndef _TEST_HODLR
struct TreeHODLR *allocate_tree_monolithic(const int height, int *ierr);
#else
struct TreeHODLR *allocate_tree_monolithic(const int height, int *ierr,
                                           void *(*malloc)(size_t size),
                                           void (*free)(void *ptr));
#endif

/* Reference: src/allocators.c:747-839. */
#ifndef _TEST_HODLR
void free_tree_hodlr(struct TreeHODLR **hodlr_ptr);
#else
void free_tree_hodlr(struct TreeHODLR **hodlr_ptr, void (*free)(void *ptr));
#endif

/* Reference: src/allocators.c:659-708. */
#ifndef _TEST_HODLR
void free_tree_data(struct TreeHODLR *hodlr);
#else
void free_tree_data(struct TreeHODLR *hodlr, void (*free)(void *ptr));
#endif

/* Reference: src/allocators.c:499-512.  Sizes are in bytes. */
void compute_construct_tree_array_sizes(const int height,
                                        size_t *size_internal_nodes,
                                        size_t *size_leaf_nodes,
                                        size_t *size_work_queue,
                                        size_t *size_innermost_leaves);

/* Reference: src/allocators.c:561-635. */
void construct_tree(const int height, struct TreeHODLR *hodlr,
                    struct HODLRInternalNode *internal_nodes,
                    struct HODLRLeafNode *leaf_nodes,
                    struct HODLRInternalNode **work_queue,
                    struct HODLRLeafNode **innermost_leaves, int *ierr);

#ifdef __cplusplus
}
#endif

#endif
