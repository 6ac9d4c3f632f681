/*
 * HODLR tree types of the hmat_lib C API.
 *
 * Drop-in contract: every struct, union and enum below has exactly the
 * member order, member types and therefore the size and layout of the
 * reference's public header (include/hmat_lib/hodlr.h:7-263), because
 * callers walk the tree and read u, v, data, m, n, s directly.  The device
 * copy of a tree lives in a side table inside libhmat_hip (keyed by the
 * TreeHODLR pointer); nothing here points to device memory.
 *
 * Storage convention (column-major FP64 throughout):
 *   NodeDiagonal.data     m x m dense diagonal leaf block.
 *   NodeOffDiagonal.u     m x s, columns already scaled by the singular
 *                         values, in descending singular-value order.
 *   NodeOffDiagonal.v     n x s, stored as V (not V^T).
 *
 * Children of an internal node, in order:
 *   [0] top-left diagonal (internal node, or a DIAGONAL leaf at the bottom)
 *   [1] top-right off-diagonal leaf   (m = size of [0], n = size of [3])
 *   [2] bottom-left off-diagonal leaf (m = size of [3], n = size of [0])
 *   [3] bottom-right diagonal (internal node, or a DIAGONAL leaf)
 */
#ifndef HMAT_LIB_HODLR_H
#define HMAT_LIB_HODLR_H

enum NodeType {
  DIAGONAL,
  OFFDIAGONAL,
  INTERNAL
};

/* Dense diagonal leaf: an m x m column-major block. */
struct NodeDiagonal {
  double *data;
  int m;
};

/* Low-rank off-diagonal leaf: block = u * v^T with u m x s and v n x s. */
struct NodeOffDiagonal {
  double *u;
  double *v;
  int m;
  int s;
  int n;
};

union HODLRData {
  struct NodeDiagonal diagonal;
  struct NodeOffDiagonal off_diagonal;
};

struct HODLRLeafNode {
  enum NodeType type;
  union HODLRData data;
  struct HODLRInternalNode *parent;
};

union HODLRNode {
  struct HODLRLeafNode *leaf;
  struct HODLRInternalNode *internal;
};

struct HODLRInternalNode {
  union HODLRNode children[4];
  struct HODLRInternalNode *parent;   /* NULL at the root */
  int m;                              /* rows (= columns) of this block */
};

/*
 * height          number of splits (= layers of internal nodes).
 * innermost_leaves the 2^height diagonal leaves, top-left to bottom-right.
 * len_work_queue  2^(height-1).
 * work_queue      scratch array of len_work_queue internal-node pointers;
 *                 it is rewritten by the library even through const trees,
 *                 so one tree must not be used by two calls at once.
 * memory_*_ptr    non-NULL when the nodes were allocated as two arrays
 *                 (allocate_tree_monolithic / construct_tree).
 */
struct TreeHODLR {
  int height;
  struct HODLRInternalNode *root;
  struct HODLRLeafNode **innermost_leaves;
  long len_work_queue;
  struct HODLRInternalNode **work_queue;
  struct HODLRInternalNode *memory_internal_ptr;
  struct HODLRLeafNode *memory_leaf_ptr;
};

#endif
