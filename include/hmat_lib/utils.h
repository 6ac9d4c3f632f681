/* Highest off-diagonal rank of a tree.  Reference src/utils.c:5-26 and
 * include/hmat_lib/utils.h:4. */
#ifndef HMAT_LIB_UTILS_H
#define HMAT_LIB_UTILS_H

#ifdef __cplusplus
extern "C" {
#endif

struct TreeHODLR;

int get_highest_s(const struct TreeHODLR *const hodlr);

#ifdef __cplusplus
}
#endif

#endif
