/*
 * kd ordering and kd partitions of point sets (see kd.cpp).
 */
#ifndef DENSE_HODLR_KD_H
#define DENSE_HODLR_KD_H

#include <vector>

#include "engine.h"

namespace hodlr {

/* perm (n entries): the kd order of n points of dim coordinates (point-major). */
void kd_sort(const double* pts, int n, int dim, int* perm);
/* The kd order and its partition: one node per segment (2n - 1, root first,
 * children by index); split rows counted in unknowns_per_point per point. */
std::vector<UserNode> kd_tree(const double* pts, int n, int dim, int unknowns_per_point,
                              int max_rank, int lr, int* perm);

}  // namespace hodlr

#endif
