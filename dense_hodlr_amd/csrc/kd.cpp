/*
 * kd ordering of point sets and the user partition that follows it, on the
 * host (rigidBody/MobFunctions.h:263-430):
 *
 *  kd_sort   get_KDTree_Sorted(locations, index): the segment is sorted by
 *            coordinate depth % dim with the reference's top-down merge sort
 *            (halves of n/2 and n - n/2; on equal keys the merge takes the
 *            right list's point first, MobFunctions.h:263-289), split into
 *            its first n/2 points and the rest, and each part is ordered the
 *            same way one depth further;
 *  kd_tree   get_KDTree_Sorted(locations, pointsPerSphere, rankUpperBound,
 *            LR_Method): the same ordering, recording one partition node per
 *            segment: split = u * (first point of the right part) - 1 in
 *            matrix rows (u unknowns per point), rank bound rankUpperBound on
 *            both blocks, the given method; single points are leaves.
 *
 * Orders are permutations: position i of the ordered list holds point
 * perm[i].
 */
#include "kd.h"

#include <stdexcept>
#include <vector>

namespace hodlr {

namespace {

struct Kd {
  const double* pts;
  int dim;
  std::vector<int> tmp;

  double key(int p, int c) const { return pts[(size_t)p * dim + c]; }

  /* mergeSort (MobFunctions.h:297-330) of perm[a, b) by coordinate c */
  void merge_sort(int* perm, int n, int c) {
    if (n <= 1) return;
    const int nl = n / 2, nr = n - nl;
    merge_sort(perm, nl, c);
    merge_sort(perm + nl, nr, c);
    tmp.assign(perm, perm + n);
    const int* l1 = tmp.data();
    const int* l2 = tmp.data() + nl;
    int i = 0, j = 0, o = 0;
    while (i < nl && j < nr) {
      if (key(l1[i], c) < key(l2[j], c))
        perm[o++] = l1[i++];
      else
        perm[o++] = l2[j++];  // equal keys: the right list first
    }
    while (i < nl) perm[o++] = l1[i++];
    while (j < nr) perm[o++] = l2[j++];
  }

  void order(int* perm, int n, int depth) {
    if (n <= 1) return;
    merge_sort(perm, n, depth % dim);
    const int nl = n / 2;
    order(perm, nl, depth + 1);
    order(perm + nl, n - nl, depth + 1);
  }

  /* returns the index of the node made for the segment */
  int tree(int* perm, int n, int depth, int start, int upp, int max_rank, int lr,
           std::vector<UserNode>& out) {
    const int me = (int)out.size();
    out.emplace_back();
    UserNode& u0 = out.back();
    u0.min_rank[0] = u0.min_rank[1] = -1;
    u0.max_rank[0] = u0.max_rank[1] = max_rank;
    u0.lr = lr;
    if (n == 1) {
      u0.split = -1;
      return me;
    }
    const int nl = n / 2;
    u0.split = (int)((long long)upp * (nl + start) - 1);
    merge_sort(perm, n, depth % dim);
    const int l = tree(perm, nl, depth + 1, start, upp, max_rank, lr, out);
    const int r = tree(perm + nl, n - nl, depth + 1, start + nl, upp, max_rank, lr, out);
    out[me].left = l;
    out[me].right = r;
    return me;
  }
};

void check(const double* pts, int n, int dim) {
  if (!pts || n <= 0) throw std::invalid_argument("points must be non-empty");
  if (dim < 1) throw std::invalid_argument("dim must be >= 1");
}

}  // namespace

void kd_sort(const double* pts, int n, int dim, int* perm) {
  check(pts, n, dim);
  for (int i = 0; i < n; i++) perm[i] = i;
  Kd k{pts, dim, {}};
  k.order(perm, n, 0);
}

std::vector<UserNode> kd_tree(const double* pts, int n, int dim, int unknowns_per_point,
                              int max_rank, int lr, int* perm) {
  check(pts, n, dim);
  if (unknowns_per_point < 1) throw std::invalid_argument("unknowns per point must be >= 1");
  if ((long long)unknowns_per_point * n > 0x7fffffffLL)
    throw std::invalid_argument("matrix size exceeds 2^31 - 1 rows");
  if (lr < -1 || lr > 1) throw std::invalid_argument("unknown low-rank method");
  for (int i = 0; i < n; i++) perm[i] = i;
  Kd k{pts, dim, {}};
  std::vector<UserNode> out;
  out.reserve((size_t)2 * n);
  k.tree(perm, n, 0, 0, unknowns_per_point, max_rank, lr, out);
  return out;
}

}  // namespace hodlr
