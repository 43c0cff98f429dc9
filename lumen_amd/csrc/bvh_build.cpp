// The navigator's BVH builder on the host, native: binned surface-area-heuristic splits, stored depth first with
// skip links (the reference's navigator-array layout, src/lumen/geometry/bvh.py:1-10, built by its
// src/lumen/geometry/bvh.py:179-210 `build_many` from the same rules: the cheapest of the bins - 1 planes per axis
// over the centroids' bins, the first axis and plane on ties, a stable partition; every centroid coincident: the
// list's halves; a leaf at <= leaf_size boxes).  Each node is one pass over its boxes: the bins' counts and boxes,
// then prefix and suffix sweeps give every plane's cost (count x surface area of each side) in O(bins).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../include/lumen_hip.h"

namespace {

struct Box {
  double lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) lo[a] = std::numeric_limits<double>::infinity(), hi[a] = -lo[a];
  }
  void grow(const Box &b) {
    for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], b.lo[a]), hi[a] = std::max(hi[a], b.hi[a]);
  }
  double area() const {
    double e[3];
    for (int a = 0; a < 3; ++a) e[a] = std::max(hi[a] - lo[a], 0.0);
    return 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
  }
};

struct Builder {
  const double *lo, *hi;
  int leaf_size, bins;
  std::vector<Box> nodes;
  std::vector<int32_t> skip, start, count, items;
  int depth = 0;

  Box box_of(int i) const {
    Box b;
    for (int a = 0; a < 3; ++a) b.lo[a] = lo[3 * i + a], b.hi[a] = hi[3 * i + a];
    return b;
  }

  // the best split of idx[0..n) into [0, nl) and [nl, n) (reordered in place, stably); false: make a leaf
  int split(std::vector<int> &idx, int b0, int n, std::vector<int> &scratch) {
    double clo[3], chi[3];
    for (int a = 0; a < 3; ++a) clo[a] = std::numeric_limits<double>::infinity(), chi[a] = -clo[a];
    for (int k = 0; k < n; ++k) {
      const int i = idx[b0 + k];
      for (int a = 0; a < 3; ++a) {
        const double c = 0.5 * (lo[3 * i + a] + hi[3 * i + a]);
        clo[a] = std::min(clo[a], c);
        chi[a] = std::max(chi[a], c);
      }
    }
    double best = std::numeric_limits<double>::infinity();
    int best_ax = -1, best_k = 0;
    std::vector<int> cnt(bins);
    std::vector<Box> bb(bins), suf(bins);
    for (int ax = 0; ax < 3; ++ax) {
      const double ext = chi[ax] - clo[ax];
      if (!(ext > 0.0)) continue;
      std::fill(cnt.begin(), cnt.end(), 0);
      for (auto &b : bb) b.reset();
      for (int k = 0; k < n; ++k) {
        const int i = idx[b0 + k];
        const double c = 0.5 * (lo[3 * i + ax] + hi[3 * i + ax]);
        const int b = std::min((int)((c - clo[ax]) / ext * bins), bins - 1);
        ++cnt[b];
        bb[b].grow(box_of(i));
      }
      suf[bins - 1] = bb[bins - 1];
      for (int b = bins - 2; b >= 0; --b) {
        suf[b] = suf[b + 1];
        suf[b].grow(bb[b]);
      }
      Box left;
      left.reset();
      int nl = 0;
      for (int k = 1; k < bins; ++k) {  // plane k: bins below k on the left
        left.grow(bb[k - 1]);
        nl += cnt[k - 1];
        if (nl == 0 || nl == n) continue;
        const double cost = nl * left.area() + (n - nl) * suf[k].area();
        if (cost < best) best = cost, best_ax = ax, best_k = k;
      }
    }
    if (best_ax < 0) return n / 2;  // every centroid coincides: the list's halves
    const double ext = chi[best_ax] - clo[best_ax];
    int nl = 0, nr = 0;
    scratch.resize(n);
    for (int k = 0; k < n; ++k) {  // stable partition: the left in order, then the right in order
      const int i = idx[b0 + k];
      const double c = 0.5 * (lo[3 * i + best_ax] + hi[3 * i + best_ax]);
      const int b = std::min((int)((c - clo[best_ax]) / ext * bins), bins - 1);
      if (b < best_k) idx[b0 + nl++] = i;
      else scratch[nr++] = i;
    }
    std::copy(scratch.begin(), scratch.begin() + nr, idx.begin() + b0 + nl);
    return nl;
  }

  void build(std::vector<int> &idx, int b0, int n, int d, std::vector<int> &scratch) {
    depth = std::max(depth, d);
    const int me = (int)nodes.size();
    Box b;
    b.reset();
    for (int k = 0; k < n; ++k) b.grow(box_of(idx[b0 + k]));
    nodes.push_back(b);
    skip.push_back(0);
    start.push_back(0);
    count.push_back(0);
    if (n <= leaf_size) {
      start[me] = (int32_t)items.size();
      count[me] = n;
      for (int k = 0; k < n; ++k) items.push_back(idx[b0 + k]);
    } else {
      const int nl = split(idx, b0, n, scratch);
      build(idx, b0, nl, d + 1, scratch);
      build(idx, b0 + nl, n - nl, d + 1, scratch);
    }
    skip[me] = (int32_t)nodes.size();
  }
};

}  // namespace

extern "C" __attribute__((visibility("default"))) int lh_bvh_build(const double *lo, const double *hi, int32_t n,
                                                                   int32_t leaf_size, int32_t bins, double *node_lo,
                                                                   double *node_hi, int32_t *skip, int32_t *start,
                                                                   int32_t *count, int32_t *items, int32_t *n_nodes,
                                                                   int32_t *depth) {
  if (!lo || !hi || n < 1 || leaf_size < 1 || leaf_size > 7 || bins < 2 || bins > 1024 || !node_lo || !node_hi ||
      !skip || !start || !count || !items || !n_nodes || !depth)
    return LH_ERR_ARG;
  Builder b{lo, hi, leaf_size, bins, {}, {}, {}, {}, {}};
  b.nodes.reserve(2 * (size_t)n);
  std::vector<int> idx(n), scratch;
  for (int i = 0; i < n; ++i) idx[i] = i;
  b.build(idx, 0, n, 1, scratch);
  const int m = (int)b.nodes.size();
  for (int k = 0; k < m; ++k) {
    for (int a = 0; a < 3; ++a) node_lo[3 * k + a] = b.nodes[k].lo[a], node_hi[3 * k + a] = b.nodes[k].hi[a];
    skip[k] = b.skip[k];
    start[k] = b.start[k];
    count[k] = b.count[k];
  }
  std::copy(b.items.begin(), b.items.end(), items);
  *n_nodes = m;
  *depth = b.depth;
  return LH_OK;
}
