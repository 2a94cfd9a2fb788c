// Binned SAH binary tree over primitive boxes, built on the host: the
// alternative to the GPU's Morton-order radix tree (accel.builder "sah").
// Top-down: the centroids of a node's primitives are binned (kBins per
// axis, all three axes), the split plane with the smallest
// area(left) x count(left) + area(right) x count(right) is taken, and the
// primitives are partitioned in place; a subtree of at most max_leaf
// primitives becomes one leaf (a run of consecutive positions). A node whose
// centroids fall into one bin is split at the median along the longest
// centroid axis. Subtrees above kTaskMin primitives are built on their own
// threads. The result is the layout the GPU refit, emit and collapse steps
// read (see the declaration in hm_internal.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <future>
#include <thread>
#include <vector>

#include "hm_internal.h"

namespace hm {

namespace {

constexpr int kBins = 32;
constexpr int kTaskMin = 1 << 15;

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = 1e30f;
      hi[a] = -1e30f;
    }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  double half_area() const {
    const double dx = std::max(0.f, hi[0] - lo[0]), dy = std::max(0.f, hi[1] - lo[1]),
                 dz = std::max(0.f, hi[2] - lo[2]);
    return dx * dy + dy * dz + dz * dx;
  }
};

struct Builder {
  const float* boxes;  // 6 floats per primitive
  int* idx;            // permutation being partitioned: leaf position -> primitive
  int* child;
  int* parent_inner;
  int* parent_leaf;
  std::atomic<int> next_node{1};
  std::atomic<int> tasks{0};
  int max_tasks = 1;
  int max_leaf = 1;

  float centroid(int p, int a) const { return 0.5f * (boxes[6 * (size_t)p + a] + boxes[6 * (size_t)p + 3 + a]); }

  // Splits [b, e) and returns the position of the first right primitive.
  int split(int b, int e) {
    float cmin[3] = {1e30f, 1e30f, 1e30f}, cmax[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = b; i < e; ++i)
      for (int a = 0; a < 3; ++a) {
        const float c = centroid(idx[i], a);
        cmin[a] = std::min(cmin[a], c);
        cmax[a] = std::max(cmax[a], c);
      }
    Box bin[3][kBins];
    int cnt[3][kBins] = {};
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < kBins; ++k) bin[a][k].reset();
    float scale[3];
    for (int a = 0; a < 3; ++a) scale[a] = cmax[a] > cmin[a] ? (float)kBins / (cmax[a] - cmin[a]) : 0.f;
    for (int i = b; i < e; ++i) {
      const int p = idx[i];
      Box pb;
      for (int a = 0; a < 3; ++a) {
        pb.lo[a] = boxes[6 * (size_t)p + a];
        pb.hi[a] = boxes[6 * (size_t)p + 3 + a];
      }
      for (int a = 0; a < 3; ++a) {
        const int k = std::min(kBins - 1, std::max(0, (int)((centroid(p, a) - cmin[a]) * scale[a])));
        bin[a][k].grow(pb);
        cnt[a][k]++;
      }
    }
    double best = 1e300;
    int best_axis = -1, best_k = -1;
    for (int a = 0; a < 3; ++a) {
      if (scale[a] == 0.f) continue;
      double right_area[kBins];
      int right_cnt[kBins];
      Box acc;
      acc.reset();
      int n = 0;
      for (int k = kBins - 1; k >= 1; --k) {
        acc.grow(bin[a][k]);
        n += cnt[a][k];
        right_area[k] = acc.half_area();
        right_cnt[k] = n;
      }
      acc.reset();
      n = 0;
      for (int k = 0; k < kBins - 1; ++k) {
        acc.grow(bin[a][k]);
        n += cnt[a][k];
        if (n == 0 || right_cnt[k + 1] == 0) continue;
        const double c = acc.half_area() * n + right_area[k + 1] * right_cnt[k + 1];
        if (c < best) {
          best = c;
          best_axis = a;
          best_k = k;
        }
      }
    }
    if (best_axis >= 0) {
      const int a = best_axis;
      const float lo = cmin[a], sc = scale[a];
      int* m = std::partition(idx + b, idx + e, [&](int p) {
        return std::min(kBins - 1, std::max(0, (int)((centroid(p, a) - lo) * sc))) <= best_k;
      });
      const int mid = (int)(m - idx);
      if (mid > b && mid < e) return mid;
    }
    // every centroid in one bin: the median along the longest centroid axis
    int a = 0;
    for (int k = 1; k < 3; ++k)
      if (cmax[k] - cmin[k] > cmax[a] - cmin[a]) a = k;
    const int mid = b + (e - b) / 2;
    std::nth_element(idx + b, idx + mid, idx + e, [&](int x, int y) {
      const float cx = centroid(x, a), cy = centroid(y, a);
      return cx < cy || (cx == cy && x < y);
    });
    return mid;
  }

  // Builds the subtree of [b, e) under inner node `node` (e - b >= 2).
  void build(int node, int b, int e) {
    const int mid = split(b, e);
    int c[2];
    const int rb[2] = {b, mid}, re[2] = {mid, e};
    for (int s = 0; s < 2; ++s) {
      if (re[s] - rb[s] <= max_leaf) {
        c[s] = leaf_code(rb[s], re[s] - rb[s]);
        parent_leaf[rb[s]] = node;
        for (int p = rb[s] + 1; p < re[s]; ++p) parent_leaf[p] = -1;
      } else {
        c[s] = next_node.fetch_add(1);
        parent_inner[c[s]] = node;
      }
    }
    child[2 * node] = c[0];
    child[2 * node + 1] = c[1];
    std::future<void> side;
    bool async = false;
    if (c[0] >= 0 && mid - b >= kTaskMin && tasks.fetch_add(1) < max_tasks) {
      side = std::async(std::launch::async, [this, c, b, mid] { build(c[0], b, mid); });
      async = true;
    } else if (c[0] >= 0) {
      if (mid - b >= kTaskMin) tasks.fetch_sub(1);
      build(c[0], b, mid);
    }
    if (c[1] >= 0) build(c[1], mid, e);
    if (async) {
      side.get();
      tasks.fetch_sub(1);
    }
  }
};

}  // namespace

void sah_build(const float* boxes, int n, int* child, int* order, int* parent_inner, int* parent_leaf,
               int max_leaf) {
  for (int i = 0; i < n; ++i) order[i] = i;
  if (n < 2) return;
  Builder bd;
  bd.max_leaf = std::max(1, std::min(max_leaf, kMaxLeafRun));
  bd.boxes = boxes;
  bd.idx = order;
  bd.child = child;
  bd.parent_inner = parent_inner;
  bd.parent_leaf = parent_leaf;
  bd.max_tasks = std::max(1, (int)std::thread::hardware_concurrency() * 2);
  parent_inner[0] = -1;
  bd.build(0, 0, n);
}

}  // namespace hm
