// Library-internal helpers: error plumbing, device allocations, and the
// device-resident objects (BVH, environment tables) shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hm_common.h"

namespace hm {

// Status codes (mirrors include/hmsnn.h).
enum Status { kOk = 0, kEInval = 1, kEIO = 2, kEDevice = 3, kENoMem = 4, kEUnsupported = 5 };

#define HM_TRY(expr)                \
  do {                              \
    int hm_rc_ = (expr);            \
    if (hm_rc_ != 0) return hm_rc_; \
  } while (0)

#define HM_HIP(expr, errstr)                                                          \
  do {                                                                                \
    hipError_t hm_e_ = (expr);                                                        \
    if (hm_e_ != hipSuccess) {                                                        \
      (errstr) = std::string(#expr) + ": " + hipGetErrorString(hm_e_);              \
      return hm_e_ == hipErrorOutOfMemory ? ::hm::kENoMem : ::hm::kEDevice;           \
    }                                                                                 \
  } while (0)

inline int check_launch(const char* what, std::string& err) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    err = std::string("kernel launch failed (") + what + "): " + hipGetErrorString(e);
    return kEDevice;
  }
  return kOk;
}

inline int dev_alloc(void** p, size_t bytes, std::string& err) {
  *p = nullptr;
  if (bytes == 0) return kOk;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    err = "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? kENoMem : kEDevice;
  }
  return kOk;
}

inline void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

// Scope-bound list of device allocations (build scratch).
struct DeviceArena {
  std::vector<void*> ptrs;
  int alloc(void** p, size_t bytes, std::string& err) {
    HM_TRY(dev_alloc(p, bytes, err));
    if (*p) ptrs.push_back(*p);
    return kOk;
  }
  ~DeviceArena() {
    (void)hipDeviceSynchronize();
    for (void* p : ptrs) dev_free(p);
  }
};

struct BvhNode;
struct BvhNodeWide;

struct DeviceBvh {
  BvhNode* nodes = nullptr;    // binary tree (build intermediate, kept for tools)
  BvhNodeWide* nodes_w = nullptr;  // wide tree that traversal reads
  float4* segs = nullptr;      // 2 per leaf record (a segment, or a copy of it per piece), in leaf order
  int num_segs = 0;    // leaf records
  int num_prims = 0;   // primitives (segments, curves or triangles)
  int num_nodes = 0;   // binary inner nodes
  int num_nodes_w = 0;  // wide nodes
  int max_depth = 0;   // binary depth
  int depth_w = 0;      // wide depth (levels of inner nodes)
  // stack entries one ray can hold: width - 1 per wide level plus one
  int stack_bound() const;
  void release() {
    dev_free(nodes);
    dev_free(nodes_w);
    dev_free(segs);
    nodes = nullptr;
    nodes_w = nullptr;
    segs = nullptr;
    num_segs = num_prims = num_nodes = num_nodes_w = max_depth = depth_w = 0;
  }
  ~DeviceBvh() { release(); }
};

// Binary tree builders, collapsed into the same wide layout: the Karras
// radix tree over Morton codes, or PLOC clustering in Morton order.
constexpr int kBvhLbvh = 0, kBvhPloc = 1, kBvhSah = 2;
// Binned SAH binary tree over n boxes (6 floats each: lo xyz, hi xyz), built
// on the host (hm_bvh_sah.cpp): child[2 (n - 1)] (>= 0 an inner node, < 0 a
// leaf run coded as leaf_code(first position, count); root 0), order[n] (leaf
// position -> box index), parent_inner[n - 1] (-1 at the root),
// parent_leaf[n] (the parent of a run's first position; -1 for the run's
// other positions). Subtrees of at most max_leaf boxes become one run.
void sah_build(const float* boxes, int n, int* child, int* order, int* parent_inner, int* parent_leaf,
               int max_leaf = 1);
// max_leaf applies to the SAH builder; the others make one-primitive leaves.
// split > 1 (SAH builder, linear segments only) builds the tree over `split`
// boxes per segment, one per equal piece of its length, each padded by the
// radius; every leaf of a piece holds the whole segment's record, so a
// segment can sit in several leaves and hits are those of the unsplit tree.
int bvh_build(const float* seg_dev, int n, DeviceBvh& out, hipStream_t stream, std::string& err,
              int builder = kBvhLbvh, int ploc_radius = 16, int max_leaf = 1, int split = 1);

// Environment map and its importance-sampling tables (RGBA32F map, FP32
// tables produced from FP64 prefix sums).
struct DeviceEnv {
  int width = 0, height = 0;
  float4* tex = nullptr;   // width*height
  float* c_pdf = nullptr;  // (width+1)*height
  float* c_cdf = nullptr;
  float* m_pdf = nullptr;  // height+1
  float* m_cdf = nullptr;
  int* m_guide = nullptr;  // kEnvGuideBins + 1
  int* c_guide = nullptr;  // (kEnvGuideBins + 1) * height
  void release() {
    dev_free(tex);
    dev_free(c_pdf);
    dev_free(c_cdf);
    dev_free(m_pdf);
    dev_free(m_cdf);
    dev_free(m_guide);
    dev_free(c_guide);
    tex = nullptr;
    c_pdf = c_cdf = m_pdf = m_cdf = nullptr;
    m_guide = c_guide = nullptr;
    width = height = 0;
  }
  ~DeviceEnv() { release(); }
};

// Builds the tables on the GPU from an RGBA32F map already in device memory.
int env_build_tables(DeviceEnv& env, hipStream_t stream, std::string& err);

}  // namespace hm
