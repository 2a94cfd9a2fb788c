// LBVH build over hair segments, on the GPU:
//   1. segment boxes and centroid bounds
//   2. 30-bit Morton codes of the centroids, radix sort (rocPRIM)
//   3. Karras (2012) radix-tree hierarchy, one thread per internal node
//   4. bottom-up box refit with per-node arrival counters
//   5. emit 64-byte traversal nodes and Morton-ordered segment records
// This replaces the OptiX curve acceleration structure the reference builds
// with owlCurvesGeomGroupCreate + owlGroupBuildAccel
// (render_path_tracing.cu:236-256, render_hair_msnn.cu:370-413).
#include <hip/hip_runtime.h>

#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "hm_bvh.h"
#include "hm_internal.h"

namespace hm {

namespace {

__device__ __forceinline__ unsigned expand_bits10(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// order-preserving float <-> int mapping for atomic min/max
__device__ __forceinline__ int f2ord(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

struct BuildBuffers {
  const float* seg_in;  // 8 floats per segment
  int n;                // leaves: segments x pieces
  int pieces;           // boxes per segment (leaf i is piece i % pieces of segment i / pieces)
  float* leaf_box;      // 6 floats per segment, padded
  int* cbounds;         // 6 ordered ints: min xyz, max xyz
  unsigned* keys;
  int* vals;
  unsigned* keys_sorted;
  int* vals_sorted;
  int* child;           // 2 per internal node (encoded)
  int* parent_internal; // per internal node
  int* parent_leaf;     // per leaf (sorted position)
  float* node_box;      // 6 floats per internal node
  int* arrive;          // per internal node
  int* max_depth;
};

__global__ void k_leaf_boxes(BuildBuffers b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int lmin[3] = {INT_MAX, INT_MAX, INT_MAX}, lmax[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (i < b.n) {
    const int P = b.pieces, piece = i % P;
    const float* s = b.seg_in + 8 * (size_t)(i / P);
    const float r = s[3];
    const float t0 = (float)piece / (float)P, t1 = (float)(piece + 1) / (float)P;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      // the piece's end points (the segment's own at its ends); the relative
      // padding below covers the rounding of the interpolation
      const float d = s[4 + k] - s[k];
      const float a = piece == 0 ? s[k] : fmaf(d, t0, s[k]);
      const float c = piece == P - 1 ? s[4 + k] : fmaf(d, t1, s[k]);
      const float pad = r + 1e-6f * (fabsf(a) + fabsf(c) + r) + 1e-7f;
      lo[k] = fminf(a, c) - pad;
      hi[k] = fmaxf(a, c) + pad;
      b.leaf_box[6 * (size_t)i + k] = lo[k];
      b.leaf_box[6 * (size_t)i + 3 + k] = hi[k];
      const float cen = 0.5f * (a + c);
      lmin[k] = f2ord(cen);
      lmax[k] = f2ord(cen);
    }
  }
  // wave reduction then one atomic per wave
  for (int k = 0; k < 3; ++k) {
    int vmin = lmin[k], vmax = lmax[k];
    for (int off = 32; off > 0; off >>= 1) {
      vmin = min(vmin, __shfl_xor(vmin, off));
      vmax = max(vmax, __shfl_xor(vmax, off));
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&b.cbounds[k], vmin);
      atomicMax(&b.cbounds[3 + k], vmax);
    }
  }
}

__global__ void k_morton(BuildBuffers b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  const float* s = b.seg_in + 8 * (size_t)i;
  unsigned code = 0;
  for (int k = 0; k < 3; ++k) {
    const float lo = ord2f(b.cbounds[k]), hi = ord2f(b.cbounds[3 + k]);
    const float ext = hi - lo;
    const float cen = 0.5f * (s[k] + s[4 + k]);
    float q = ext > 0.f ? (cen - lo) / ext : 0.5f;
    unsigned qi = (unsigned)fminf(fmaxf(q * 1024.f, 0.f), 1023.f);
    code |= expand_bits10(qi) << (2 - k);
  }
  b.keys[i] = code;
  b.vals[i] = i;
}

__device__ __forceinline__ int delta(const unsigned* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const unsigned ki = keys[i], kj = keys[j];
  if (ki == kj) return 32 + __clz((unsigned)(i ^ j));
  return __clz(ki ^ kj);
}

__global__ void k_karras(BuildBuffers b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = b.n;
  if (i >= n - 1) return;
  const unsigned* keys = b.keys_sorted;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0, step = l;
  do {
    step = (step + 1) >> 1;
    if (delta(keys, n, i, i + (s + step) * d) > dnode) s += step;
  } while (step > 1);
  const int gamma = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  int left, right;
  if (lo == gamma) {
    left = ~gamma;
    b.parent_leaf[gamma] = i;
  } else {
    left = gamma;
    b.parent_internal[gamma] = i;
  }
  if (hi == gamma + 1) {
    right = ~(gamma + 1);
    b.parent_leaf[gamma + 1] = i;
  } else {
    right = gamma + 1;
    b.parent_internal[gamma + 1] = i;
  }
  b.child[2 * i] = left;
  b.child[2 * i + 1] = right;
  if (i == 0) b.parent_internal[0] = -1;
}

__device__ __forceinline__ void child_box(const BuildBuffers& b, int c, float box[6]) {
  if (c < 0) {
    // a leaf run: the union of its primitives' boxes
    const int code = ~c, first = code & kLeafPosMask, count = (code >> kLeafCountShift) + 1;
    for (int k = 0; k < 3; ++k) {
      box[k] = 3.0e38f;
      box[3 + k] = -3.0e38f;
    }
    for (int j = 0; j < count; ++j) {
      const int orig = b.vals_sorted[first + j];
      for (int k = 0; k < 3; ++k) {
        box[k] = fminf(box[k], b.leaf_box[6 * (size_t)orig + k]);
        box[3 + k] = fmaxf(box[3 + k], b.leaf_box[6 * (size_t)orig + 3 + k]);
      }
    }
  } else {
    for (int k = 0; k < 6; ++k)
      box[k] = __hip_atomic_load(&b.node_box[6 * (size_t)c + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void k_refit(BuildBuffers b) {
  const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= b.n) return;
  int node = b.parent_leaf[leaf];
  for (int guard = 0; node >= 0 && guard < b.n; ++guard) {
    __threadfence();
    const int prev = atomicAdd(&b.arrive[node], 1);
    if (prev == 0) return;  // the sibling subtree is not done yet
    __threadfence();
    float bl[6], br[6];
    child_box(b, b.child[2 * node], bl);
    child_box(b, b.child[2 * node + 1], br);
    for (int k = 0; k < 3; ++k) {
      __hip_atomic_store(&b.node_box[6 * (size_t)node + k], fminf(bl[k], br[k]), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&b.node_box[6 * (size_t)node + 3 + k], fmaxf(bl[3 + k], br[3 + k]),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    node = b.parent_internal[node];
  }
}

// Number of internal ancestors of each leaf; the maximum sizes the stack spill.
__global__ void k_depth(BuildBuffers b) {
  const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= b.n) return;
  int depth = 0;
  for (int node = b.parent_leaf[leaf]; node >= 0 && depth < b.n; node = b.parent_internal[node]) ++depth;
  atomicMax(b.max_depth, depth);
}

__global__ void k_emit(BuildBuffers b, BvhNode* nodes, float4* segs_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b.n - 1) {
    float bl[6], br[6];
    const int c0 = b.child[2 * i], c1 = b.child[2 * i + 1];
    child_box(b, c0, bl);
    child_box(b, c1, br);
    BvhNode nd;
    nd.a = make_float4(bl[0], bl[1], bl[2], bl[3]);
    nd.b = make_float4(bl[4], bl[5], br[0], br[1]);
    nd.c = make_float4(br[2], br[3], br[4], br[5]);
    nd.d = make_int4(c0, c1, 0, 0);
    nodes[i] = nd;
  }
  if (i < b.n) {
    const int orig = b.vals_sorted[i] / b.pieces;
    const float* s = b.seg_in + 8 * (size_t)orig;
    segs_out[2 * (size_t)i] = make_float4(s[0], s[1], s[2], s[3]);
    segs_out[2 * (size_t)i + 1] = make_float4(s[4], s[5], s[6], __int_as_float(orig));
  }
}

// Single-segment tree: one node whose two children are the same leaf.
__global__ void k_emit_single(const float* seg_in, BvhNode* nodes, float4* segs_out) {
  const float* s = seg_in;
  const float r = s[3];
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    const float pad = r + 1e-6f * (fabsf(s[k]) + fabsf(s[4 + k]) + r) + 1e-7f;
    lo[k] = fminf(s[k], s[4 + k]) - pad;
    hi[k] = fmaxf(s[k], s[4 + k]) + pad;
  }
  BvhNode nd;
  nd.a = make_float4(lo[0], lo[1], lo[2], hi[0]);
  nd.b = make_float4(hi[1], hi[2], 1.f, 1.f);
  nd.c = make_float4(1.f, -1.f, -1.f, -1.f);  // empty second box
  nd.d = make_int4(~0, ~0, 0, 0);
  nodes[0] = nd;
  segs_out[0] = make_float4(s[0], s[1], s[2], s[3]);
  segs_out[1] = make_float4(s[4], s[5], s[6], __int_as_float(0));
}

// ---- collapse to wide nodes ------------------------------------------------

struct CollapseArgs {
  const BvhNode* bin;
  BvhNodeWide* out;
  const int2* frontier;  // (binary inner node, wide-node index)
  int nfront;
  int2* next;
  int* next_count;
  int* node_count;
};

__device__ inline float half_area(float lx, float ly, float lz, float hx, float hy, float hz) {
  const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
  return dx * dy + dy * dz + dz * dx;
}

// ---- PLOC: parallel locally-ordered clustering (Meister and Bittner, IEEE
// TVCG 2018). Starting from the leaves in Morton order, every cluster finds
// the neighbour within +-radius positions whose merged box has the smallest
// surface area; mutual nearest neighbours merge into a new inner node, and
// the surviving clusters are compacted in order. The order (area, index) is
// total, so the pair that is smallest under it is always mutual and each
// round merges at least one pair. Inner nodes are numbered from n - 2 down,
// so the last merge, the root, is node 0. The tree replaces the Karras radix
// tree (child, node_box, parent links); emission and collapse are shared.

constexpr int kPlocBlock = 256;
constexpr int kPlocMaxRadius = 64;

__global__ void k_ploc_init(BuildBuffers b, int* cid, float* cbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  cid[i] = ~i;
  const int orig = b.vals_sorted[i];
  for (int k = 0; k < 6; ++k) cbox[6 * (size_t)i + k] = b.leaf_box[6 * (size_t)orig + k];
}

// Nearest neighbour of every cluster among positions [i - R, i + R], boxes
// staged in LDS per block with an R-wide halo on both sides.
__global__ void __launch_bounds__(kPlocBlock) k_ploc_nn(const float* cbox, int m, int R, int* nn) {
  __shared__ float tile[(kPlocBlock + 2 * kPlocMaxRadius) * 6];
  const int base = blockIdx.x * kPlocBlock - R;
  const int span = kPlocBlock + 2 * R;
  for (int t = threadIdx.x; t < span; t += blockDim.x) {
    const int g = base + t;
    for (int k = 0; k < 6; ++k) tile[6 * t + k] = (g >= 0 && g < m) ? cbox[6 * (size_t)g + k] : 0.f;
  }
  __syncthreads();
  const int i = blockIdx.x * kPlocBlock + threadIdx.x;
  if (i >= m) return;
  const float* bi = &tile[6 * (threadIdx.x + R)];
  float best = __builtin_inff();
  int bj = -1;
  const int j0 = max(0, i - R), j1 = min(m - 1, i + R);
  for (int j = j0; j <= j1; ++j) {
    if (j == i) continue;
    const float* bj_box = &tile[6 * (j - base)];
    const float a = half_area(fminf(bi[0], bj_box[0]), fminf(bi[1], bj_box[1]), fminf(bi[2], bj_box[2]),
                              fmaxf(bi[3], bj_box[3]), fmaxf(bi[4], bj_box[4]), fmaxf(bi[5], bj_box[5]));
    if (a < best || (a == best && j < bj)) {  // ties: the lower index (a total order)
      best = a;
      bj = j;
    }
  }
  nn[i] = bj;
}

__global__ void k_ploc_merge(BuildBuffers b, int m, const int* nn, int* cid, float* cbox, int* valid,
                             int* next_free) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int j = nn[i];
  const bool mutual = j >= 0 && nn[j] == i;
  if (mutual && i < j) {
    const int node = atomicSub(next_free, 1);
    if (node < 0) {  // more merges than inner nodes: impossible for a valid run; stay in bounds
      valid[i] = 1;
      return;
    }
    const int ci = cid[i], cj = cid[j];
    float u[6];
    for (int k = 0; k < 3; ++k) {
      u[k] = fminf(cbox[6 * (size_t)i + k], cbox[6 * (size_t)j + k]);
      u[3 + k] = fmaxf(cbox[6 * (size_t)i + 3 + k], cbox[6 * (size_t)j + 3 + k]);
    }
    b.child[2 * node] = ci;
    b.child[2 * node + 1] = cj;
    for (int k = 0; k < 6; ++k) b.node_box[6 * (size_t)node + k] = u[k];
    if (ci >= 0) b.parent_internal[ci] = node;
    else b.parent_leaf[~ci] = node;
    if (cj >= 0) b.parent_internal[cj] = node;
    else b.parent_leaf[~cj] = node;
    // the merged cluster takes position i (only this thread touches i and j)
    cid[i] = node;
    for (int k = 0; k < 6; ++k) cbox[6 * (size_t)i + k] = u[k];
    valid[i] = 1;
  } else {
    valid[i] = (mutual && i > j) ? 0 : 1;
  }
}

__global__ void k_ploc_compact(int m, const int* valid, const int* pos, const int* cid, const float* cbox, int* cid2,
                               float* cbox2, int* count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (valid[i]) {
    const int p = pos[i];
    cid2[p] = cid[i];
    for (int k = 0; k < 6; ++k) cbox2[6 * (size_t)p + k] = cbox[6 * (size_t)i + k];
  }
  if (i == m - 1) *count = pos[i] + valid[i];
}

__global__ void k_ploc_root(BuildBuffers b) { b.parent_internal[0] = -1; }

// Builds the binary tree of b (leaf boxes, Morton order in vals_sorted) by
// PLOC with search radius R.
static int ploc_build(BuildBuffers& b, DeviceArena& arena, int R, hipStream_t stream, std::string& err) {
  const int n = b.n;
  R = std::max(1, std::min(R, kPlocMaxRadius));
  int *cid[2], *nn, *valid, *pos, *dev_cnt, *next_free;
  float* cbox[2];
  for (int k = 0; k < 2; ++k) {
    HM_TRY(arena.alloc((void**)&cid[k], sizeof(int) * (size_t)n, err));
    HM_TRY(arena.alloc((void**)&cbox[k], sizeof(float) * 6 * (size_t)n, err));
  }
  HM_TRY(arena.alloc((void**)&nn, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&valid, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&pos, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&dev_cnt, sizeof(int) * 2, err));
  next_free = dev_cnt + 1;
  const int top = n - 2;
  HM_HIP(hipMemcpyAsync(next_free, &top, sizeof(int), hipMemcpyHostToDevice, stream), err);
  size_t scan_bytes = 0;
  HM_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, valid, pos, 0, (size_t)n, rocprim::plus<int>(), stream), err);
  void* scan_tmp = nullptr;
  HM_TRY(arena.alloc(&scan_tmp, scan_bytes, err));
  const int gs = (n + kPlocBlock - 1) / kPlocBlock;
  hipLaunchKernelGGL(k_ploc_init, dim3(gs), dim3(kPlocBlock), 0, stream, b, cid[0], cbox[0]);
  HM_TRY(check_launch("ploc init", err));
  int m = n, cur = 0;
  int* cnt_host = nullptr;
  HM_HIP(hipHostMalloc((void**)&cnt_host, sizeof(int)), err);
  int rc = kOk;
  for (int round = 0; m > 1; ++round) {
    if (round > n) {
      err = "PLOC did not converge";
      rc = kEDevice;
      break;
    }
    const int g = (m + kPlocBlock - 1) / kPlocBlock;
    hipLaunchKernelGGL(k_ploc_nn, dim3(g), dim3(kPlocBlock), 0, stream, cbox[cur], m, R, nn);
    hipLaunchKernelGGL(k_ploc_merge, dim3(g), dim3(kPlocBlock), 0, stream, b, m, nn, cid[cur], cbox[cur], valid,
                       next_free);
    if (hipSuccess != rocprim::exclusive_scan(scan_tmp, scan_bytes, valid, pos, 0, (size_t)m, rocprim::plus<int>(),
                                              stream)) {
      err = "PLOC scan failed";
      rc = kEDevice;
      break;
    }
    hipLaunchKernelGGL(k_ploc_compact, dim3(g), dim3(kPlocBlock), 0, stream, m, valid, pos, cid[cur], cbox[cur],
                       cid[cur ^ 1], cbox[cur ^ 1], dev_cnt);
    rc = check_launch("ploc round", err);
    if (rc) break;
    if (hipMemcpyAsync(cnt_host, dev_cnt, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      err = "PLOC count read-back failed";
      rc = kEDevice;
      break;
    }
    if (*cnt_host >= m) {
      err = "PLOC made no merge in a round";
      rc = kEDevice;
      break;
    }
    m = *cnt_host;
    cur ^= 1;
  }
  (void)hipHostFree(cnt_host);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ploc_root, dim3(1), dim3(1), 0, stream, b);
  return check_launch("ploc root", err);
}

// One wide node per frontier entry: start from the binary node's two
// children and open the inner child of largest surface area until there are
// kBvhWidth (or only leaves remain).
__global__ void k_collapse(CollapseArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nfront) return;
  const int2 f = a.frontier[i];
  float lo[kBvhWidth][3], hi[kBvhWidth][3];
  int ch[kBvhWidth];
  auto load2 = [&](int b, int s0, int s1) {
    const BvhNode n = a.bin[b];
    lo[s0][0] = n.a.x; lo[s0][1] = n.a.y; lo[s0][2] = n.a.z;
    hi[s0][0] = n.a.w; hi[s0][1] = n.b.x; hi[s0][2] = n.b.y;
    lo[s1][0] = n.b.z; lo[s1][1] = n.b.w; lo[s1][2] = n.c.x;
    hi[s1][0] = n.c.y; hi[s1][1] = n.c.z; hi[s1][2] = n.c.w;
    ch[s0] = n.d.x;
    ch[s1] = n.d.y;
  };
  load2(f.x, 0, 1);
  int cnt = 2;
  while (cnt < kBvhWidth) {
    int best = -1;
    float ba = -1.f;
    for (int k = 0; k < cnt; ++k) {
      if (ch[k] < 0) continue;
      const float ar = half_area(lo[k][0], lo[k][1], lo[k][2], hi[k][0], hi[k][1], hi[k][2]);
      if (ar > ba) {
        ba = ar;
        best = k;
      }
    }
    if (best < 0) break;
    load2(ch[best], best, cnt);
    ++cnt;
  }
  BvhNodeWide o;
  for (int k = 0; k < kBvhWidth; ++k) {
    if (k < cnt) {
      o.lox[k] = lo[k][0]; o.loy[k] = lo[k][1]; o.loz[k] = lo[k][2];
      o.hix[k] = hi[k][0]; o.hiy[k] = hi[k][1]; o.hiz[k] = hi[k][2];
      if (ch[k] >= 0) {
        const int j = atomicAdd(a.node_count, 1);
        a.next[atomicAdd(a.next_count, 1)] = make_int2(ch[k], j);
        o.child[k] = j;
      } else {
        o.child[k] = ch[k];
      }
    } else {
      o.lox[k] = o.loy[k] = o.loz[k] = 0.f;
      o.hix[k] = o.hiy[k] = o.hiz[k] = 0.f;
      o.child[k] = kEmptyChild;
    }
  }
  for (int k = 0; k < (int)(sizeof(o.pad) / sizeof(int)); ++k) o.pad[k] = 0;
  a.out[f.y] = o;
}

}  // namespace

int DeviceBvh::stack_bound() const { return (kBvhWidth - 1) * depth_w + 1; }

// Collapses the binary tree into wide nodes, one frontier level per
// launch (the node count per level comes back to the host, so the loop
// knows when the tree is done).
static int bvh_collapse(DeviceBvh& out, hipStream_t stream, std::string& err) {
  const int cap = std::max(1, out.num_nodes);  // a wide tree has fewer nodes than the binary one
  HM_TRY(dev_alloc((void**)&out.nodes_w, sizeof(BvhNodeWide) * (size_t)cap, err));
  DeviceArena arena;
  int2 *fa = nullptr, *fb = nullptr;
  int* counters = nullptr;
  HM_TRY(arena.alloc((void**)&fa, sizeof(int2) * (size_t)cap, err));
  HM_TRY(arena.alloc((void**)&fb, sizeof(int2) * (size_t)cap, err));
  HM_TRY(arena.alloc((void**)&counters, sizeof(int) * 2, err));
  const int2 root = make_int2(0, 0);
  const int init[2] = {1, 0};  // node count (root allocated), next-frontier count
  HM_HIP(hipMemcpyAsync(fa, &root, sizeof(root), hipMemcpyHostToDevice, stream), err);
  HM_HIP(hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, stream), err);
  int nfront = 1, levels = 0;
  while (nfront > 0) {
    HM_HIP(hipMemsetAsync(counters + 1, 0, sizeof(int), stream), err);
    CollapseArgs a{out.nodes, out.nodes_w, fa, nfront, fb, counters + 1, counters};
    hipLaunchKernelGGL(k_collapse, dim3((nfront + 255) / 256), dim3(256), 0, stream, a);
    HM_TRY(check_launch("bvh collapse", err));
    int h[2];
    HM_HIP(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, stream), err);
    HM_HIP(hipStreamSynchronize(stream), err);
    if (h[0] > cap) {
      err = "bvh collapse: node count exceeds the binary tree's";
      return kEDevice;
    }
    out.num_nodes_w = h[0];
    nfront = h[1];
    std::swap(fa, fb);
    ++levels;
  }
  out.depth_w = levels;
  return kOk;
}

// Builds the BVH from device segments (8 floats each). All scratch is
// released before returning; the result owns nodes and segment records.
int bvh_build(const float* seg_dev, int n, DeviceBvh& out, hipStream_t stream, std::string& err, int builder,
              int ploc_radius, int max_leaf, int split) {
  out.release();
  out.num_prims = std::max(n, 0);
  if (split < 1 || split > 4) {
    err = "BVH: split must be 1..4";
    return kEInval;
  }
  const int pieces = (builder == kBvhSah && n > 1) ? split : 1;
  if ((long long)n * pieces > kLeafPosMask + 1) {
    err = "BVH: at most 16,777,216 leaf boxes (primitives x accel.split)";
    return kEInval;
  }
  n *= pieces;  // from here on: leaves
  out.num_segs = n;
  if (n <= 0) {
    out.max_depth = 0;
    return 0;
  }
  const int num_nodes = n > 1 ? n - 1 : 1;
  HM_TRY(dev_alloc((void**)&out.nodes, sizeof(BvhNode) * (size_t)num_nodes, err));
  HM_TRY(dev_alloc((void**)&out.segs, sizeof(float4) * 2 * (size_t)n, err));
  out.num_nodes = num_nodes;
  if (n == 1) {
    hipLaunchKernelGGL(k_emit_single, dim3(1), dim3(1), 0, stream, seg_dev, out.nodes, out.segs);
    out.max_depth = 1;
    HM_TRY(check_launch("bvh single", err));
    return bvh_collapse(out, stream, err);
  }
  DeviceArena arena;
  BuildBuffers b{};
  b.seg_in = seg_dev;
  b.n = n;
  b.pieces = pieces;
  HM_TRY(arena.alloc((void**)&b.leaf_box, sizeof(float) * 6 * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.cbounds, sizeof(int) * 8, err));
  HM_TRY(arena.alloc((void**)&b.keys, sizeof(unsigned) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.vals, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.keys_sorted, sizeof(unsigned) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.vals_sorted, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.child, sizeof(int) * 2 * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.parent_internal, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.parent_leaf, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.node_box, sizeof(float) * 6 * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.arrive, sizeof(int) * (size_t)n, err));
  HM_TRY(arena.alloc((void**)&b.max_depth, sizeof(int) * 4, err));
  const int init_b[8] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN, 0, 0};
  HM_HIP(hipMemcpyAsync(b.cbounds, init_b, sizeof(init_b), hipMemcpyHostToDevice, stream), err);
  HM_HIP(hipMemsetAsync(b.arrive, 0, sizeof(int) * (size_t)n, stream), err);
  HM_HIP(hipMemsetAsync(b.max_depth, 0, sizeof(int) * 4, stream), err);
  const int bs = 256, gs = (n + bs - 1) / bs;
  hipLaunchKernelGGL(k_leaf_boxes, dim3(gs), dim3(bs), 0, stream, b);
  HM_TRY(check_launch("bvh leaf boxes", err));
  if (builder == kBvhSah) {
    // the tree on the host from the padded leaf boxes, then the same refit,
    // emission and collapse as the radix tree
    std::vector<float> boxes((size_t)6 * n);
    std::vector<int> child((size_t)2 * (n - 1)), order(n), pin(n - 1), pleaf(n);
    HM_HIP(hipMemcpyAsync(boxes.data(), b.leaf_box, sizeof(float) * boxes.size(), hipMemcpyDeviceToHost, stream),
           err);
    HM_HIP(hipStreamSynchronize(stream), err);
    sah_build(boxes.data(), n, child.data(), order.data(), pin.data(), pleaf.data(), max_leaf);
    HM_HIP(hipMemcpyAsync(b.child, child.data(), sizeof(int) * child.size(), hipMemcpyHostToDevice, stream), err);
    HM_HIP(hipMemcpyAsync(b.vals_sorted, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, stream),
           err);
    HM_HIP(hipMemcpyAsync(b.parent_internal, pin.data(), sizeof(int) * pin.size(), hipMemcpyHostToDevice, stream),
           err);
    HM_HIP(hipMemcpyAsync(b.parent_leaf, pleaf.data(), sizeof(int) * pleaf.size(), hipMemcpyHostToDevice, stream),
           err);
    hipLaunchKernelGGL(k_refit, dim3(gs), dim3(bs), 0, stream, b);
    HM_TRY(check_launch("bvh refit", err));
    hipLaunchKernelGGL(k_emit, dim3(gs), dim3(bs), 0, stream, b, out.nodes, out.segs);
    HM_TRY(check_launch("bvh emit", err));
    hipLaunchKernelGGL(k_depth, dim3(gs), dim3(bs), 0, stream, b);
    HM_TRY(check_launch("bvh depth", err));
    int depth = 0;
    HM_HIP(hipMemcpyAsync(&depth, b.max_depth, sizeof(int), hipMemcpyDeviceToHost, stream), err);
    HM_HIP(hipStreamSynchronize(stream), err);
    out.max_depth = depth;
    return bvh_collapse(out, stream, err);
  }
  hipLaunchKernelGGL(k_morton, dim3(gs), dim3(bs), 0, stream, b);
  HM_TRY(check_launch("bvh morton", err));
  size_t tmp_bytes = 0;
  HM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, b.keys, b.keys_sorted, b.vals, b.vals_sorted, n, 0,
                                   30, stream),
         err);
  void* tmp = nullptr;
  HM_TRY(arena.alloc(&tmp, tmp_bytes, err));
  HM_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, b.keys, b.keys_sorted, b.vals, b.vals_sorted, n, 0, 30,
                                   stream),
         err);
  if (builder == kBvhPloc) {
    HM_TRY(ploc_build(b, arena, ploc_radius, stream, err));
  } else {
    hipLaunchKernelGGL(k_karras, dim3(gs), dim3(bs), 0, stream, b);
    HM_TRY(check_launch("bvh karras", err));
    hipLaunchKernelGGL(k_refit, dim3(gs), dim3(bs), 0, stream, b);
    HM_TRY(check_launch("bvh refit", err));
  }
  hipLaunchKernelGGL(k_emit, dim3(gs), dim3(bs), 0, stream, b, out.nodes, out.segs);
  HM_TRY(check_launch("bvh emit", err));
  hipLaunchKernelGGL(k_depth, dim3(gs), dim3(bs), 0, stream, b);
  HM_TRY(check_launch("bvh depth", err));
  int depth = 0;
  HM_HIP(hipMemcpyAsync(&depth, b.max_depth, sizeof(int), hipMemcpyDeviceToHost, stream), err);
  HM_HIP(hipStreamSynchronize(stream), err);
  out.max_depth = depth;
  return bvh_collapse(out, stream, err);
}

}  // namespace hm
