// femrt_lbvh.hip -- BVH builder 1: a linear BVH built on the GPU.
//
// The alternative to the host binned-SAH build (femrt_host.cpp, builder 0)
// for meshes whose host build time matters.  Every stage runs on the
// context's stream:
//   1. k_lbvh_bounds    bounds of the element-box centroids (block reduction,
//                       one atomic per block on order-preserving integer keys);
//   2. k_lbvh_morton    30-bit Morton code of each centroid (10 bits an axis);
//   3. radix sort of (code, element) pairs (rocPRIM);
//   4. k_lbvh_tree      the binary radix tree over the sorted codes, one
//                       thread per internal node: direction and extent of the
//                       node's key range by longest common prefix, split by
//                       binary search (Karras, "Maximizing parallelism in the
//                       construction of BVHs, octrees, and k-d trees", HPG
//                       2012); equal codes are told apart by their sorted
//                       index, so the tree is well defined for any input;
//   5. k_lbvh_boxes     boxes bottom-up, one thread per leaf: the second of a
//                       node's two children to finish computes its box and
//                       climbs on (one atomic counter per node, no waiting);
//   6. k_lbvh_collapse  internal nodes whose range holds at most 4 elements
//                       become leaves, as the SAH build's leaves hold <= 4.
// The node layout is the traversal's (two float4 per node: lo.xyz + a,
// hi.xyz + b; internal a, b = children; leaf b = -count, a = first slot of
// the element list), root at node 0.  Internal nodes occupy [0, n-1), the
// single-element leaves [n-1, 2n-1); nodes under a collapsed leaf stay in the
// array unreferenced.
// Reference: BVH.cu:9-247 (the host SAH build this replaces; the reference
// has no GPU builder).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "femrt_internal.hpp"
#include "femrt_runtime.hpp"

namespace {

constexpr int LBVH_LEAF_MAX = 4;
constexpr int LBVH_BLOCK = 256;

// float <-> unsigned key whose unsigned order is the float order
__device__ __forceinline__ unsigned flt_key(float f)
{
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_flt(unsigned k)
{
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ void k_lbvh_init(unsigned* __restrict__ bounds)
{
  if (threadIdx.x < 3) bounds[threadIdx.x] = 0xFFFFFFFFu;     // min keys
  else if (threadIdx.x < 6) bounds[threadIdx.x] = 0u;         // max keys
}

__global__ __launch_bounds__(LBVH_BLOCK) void k_lbvh_bounds(const float* __restrict__ boxes, int n,
                                                            unsigned* __restrict__ bounds)
{
  __shared__ unsigned s[6][LBVH_BLOCK];
  unsigned lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const float* b = boxes + (size_t)e * FEMRT_BOX_STRIDE;
    for (int c = 0; c < 3; ++c) {
      const unsigned k = flt_key(0.5f * (b[c] + b[3 + c]));
      lo[c] = min(lo[c], k);
      hi[c] = max(hi[c], k);
    }
  }
  for (int c = 0; c < 3; ++c) {
    s[c][threadIdx.x] = lo[c];
    s[3 + c][threadIdx.x] = hi[c];
  }
  __syncthreads();
  for (int w = LBVH_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = min(s[c][threadIdx.x], s[c][threadIdx.x + w]);
        s[3 + c][threadIdx.x] = max(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicMin(&bounds[threadIdx.x], s[threadIdx.x][0]);
  else if (threadIdx.x < 6) atomicMax(&bounds[threadIdx.x], s[threadIdx.x][0]);
}

// 10 bits of v spread to every third bit
__device__ __forceinline__ unsigned spread3(unsigned v)
{
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(LBVH_BLOCK) void k_lbvh_morton(const float* __restrict__ boxes, int n,
                                                            const unsigned* __restrict__ bounds,
                                                            unsigned* __restrict__ codes, int* __restrict__ ids)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float* b = boxes + (size_t)e * FEMRT_BOX_STRIDE;
  unsigned q[3];
  for (int c = 0; c < 3; ++c) {
    const float lo = key_flt(bounds[c]), hi = key_flt(bounds[3 + c]);
    const float ext = hi - lo;
    const float x = ext > 0.f ? (0.5f * (b[c] + b[3 + c]) - lo) / ext : 0.f;
    q[c] = (unsigned)min(1023.f, fmaxf(0.f, x * 1024.f));
  }
  codes[e] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
  ids[e] = e;
}

// Common-prefix length of sorted keys i and j; equal keys by their indices.
__device__ __forceinline__ int lcp(const unsigned* __restrict__ codes, int n, int i, int j)
{
  if (j < 0 || j >= n) return -1;
  const unsigned a = codes[i], b = codes[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((unsigned)i ^ (unsigned)j);
}

__global__ __launch_bounds__(LBVH_BLOCK) void k_lbvh_tree(const unsigned* __restrict__ codes, int n,
                                                          int* __restrict__ child, int* __restrict__ parent,
                                                          int* __restrict__ range)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  // direction of the range: towards the neighbour with the longer prefix
  const int d = lcp(codes, n, i, i + 1) - lcp(codes, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = lcp(codes, n, i, i - d);
  int lmax = 2;
  while (lcp(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (lcp(codes, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = lcp(codes, n, i, j);
  // split: the last index whose prefix with i is longer than the node's
  int s = 0, step = l;
  do {
    step = (step + 1) >> 1;
    if (s + step < l && lcp(codes, n, i, i + (s + step) * d) > dnode) s += step;
  } while (step > 1);
  const int gamma = i + s * d + min(d, 0);
  const int first = min(i, j), last = max(i, j);
  const int left = first == gamma ? (n - 1) + gamma : gamma;
  const int right = last == gamma + 1 ? (n - 1) + gamma + 1 : gamma + 1;
  child[2 * i] = left;
  child[2 * i + 1] = right;
  parent[left] = i;
  parent[right] = i;
  range[2 * i] = first;
  range[2 * i + 1] = last;
  if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ void store_box(float4* nodes, int node, const float* lo, const float* hi, int a, int b)
{
  float* p = (float*)(nodes + 2 * (size_t)node);
  for (int c = 0; c < 3; ++c) {
    __hip_atomic_store(p + c, lo[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 4 + c, hi[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __hip_atomic_store((int*)p + 3, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store((int*)p + 7, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void load_box(const float4* nodes, int node, float* lo, float* hi)
{
  const float* p = (const float*)(nodes + 2 * (size_t)node);
  for (int c = 0; c < 3; ++c) {
    lo[c] = __hip_atomic_load(p + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    hi[c] = __hip_atomic_load(p + 4 + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Leaf i (sorted slot i) stores its element's box, then climbs: at each
// parent the first child to arrive stops, the second (which sees the first's
// box through the acquire-release counter) writes the union and goes on.
// Every thread stops after at most the tree depth steps.
__global__ __launch_bounds__(LBVH_BLOCK) void k_lbvh_boxes(const float* __restrict__ boxes, int n,
                                                           const int* __restrict__ ids,
                                                           const int* __restrict__ child,
                                                           const int* __restrict__ parent,
                                                           int* __restrict__ arrived, float4* __restrict__ nodes)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* b = boxes + (size_t)ids[i] * FEMRT_BOX_STRIDE;
  float lo[3] = {b[0], b[1], b[2]}, hi[3] = {b[3], b[4], b[5]};
  const int leaf = (n - 1) + i;
  store_box(nodes, leaf, lo, hi, i, -1);
  int node = parent[leaf];
  while (node >= 0) {
    if (__hip_atomic_fetch_add(&arrived[node], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    const int a = child[2 * node], c = child[2 * node + 1];
    float lo2[3], hi2[3];
    load_box(nodes, a, lo, hi);
    load_box(nodes, c, lo2, hi2);
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(lo[k], lo2[k]);
      hi[k] = fmaxf(hi[k], hi2[k]);
    }
    store_box(nodes, node, lo, hi, a, c);
    node = parent[node];
  }
}

__global__ __launch_bounds__(LBVH_BLOCK) void k_lbvh_collapse(int n, const int* __restrict__ range,
                                                              float4* __restrict__ nodes)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int first = range[2 * i], cnt = range[2 * i + 1] - first + 1;
  if (cnt <= LBVH_LEAF_MAX) {
    int* p = (int*)(nodes + 2 * (size_t)i);
    p[3] = first;
    p[7] = -cnt;
  }
}

}  // namespace

// Builds the mesh's BVH on the device (builder 1) into m->d_nodes / d_prims
// and keeps a host copy in m->bvh.  Returns FEMRT_OK or an error code (the
// context's last error says why).
int lbvh_build(femrt_ctx* c, femrt_mesh* m)
{
  const int n = m->n_elem;
  if (n < 1) return fail(c, FEMRT_EINVAL, "build_bvh: no elements");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n_nodes = n > 1 ? 2 * (size_t)n - 1 : 1;
  float4* d_nodes = nullptr;
  int* d_prims = nullptr;
  unsigned *d_bounds = nullptr, *d_codes = nullptr, *d_codes2 = nullptr;
  int *d_ids = nullptr, *d_child = nullptr, *d_parent = nullptr, *d_range = nullptr, *d_arrived = nullptr;
  void* d_tmp = nullptr;
  size_t tmp_bytes = 0;
  auto cleanup = [&]() {
    // best-effort frees of the build's temporaries
    (void)hipFree(d_bounds);
    (void)hipFree(d_codes);
    (void)hipFree(d_codes2);
    (void)hipFree(d_ids);
    (void)hipFree(d_child);
    (void)hipFree(d_parent);
    (void)hipFree(d_range);
    (void)hipFree(d_arrived);
    (void)hipFree(d_tmp);
  };
  auto bad = [&](hipError_t e, const char* what) {
    cleanup();
    (void)hipFree(d_nodes);
    (void)hipFree(d_prims);
    return fail(c, FEMRT_EHIP, std::string("build_bvh (GPU LBVH): ") + what + ": " + hipGetErrorString(e));
  };
  hipError_t e;
#define LB(call, what)                        \
  if ((e = (call)) != hipSuccess) return bad(e, what);
  LB(hipMalloc(&d_nodes, n_nodes * 2 * sizeof(float4)), "hipMalloc nodes");
  LB(hipMalloc(&d_prims, (size_t)n * 4), "hipMalloc prims");
  LB(hipMalloc(&d_bounds, 6 * 4), "hipMalloc bounds");
  LB(hipMalloc(&d_codes, (size_t)n * 4), "hipMalloc codes");
  LB(hipMalloc(&d_codes2, (size_t)n * 4), "hipMalloc codes");
  LB(hipMalloc(&d_ids, (size_t)n * 4), "hipMalloc ids");
  const int nb = (n + LBVH_BLOCK - 1) / LBVH_BLOCK;
  const int nbr = std::min(nb, 4 * c->cus);
  hipLaunchKernelGGL(k_lbvh_init, dim3(1), dim3(64), 0, c->stream, d_bounds);
  hipLaunchKernelGGL(k_lbvh_bounds, dim3(nbr), dim3(LBVH_BLOCK), 0, c->stream, m->d_boxes, n, d_bounds);
  hipLaunchKernelGGL(k_lbvh_morton, dim3(nb), dim3(LBVH_BLOCK), 0, c->stream, m->d_boxes, n, d_bounds,
                     d_codes, d_ids);
  LB(hipGetLastError(), "Morton codes");
  LB(rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_codes, d_codes2, d_ids, d_prims, n, 0, 30, c->stream),
     "radix sort size");
  LB(hipMalloc(&d_tmp, std::max<size_t>(tmp_bytes, 4)), "hipMalloc sort storage");
  LB(rocprim::radix_sort_pairs(d_tmp, tmp_bytes, d_codes, d_codes2, d_ids, d_prims, n, 0, 30, c->stream),
     "radix sort");
  if (n > 1) {
    LB(hipMalloc(&d_child, (size_t)(n - 1) * 2 * 4), "hipMalloc children");
    LB(hipMalloc(&d_parent, n_nodes * 4), "hipMalloc parents");
    LB(hipMalloc(&d_range, (size_t)(n - 1) * 2 * 4), "hipMalloc ranges");
    LB(hipMalloc(&d_arrived, (size_t)(n - 1) * 4), "hipMalloc counters");
    LB(hipMemsetAsync(d_arrived, 0, (size_t)(n - 1) * 4, c->stream), "clear counters");
    const int ni = (n - 1 + LBVH_BLOCK - 1) / LBVH_BLOCK;
    hipLaunchKernelGGL(k_lbvh_tree, dim3(ni), dim3(LBVH_BLOCK), 0, c->stream, d_codes2, n, d_child, d_parent,
                       d_range);
    hipLaunchKernelGGL(k_lbvh_boxes, dim3(nb), dim3(LBVH_BLOCK), 0, c->stream, m->d_boxes, n, d_prims, d_child,
                       d_parent, d_arrived, d_nodes);
    hipLaunchKernelGGL(k_lbvh_collapse, dim3(ni), dim3(LBVH_BLOCK), 0, c->stream, n, d_range, d_nodes);
    LB(hipGetLastError(), "tree kernels");
  } else {
    // one element: the root is its leaf
    const float* b = m->h_boxes.data();
    float node[8] = {b[0], b[1], b[2], 0.f, b[3], b[4], b[5], 0.f};
    const int a = 0, bb = -1;
    std::memcpy(&node[3], &a, 4);
    std::memcpy(&node[7], &bb, 4);
    LB(hipMemcpyAsync(d_nodes, node, sizeof(node), hipMemcpyHostToDevice, c->stream), "root leaf");
  }
  // host copy of the tree (depth check, femrt_mesh_get_bvh)
  m->bvh.nodes.resize(n_nodes * 8);
  m->bvh.prims.resize(n);
  LB(hipMemcpyAsync(m->bvh.nodes.data(), d_nodes, n_nodes * 32, hipMemcpyDeviceToHost, c->stream), "copy nodes");
  LB(hipMemcpyAsync(m->bvh.prims.data(), d_prims, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream), "copy prims");
  LB(hipStreamSynchronize(c->stream), "build");
#undef LB
  cleanup();
  (void)hipFree(m->d_nodes);
  (void)hipFree(m->d_prims);
  m->d_nodes = d_nodes;
  m->d_prims = d_prims;
  m->n_nodes = (int)n_nodes;
  return FEMRT_OK;
}
