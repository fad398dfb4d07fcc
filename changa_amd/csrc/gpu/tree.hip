// GPU tree build: Morton order, binary splits and breadth-first numbering.
//
// The tree is the one the host builder makes (the caller side, playing
// TreePiece::buildTree, TreePiece.cpp:2619, and LocalTreeBuilder,
// Compute.cpp:2440-2580): particles ordered along a Morton curve whose top
// key bit splits x, then y, then z (the axis cycle of
// BinaryTreeNode::makeOctChildren, GenericTreeNode.h:473-575); a node holding
// more than nBucket particles (Compute.cpp:2476) splits on the highest key
// bit on which its particles differ, so chains of single-child nodes never
// appear; coincident keys leave an oversized bucket. Nodes are numbered
// breadth first (children after parents, each depth contiguous, children in
// their parents' order) and buckets in particle order, so every array equals
// the host builder's.
//
// Steps: 63-bit Morton keys (21 bits per axis, the host's FP64 arithmetic),
// a stable rocPRIM radix sort of (key, input index), then one pass per depth:
// each node finds its split bit from the xor of its first and last keys and
// its split point by binary search; a scan over the depth's internal nodes
// numbers the next depth. The host reads one count per depth.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <vector>

#include "common.h"

namespace cgpu {

constexpr int kMortonBits = 21;

__device__ __forceinline__ unsigned long long spread3(unsigned long long v) {
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

#pragma clang fp contract(off)

__global__ void morton_kernel(int n, const double* pos, double lx, double ly, double lz,
                              double scale, unsigned long long* key, int* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lx, ly, lz};
  unsigned long long c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double u = (pos[3 * (size_t)i + k] - lo[k]) * scale;
    long long v = (long long)floor(u);
    if (v < 0) v = 0;
    if (v > (1 << kMortonBits) - 1) v = (1 << kMortonBits) - 1;
    c[k] = (unsigned long long)v;
  }
  key[i] = (spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]);
  idx[i] = i;
}

// Nodes [lo, hi) of one depth: box radius from the creation bit; split flag
// (1 for a node that splits), split point and split bit.
__global__ void split_kernel(int lo, int hi, int bucket, double box_size,
                             const unsigned long long* key, const int* first, const int* count,
                             const int* nbit, double* box_radius, int* flag, int* split,
                             int* sbit) {
  const int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= hi) return;
  const int level = 62 - nbit[nd];
  const int nx = (level + 2) / 3, ny = (level + 1) / 3, nz = level / 3;
  const double ex = box_size / ldexp(1.0, nx), ey = box_size / ldexp(1.0, ny),
               ez = box_size / ldexp(1.0, nz);
  box_radius[nd] = 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
  int f = 0;
  const int cnt = count[nd];
  if (cnt > bucket) {
    const int a = first[nd], b = a + cnt - 1;
    const unsigned long long x = key[a] ^ key[b];
    if (x != 0ull) {
      // keys of the node share every bit above the highest differing one
      const int bit = 63 - __clzll((long long)x);
      const unsigned long long mask = 1ull << bit;
      int l = a, r = b;  // first key in [a, b] with the bit set (key[b] has it)
      while (l < r) {
        const int m = l + ((r - l) >> 1);
        if (key[m] & mask)
          r = m;
        else
          l = m + 1;
      }
      split[nd - lo] = l;
      sbit[nd - lo] = bit;
      f = 1;
    }
  }
  flag[nd - lo] = f;
}

#pragma clang fp contract(on)

// children of the splitting nodes of [lo, hi), numbered from `base` in
// their parents' order
__global__ void emit_kernel(int lo, int hi, int base, const int* flag, const int* rank,
                            const int* split, const int* sbit, int* first, int* count, int* child,
                            int* nbit) {
  const int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= hi) return;
  const int j = nd - lo;
  if (!flag[j]) {
    child[2 * nd] = -1;
    child[2 * nd + 1] = -1;
    return;
  }
  const int c0 = base + 2 * rank[j], c1 = c0 + 1;
  child[2 * nd] = c0;
  child[2 * nd + 1] = c1;
  const int a = first[nd], s = split[j], e = a + count[nd];
  first[c0] = a;
  count[c0] = s - a;
  first[c1] = s;
  count[c1] = e - s;
  nbit[c0] = sbit[j] - 1;
  nbit[c1] = sbit[j] - 1;
}

__global__ void root_kernel(int n, int* first, int* count, int* nbit) {
  first[0] = 0;
  count[0] = n;
  nbit[0] = 62;
}

// buckets in particle order: mark each bucket's first particle, scan, place
__global__ void bucket_mark_kernel(int n_nodes, const int* first, const int* child, int* mark,
                                   int* owner) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= n_nodes || child[2 * nd] >= 0) return;
  mark[first[nd]] = 1;
  owner[first[nd]] = nd;
}

__global__ void bucket_place_kernel(int n, const int* mark, const int* rank, const int* owner,
                                    int* bucket_node) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n && mark[p]) bucket_node[rank[p]] = owner[p];
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

struct TreeWork {
  size_t key, key_out, idx, flag, rank, split, sbit, nbit, temp, total, temp_bytes;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static bool tree_layout(int n, TreeWork& w) {
  size_t sort_b = 0, scan_b = 0;
  if (rocprim::radix_sort_pairs(nullptr, sort_b, (const unsigned long long*)nullptr,
                                (unsigned long long*)nullptr, (const int*)nullptr, (int*)nullptr,
                                (size_t)n, 0, 3 * kMortonBits) != hipSuccess)
    return false;
  if (rocprim::exclusive_scan(nullptr, scan_b, (const int*)nullptr, (int*)nullptr, 0,
                              (size_t)n + 1, rocprim::plus<int>()) != hipSuccess)
    return false;
  w.temp_bytes = sort_b > scan_b ? sort_b : scan_b;
  const size_t nn = 2 * (size_t)n + 1;  // nodes of a binary tree over n particles: < 2n
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + bytes);
    return at;
  };
  w.key = take(8 * (size_t)n);
  w.key_out = take(8 * (size_t)n);
  w.idx = take(4 * (size_t)n);
  w.flag = take(4 * ((size_t)n + 1));
  w.rank = take(4 * ((size_t)n + 1));
  w.split = take(4 * (size_t)n);
  w.sbit = take(4 * (size_t)n);
  w.nbit = take(4 * nn);
  w.temp = take(w.temp_bytes);
  w.total = o;
  return true;
}

}  // namespace cgpu

using namespace cgpu;

extern "C" size_t changaBuildTreeWorkBytes(int n) {
  TreeWork w;
  if (n <= 0 || !tree_layout(n, w)) return 0;
  return w.total;
}

extern "C" int changaBuildTree(ChangaTreeBuild* t, void* stream) {
  if (!t || t->n <= 0 || !t->pos || !t->perm || !t->first || !t->count || !t->child ||
      !t->boxRadius || !t->bucketNode || !t->levelOffsets || !t->work || t->bucketSize < 1 ||
      !(t->boxSize > 0.0))
    return fail("changaBuildTree: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  const int n = t->n;
  if (t->maxNodes < 2 * n - 1)
    return fail("changaBuildTree: node arrays need 2n - 1 entries", hipErrorInvalidValue,
                __FILE__, __LINE__);
  TreeWork W;
  if (!tree_layout(n, W) || t->workBytes < W.total)
    return fail("changaBuildTree: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  hipStream_t s = (hipStream_t)stream;
  char* wb = static_cast<char*>(t->work);
  auto* key = reinterpret_cast<unsigned long long*>(wb + W.key);
  auto* key_out = reinterpret_cast<unsigned long long*>(wb + W.key_out);
  int* idx = reinterpret_cast<int*>(wb + W.idx);
  int* flag = reinterpret_cast<int*>(wb + W.flag);
  int* rank = reinterpret_cast<int*>(wb + W.rank);
  int* split = reinterpret_cast<int*>(wb + W.split);
  int* sbit = reinterpret_cast<int*>(wb + W.sbit);
  int* nbit = reinterpret_cast<int*>(wb + W.nbit);
  void* temp = wb + W.temp;
  const int B = 256;
  // keys and the stable sort (ties keep input order)
  const double scale = double(1 << kMortonBits) / t->boxSize;
  hipLaunchKernelGGL(morton_kernel, dim3((n + B - 1) / B), dim3(B), 0, s, n, t->pos,
                     t->boxLo[0], t->boxLo[1], t->boxLo[2], scale, key, idx);
  CGPU_CHECK(hipGetLastError());
  size_t tb = W.temp_bytes;
  CGPU_CHECK(rocprim::radix_sort_pairs(temp, tb, key, key_out, idx, t->perm, (size_t)n, 0,
                                       3 * kMortonBits, s));
  // depths, top down
  hipLaunchKernelGGL(root_kernel, dim3(1), dim3(1), 0, s, n, t->first, t->count, nbit);
  CGPU_CHECK(hipGetLastError());
  std::vector<int> offs;
  int lo = 0, hi = 1;
  while (hi > lo) {
    if ((int)offs.size() + 2 > t->maxDepths)
      return fail("changaBuildTree: deeper than maxDepths", hipErrorInvalidValue, __FILE__,
                  __LINE__);
    offs.push_back(lo);
    const int m = hi - lo;
    hipLaunchKernelGGL(split_kernel, dim3((m + B - 1) / B), dim3(B), 0, s, lo, hi, t->bucketSize,
                       t->boxSize, key_out, t->first, t->count, nbit, t->boxRadius, flag, split,
                       sbit);
    CGPU_CHECK(hipGetLastError());
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, s, flag + m, 1, 0);
    CGPU_CHECK(hipGetLastError());
    tb = W.temp_bytes;
    CGPU_CHECK(rocprim::exclusive_scan(temp, tb, flag, rank, 0, (size_t)m + 1,
                                       rocprim::plus<int>(), s));
    hipLaunchKernelGGL(emit_kernel, dim3((m + B - 1) / B), dim3(B), 0, s, lo, hi, hi, flag, rank,
                       split, sbit, t->first, t->count, t->child, nbit);
    CGPU_CHECK(hipGetLastError());
    int internal = 0;
    CGPU_CHECK(hipMemcpyAsync(&internal, rank + m, sizeof(int), hipMemcpyDeviceToHost, s));
    CGPU_CHECK(hipStreamSynchronize(s));
    lo = hi;
    hi = hi + 2 * internal;
  }
  offs.push_back(hi);
  const int n_nodes = hi;
  // buckets in particle order (flag / rank / split reused as mark / rank / owner)
  hipLaunchKernelGGL(fill_int_kernel, dim3((n + 1 + B - 1) / B), dim3(B), 0, s, flag, n + 1, 0);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(bucket_mark_kernel, dim3((n_nodes + B - 1) / B), dim3(B), 0, s, n_nodes,
                     t->first, t->child, flag, split);
  CGPU_CHECK(hipGetLastError());
  tb = W.temp_bytes;
  CGPU_CHECK(rocprim::exclusive_scan(temp, tb, flag, rank, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), s));
  hipLaunchKernelGGL(bucket_place_kernel, dim3((n + B - 1) / B), dim3(B), 0, s, n, flag, rank,
                     split, t->bucketNode);
  CGPU_CHECK(hipGetLastError());
  int nb = 0;
  CGPU_CHECK(hipMemcpyAsync(&nb, rank + n, sizeof(int), hipMemcpyDeviceToHost, s));
  CGPU_CHECK(hipStreamSynchronize(s));
  t->nNodes = n_nodes;
  t->nBuckets = nb;
  t->nDepths = (int)offs.size() - 1;
  for (size_t k = 0; k < offs.size(); ++k) t->levelOffsets[k] = offs[k];
  return 0;
}
