// Host-side tree build and node geometry for the gravity hot path.
//
// This is the caller side of the GPU boundary: it plays the part of ChaNGa's
// TreePiece tree build (TreePiece::buildTree, TreePiece.cpp:2619;
// LocalTreeBuilder, Compute.cpp:2440-2580) for one domain. Particles are
// ordered along a Morton space-filling curve whose top key bit splits x, then
// y, then z, the axis cycle of BinaryTreeNode::makeOctChildren
// (GenericTreeNode.h:473-575). A node becomes a bucket when it holds at most
// nBucket particles (Compute.cpp:2476). Chains of nodes with a single
// non-empty child are collapsed: such a node has the same particles, centre
// of mass and bounding box as its child.
//
// Node numbering is breadth-first, so every child index exceeds its parent's
// and the nodes of one depth are contiguous (the GPU moments build sweeps the
// depths bottom-up). Buckets are numbered in particle order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "host_tree.h"

namespace cgh {

static inline uint64_t spread3(uint64_t v) {
  // 21 bits -> every third bit
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

static void radix_sort_keys(std::vector<uint64_t>& keys, std::vector<int32_t>& idx) {
  const size_t n = keys.size();
  std::vector<uint64_t> k2(n);
  std::vector<int32_t> i2(n);
  const int BITS = 11, NB = 1 << BITS;
  std::vector<size_t> cnt(NB);
  for (int shift = 0; shift < 63; shift += BITS) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) cnt[(keys[i] >> shift) & (NB - 1)]++;
    size_t s = 0;
    for (int b = 0; b < NB; ++b) {
      size_t c = cnt[b];
      cnt[b] = s;
      s += c;
    }
    for (size_t i = 0; i < n; ++i) {
      size_t d = cnt[(keys[i] >> shift) & (NB - 1)]++;
      k2[d] = keys[i];
      i2[d] = idx[i];
    }
    keys.swap(k2);
    idx.swap(i2);
  }
}

Tree* build_tree(int n, const double* pos, const double box_lo[3], double box_size,
                 int bucket_size) {
  Tree* t = new Tree();
  t->n = n;
  t->bucket_size = bucket_size;
  t->box_size = box_size;
  for (int k = 0; k < 3; ++k) t->box_lo[k] = box_lo[k];
  std::vector<uint64_t> keys(n);
  t->perm.resize(n);
  const double scale = double(1 << 21) / box_size;
  for (int i = 0; i < n; ++i) {
    uint64_t c[3];
    for (int k = 0; k < 3; ++k) {
      double u = (pos[3 * i + k] - box_lo[k]) * scale;
      int64_t v = (int64_t)std::floor(u);
      if (v < 0) v = 0;
      if (v > (1 << 21) - 1) v = (1 << 21) - 1;
      c[k] = (uint64_t)v;
    }
    keys[i] = (spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]);
    t->perm[i] = i;
  }
  radix_sort_keys(keys, t->perm);

  struct Item {
    int32_t first, count, bit, depth;
  };
  std::vector<Item> cur, next;
  cur.push_back({0, n, 62, 0});
  int depth = 0;
  while (!cur.empty()) {
    t->level_offsets.push_back((int32_t)t->first.size());
    next.clear();
    for (const Item& it : cur) {
      int32_t id = (int32_t)t->first.size();
      t->first.push_back(it.first);
      t->count.push_back(it.count);
      t->child.push_back(-1);
      t->child.push_back(-1);
      t->depth.push_back(depth);
      // half diagonal of the Morton cell this node was created on
      int level = 62 - it.bit;
      int nx = (level + 2) / 3, ny = (level + 1) / 3, nz = level / 3;
      double ex = box_size / std::ldexp(1.0, nx), ey = box_size / std::ldexp(1.0, ny),
             ez = box_size / std::ldexp(1.0, nz);
      t->box_radius.push_back(0.5 * std::sqrt(ex * ex + ey * ey + ez * ez));
      if (it.count <= bucket_size) continue;
      // find the first key bit on which the range splits
      int bit = it.bit;
      const int32_t lo = it.first, hi = it.first + it.count - 1;
      int32_t split = -1;
      while (bit >= 0) {
        uint64_t mask = 1ULL << bit;
        if ((keys[lo] & mask) != (keys[hi] & mask)) {
          // keys share every bit above `bit`: the first key with the bit set
          auto p = std::partition_point(keys.begin() + lo, keys.begin() + hi + 1,
                                        [mask](uint64_t k) { return (k & mask) == 0; });
          split = (int32_t)(p - keys.begin());
          break;
        }
        --bit;
      }
      if (split < 0) continue;  // coincident keys: an oversized bucket
      t->child[2 * id] = -2;    // filled when the children are numbered
      next.push_back({lo, split - lo, bit - 1, depth + 1});
      next.push_back({split, hi - split + 1, bit - 1, depth + 1});
    }
    // number the children: they follow in `next` order
    int32_t base = (int32_t)t->first.size();
    int32_t k = 0;
    for (int32_t id = t->level_offsets.back(); id < base; ++id) {
      if (t->child[2 * id] == -2) {
        t->child[2 * id] = base + k;
        t->child[2 * id + 1] = base + k + 1;
        k += 2;
      }
    }
    cur.swap(next);
    ++depth;
  }
  t->level_offsets.push_back((int32_t)t->first.size());
  t->n_nodes = (int32_t)t->first.size();
  // buckets in particle order
  for (int32_t id = 0; id < t->n_nodes; ++id)
    if (t->child[2 * id] < 0) t->bucket_node.push_back(id);
  std::sort(t->bucket_node.begin(), t->bucket_node.end(),
            [t](int32_t a, int32_t b) { return t->first[a] < t->first[b]; });
  t->node_bucket.assign(t->n_nodes, -1);
  for (size_t b = 0; b < t->bucket_node.size(); ++b) t->node_bucket[t->bucket_node[b]] = (int32_t)b;
  return t;
}

// Node geometry in FP64: mass, centre of mass, mass-weighted softening,
// particle bounding box and the opening/expansion radius (farthest particle
// for a bucket, calculateRadiusFarthestParticle, MultipoleMoments.h:489-505;
// farthest bounding-box corner otherwise, calculateRadiusFarthestCorner,
// MultipoleMoments.h:460-471). geo[12*n] = cm.xyz, radius, soft, mass,
// lo.xyz, hi.xyz. The arithmetic is ordered as in the CPU checker so the two
// agree bit for bit.
void node_geometry(int n_nodes, const int32_t* first, const int32_t* count, const int32_t* child,
                   const double* box_radius, const double* pos, const double* mass,
                   const double* soft, double* geo) {
  for (int nd = n_nodes - 1; nd >= 0; --nd) {
    double* g = geo + (size_t)nd * 12;
    int c0 = child[2 * nd], c1 = child[2 * nd + 1];
    if (c0 < 0) {
      int f = first[nd], cnt = count[nd];
      double M = 0, sx = 0, sy = 0, sz = 0, ss = 0;
      for (int k = 0; k < 3; ++k) {
        g[6 + k] = HUGE_VAL;
        g[9 + k] = -HUGE_VAL;
      }
      for (int i = f; i < f + cnt; ++i) {
        M += mass[i];
        sx += mass[i] * pos[3 * i];
        sy += mass[i] * pos[3 * i + 1];
        sz += mass[i] * pos[3 * i + 2];
        ss += mass[i] * soft[i];
        for (int k = 0; k < 3; ++k) {
          if (pos[3 * i + k] < g[6 + k]) g[6 + k] = pos[3 * i + k];
          if (pos[3 * i + k] > g[9 + k]) g[9 + k] = pos[3 * i + k];
        }
      }
      g[0] = sx / M;
      g[1] = sy / M;
      g[2] = sz / M;
      g[4] = ss / M;
      g[5] = M;
      double r2max = 0.0;
      for (int i = f; i < f + cnt; ++i) {
        double dx = g[0] - pos[3 * i], dy = g[1] - pos[3 * i + 1], dz = g[2] - pos[3 * i + 2];
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > r2max) r2max = d2;
      }
      g[3] = (cnt > 1 && r2max > 0.0) ? std::sqrt(r2max) : box_radius[nd];
    } else {
      const double* g0 = geo + (size_t)c0 * 12;
      const double* g1 = geo + (size_t)c1 * 12;
      double M0 = g0[5], M1 = g1[5], M = M0 + M1;
      g[5] = M;
      g[4] = (M0 * g0[4] + M1 * g1[4]) / M;
      for (int k = 0; k < 3; ++k) g[k] = (M0 * g0[k] + M1 * g1[k]) / M;
      for (int k = 0; k < 3; ++k) {
        g[6 + k] = g0[6 + k] < g1[6 + k] ? g0[6 + k] : g1[6 + k];
        g[9 + k] = g0[9 + k] > g1[9 + k] ? g0[9 + k] : g1[9 + k];
      }
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        double e1 = g[k] - g[6 + k], e2 = g[9 + k] - g[k];
        double e = e1 > e2 ? e1 : e2;
        d2 += e * e;
      }
      g[3] = std::sqrt(d2);
    }
  }
}

}  // namespace cgh
