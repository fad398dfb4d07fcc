// Interaction-list builder: for every target bucket, walks the tree (and its
// periodic replicas) and emits ChaNGa's GPU wire lists.
//
// The opening test is openCriterionBucket (gravity.h:597-635) with the
// hexadecapole softening branch: a node with at most nMinParticleNode = 6
// particles is always opened; otherwise its opening sphere, radius
// max(2/sqrt(3) * r / theta, r) (TreeNode.h:35), is tested against the
// bucket's bounding box; a well-separated node whose softening overlaps the
// bucket (openSoftening, gravity.h:251-260) is re-tested with the monopole
// radius 2/sqrt(3) * r / theta^4 (theta_mono, TreePiece.cpp:5036). An
// accepted node becomes an ILCell; an opened bucket becomes an ILPart over its
// particles (ListCompute::doWork, Compute.cpp:690-884, with the LocalOpt
// action table, Opt.h:85-128). A per-bucket walk yields the same lists as
// ChaNGa's double walk: decisions made against an ancestor of the bucket are
// ones the bucket's own test would make too, because both tests are monotone
// in the target box.
//
// Wire layout (cuda_typedef.h:208-235, ListCompute::stateReady,
// Compute.cpp:1735-1741, 1806): ILCell{index, offsetID} with offsetID =
// encodeOffset(bucket, x, y, z) (TreePiece.cpp:3631-3645); ILPart{first,
// encodeOffset(0, x, y, z), num}. Cells whose softening overlaps the bucket,
// which ChaNGa evaluates on the CPU as a monopole pseudo-particle
// (Compute.cpp:1683-1698), are kept on the GPU list, flagged in offsetID bit
// 31 and placed after the bucket's hexadecapole cells.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "host_tree.h"

namespace cgh {

namespace {

inline bool box_sphere(const double* lo, const double* hi, const double c[3], double r) {
  double dsq = 0.0;
  for (int k = 0; k < 3; ++k) {
    double e = 0.0;
    if (c[k] < lo[k])
      e = lo[k] - c[k];
    else if (c[k] > hi[k])
      e = c[k] - hi[k];
    dsq += e * e;
  }
  return dsq <= r * r;
}

inline bool open_softening(const double* cg, const double* bg, const double off[3]) {
  double c[3] = {cg[0] + off[0], cg[1] + off[1], cg[2] + off[2]};
  double rs = 2.0 * cg[4], rb = 2.0 * bg[4];
  double dx = c[0] - bg[0], dy = c[1] - bg[1], dz = c[2] - bg[2];
  if (dx * dx + dy * dy + dz * dz <= (rs + rb) * (rs + rb)) return true;
  return box_sphere(bg + 6, bg + 9, c, rs);
}

inline int32_t encode_offset(int32_t req, int x, int y, int z) {
  int32_t code = (x + 3) | ((y + 3) << 3) | ((z + 3) << 6);
  return req | (code << 22);
}

struct Chunk {
  std::vector<int32_t> cells, parts;
  std::vector<int64_t> cell_len, soft_len, part_len;
  int64_t part_sources = 0;
};

}  // namespace

Lists* walk(int n_buckets, const int32_t* bucket_node, const int32_t* node_count,
            const int32_t* node_first, const int32_t* child, const double* geo,
            const WalkParams& p) {
  const double kgeom = 2.0 / std::sqrt(3.0);
  const double theta = p.theta, theta_mono = theta * theta * theta * theta;
  const int R = p.n_replicas;
  const int CH = 256;  // buckets per work chunk
  const int n_chunks = (n_buckets + CH - 1) / CH;
  std::vector<Chunk> chunks(n_chunks);
  std::atomic<int> next(0);
  int nt = p.n_threads > 0 ? p.n_threads : (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;

  auto worker = [&]() {
    std::vector<std::pair<int32_t, int32_t>> stack;
    std::vector<int32_t> hex, soft;
    for (;;) {
      int c = next.fetch_add(1);
      if (c >= n_chunks) break;
      Chunk& ch = chunks[c];
      int b0 = c * CH, b1 = std::min(n_buckets, b0 + CH);
      for (int b = b0; b < b1; ++b) {
        const double* bg = geo + (size_t)bucket_node[b] * 12;
        hex.clear();
        soft.clear();
        size_t parts0 = ch.parts.size();
        int64_t np = 0;
        stack.clear();
        for (int x = R; x >= -R; --x)
          for (int y = R; y >= -R; --y)
            for (int z = R; z >= -R; --z) stack.push_back({0, encode_offset(0, x, y, z)});
        while (!stack.empty()) {
          auto [nd, oid] = stack.back();
          stack.pop_back();
          int code = (oid >> 22) & 0x1ff;
          double off[3] = {((code & 7) - 3) * p.period, (((code >> 3) & 7) - 3) * p.period,
                           (((code >> 6) & 7) - 3) * p.period};
          const double* g = geo + (size_t)nd * 12;
          int open = 1, softened = 0;
          if (node_count[nd] > 6) {
            double c3[3] = {g[0] + off[0], g[1] + off[1], g[2] + off[2]};
            double rad = kgeom * g[3] / theta;
            if (rad < g[3]) rad = g[3];
            if (!box_sphere(bg + 6, bg + 9, c3, rad)) {
              if (!open_softening(g, bg, off)) {
                open = 0;
              } else {
                double rm = kgeom * g[3] / theta_mono;
                if (!box_sphere(bg + 6, bg + 9, c3, rm)) {
                  open = 0;
                  softened = 1;
                }
              }
            }
          }
          if (!open) {
            int32_t o = (oid & ~0x3fffff) | b;
            if (softened)
              soft.push_back(nd), soft.push_back(o | (int32_t)0x80000000u);
            else
              hex.push_back(nd), hex.push_back(o);
          } else if (child[2 * nd] < 0) {
            ch.parts.push_back(node_first[nd]);
            ch.parts.push_back(oid & ~0x3fffff);
            ch.parts.push_back(node_count[nd]);
            np += node_count[nd];
          } else {
            stack.push_back({child[2 * nd + 1], oid});
            stack.push_back({child[2 * nd], oid});
          }
        }
        // openSoftening of a cell that is not flagged is false by
        // construction; a flagged one is true
        ch.cells.insert(ch.cells.end(), hex.begin(), hex.end());
        ch.cells.insert(ch.cells.end(), soft.begin(), soft.end());
        ch.cell_len.push_back((int64_t)(hex.size() + soft.size()) / 2);
        ch.soft_len.push_back((int64_t)soft.size() / 2);
        ch.part_len.push_back((int64_t)(ch.parts.size() - parts0) / 3);
        ch.part_sources += np;
      }
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i) th.emplace_back(worker);
  for (auto& x : th) x.join();

  Lists* L = new Lists();
  L->n_buckets = n_buckets;
  int64_t nc = 0, npp = 0;
  for (auto& ch : chunks) {
    nc += (int64_t)ch.cells.size() / 2;
    npp += (int64_t)ch.parts.size() / 3;
    L->n_part_sources += ch.part_sources;
  }
  L->cells.resize(2 * nc);
  L->parts.resize(3 * npp);
  L->cell_marks.resize(n_buckets + 1);
  L->part_marks.resize(n_buckets + 1);
  L->soft_start.resize(n_buckets);
  int64_t cpos = 0, ppos = 0;
  int b = 0;
  for (auto& ch : chunks) {
    std::copy(ch.cells.begin(), ch.cells.end(), L->cells.begin() + 2 * cpos);
    std::copy(ch.parts.begin(), ch.parts.end(), L->parts.begin() + 3 * ppos);
    for (size_t i = 0; i < ch.cell_len.size(); ++i, ++b) {
      L->cell_marks[b] = cpos;
      L->soft_start[b] = cpos + ch.cell_len[i] - ch.soft_len[i];
      L->part_marks[b] = ppos;
      cpos += ch.cell_len[i];
      ppos += ch.part_len[i];
    }
    std::vector<int32_t>().swap(ch.cells);
    std::vector<int32_t>().swap(ch.parts);
  }
  L->cell_marks[n_buckets] = cpos;
  L->part_marks[n_buckets] = ppos;
  return L;
}

}  // namespace cgh
