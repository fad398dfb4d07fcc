// What a tree walk reads of a node, and the opening test, shared by the
// list builder (walk.hip) and the fused walk (fused.hip) so that both make
// the same decision on every node.
//
// openCriterionBucket (gravity.h:597-635): a node with at most
// nMinParticleNode = 6 particles is always opened; otherwise its opening
// sphere, radius max(2/sqrt(3) r / theta, r) (TreeNode.h:35), is tested
// against the target bucket's bounding box. A well-separated node whose
// softening overlaps the bucket (openSoftening, gravity.h:251-260) is
// re-tested with the monopole radius 2/sqrt(3) r / theta^4 (theta_mono,
// TreePiece.cpp:5036): opened if it overlaps, else a softened cell.
// Everything in FP64 with contraction off, in the host walk's operation
// order, so the decisions are the host walk's on the same node geometry.
#pragma once
#include "common.h"

namespace cgpu {

// 64 bytes, one cache line per visited node: the centre of mass, the
// opening radius, twice the softening, the monopole opening radius, the
// particle count, the children (c0 < 0: a bucket) and the first particle.
struct alignas(16) WalkRec {
  double cx, cy, cz, rad, rs, rm;
  int count, c0, c1, first;
};
static_assert(sizeof(WalkRec) == 64, "one cache line per node");

// The records of nodes [0, nn) from the node records of changaBuildMoments
// (node64: cm.xyz, radius, soft, mass, lo.xyz, hi.xyz, moments; 34 doubles).
hipError_t walk_records(int nn, const double* node64, const int* first, const int* count,
                        const int* child, double theta, WalkRec* rec, hipStream_t s);

#pragma clang fp contract(off)

__device__ __forceinline__ bool walk_box_sphere(const double* lo, const double* hi,
                                                const double c[3], double r) {
  double dsq = 0.0;
#pragma unroll
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

// 0 open, 1 hexadecapole cell, 2 softened cell. bg: the target bucket's
// node record (cm at 0-2, lo at 6-8, hi at 9-11), rb twice its softening;
// (ox, oy, oz) the replica offset.
__device__ __forceinline__ int walk_decide(const WalkRec& v, const double* bg, double rb, double ox,
                                           double oy, double oz) {
  if (v.count <= 6) return 0;
  const double c3[3] = {v.cx + ox, v.cy + oy, v.cz + oz};
  if (walk_box_sphere(bg + 6, bg + 9, c3, v.rad)) return 0;
  const double rs = v.rs;
  const double dx = c3[0] - bg[0], dy = c3[1] - bg[1], dz = c3[2] - bg[2];
  const bool soft_overlap = dx * dx + dy * dy + dz * dz <= (rs + rb) * (rs + rb) ||
                            walk_box_sphere(bg + 6, bg + 9, c3, rs);
  if (!soft_overlap) return 1;
  return walk_box_sphere(bg + 6, bg + 9, c3, v.rm) ? 0 : 2;
}

#pragma clang fp contract(on)

}  // namespace cgpu
