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
// particle count, the first child (c0 < 0: a bucket), a link and the first
// particle. The link is the second child for the cooperative walk and the
// escape node (the depth-first successor of the node's subtree) for the
// stackless depth-first walk of the list builder.
struct alignas(16) WalkRec {
  double cx, cy, cz, rad, rs, rm;
  int count, c0, link, first;
};
static_assert(sizeof(WalkRec) == 64, "one cache line per node");

// The records of nodes [0, nn) from the node records of changaBuildMoments
// (node64: cm.xyz, radius, soft, mass, lo.xyz, hi.xyz, moments; 34 doubles);
// link = esc[node] when esc is given, else the second child.
hipError_t walk_records(int nn, const double* node64, const int* first, const int* count,
                        const int* child, const int* esc, double theta, WalkRec* rec,
                        hipStream_t s);

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

// ---------------------------------------------------------------------------
// Wavefront-cooperative walk of one target bucket over an LDS stack of
// 32-bit entries (node index in the low kCoopNodeBits bits, the replica code
// (x+3) | (y+3) << 3 | (z+3) << 6 above). Each iteration pops up to 64
// entries, one per lane, tests them, and reports the accepted ones through
// `emit(kind, nd, code, v, slot)` (kind 1 hexadecapole cell, 2 softened
// cell, 3 opened bucket; slot = the entry's index among the bucket's
// entries of that kind, from wavefront prefix counts); opened internal nodes
// push their two children. The set of accepted entries is the depth-first
// walk's; their order is not. Returns false when a capacity check fails
// (`room(hex, soft, runs)` false) or the stack would overflow: near a full
// stack fewer entries are popped, so the walk turns depth first.
// ---------------------------------------------------------------------------
constexpr int kCoopStack = 2304;  // entries: 9 KB of LDS
constexpr int kCoopNodeBits = 23;

__device__ __forceinline__ int lanes_below(unsigned long long mask, int lane) {
  return __popcll(mask & ((1ull << lane) - 1ull));
}

template <typename Room, typename Emit>
__device__ __forceinline__ bool coop_walk(const WalkRec* rec, const double* bg, double rb, int R,
                                          double period, int lane, unsigned* stack, int& nhex,
                                          int& nsoft, int& nruns, Room&& room, Emit&& emit) {
  nhex = nsoft = nruns = 0;
  const int W = 2 * R + 1, nrep = W * W * W;
  for (int k = lane; k < nrep; k += kWave) {
    const int x = k / (W * W) - R, y = (k / W) % W - R, z = k % W - R;
    const unsigned code = (unsigned)((x + 3) | ((y + 3) << 3) | ((z + 3) << 6));
    stack[k] = code << kCoopNodeBits;  // the root of every replica
  }
  int sp = nrep;
  bool ok = true;
  __syncthreads();
  while (sp > 0) {
    // each popped entry pushes at most two: popping at most (capacity - sp)
    // keeps the stack within capacity
    const int n = min(min(kWave, sp), max(1, kCoopStack - sp));
    unsigned e = 0;
    if (lane < n) e = stack[sp - n + lane];
    __syncthreads();
    sp -= n;
    int kind = -1, nd = 0, code = 0;  // -1 none, 0 open, 1 hex, 2 soft
    WalkRec v = {};
    if (lane < n) {
      nd = (int)(e & ((1u << kCoopNodeBits) - 1u));
      code = (int)(e >> kCoopNodeBits);
      v = rec[nd];
      const int x = (code & 7) - 3, y = ((code >> 3) & 7) - 3, z = ((code >> 6) & 7) - 3;
      kind = walk_decide(v, bg, rb, x * period, y * period, z * period);
    }
    const bool is_bucket = kind == 0 && v.c0 < 0;
    const bool is_push = kind == 0 && !is_bucket;
    const unsigned long long mh = __ballot(kind == 1), ms = __ballot(kind == 2),
                             mr = __ballot(is_bucket), mp = __ballot(is_push);
    const int ch = __popcll(mh), cs = __popcll(ms), cr = __popcll(mr), cp = __popcll(mp);
    if (!room(nhex + ch, nsoft + cs, nruns + cr) || sp + 2 * cp > kCoopStack) {
      ok = false;
      break;
    }
    if (kind == 1) emit(1, nd, code, v, nhex + lanes_below(mh, lane));
    if (kind == 2) emit(2, nd, code, v, nsoft + lanes_below(ms, lane));
    if (is_bucket) emit(3, nd, code, v, nruns + lanes_below(mr, lane));
    if (is_push) {
      const int at = sp + 2 * lanes_below(mp, lane);
      const unsigned hi = (unsigned)code << kCoopNodeBits;
      stack[at] = hi | (unsigned)v.link;  // the second child
      stack[at + 1] = hi | (unsigned)v.c0;
    }
    nhex += ch;
    nsoft += cs;
    nruns += cr;
    sp += 2 * cp;
    __syncthreads();
  }
  __syncthreads();
  return ok;
}

}  // namespace cgpu
