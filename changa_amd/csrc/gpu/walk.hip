// GPU interaction-list construction: for every target bucket, the walk of the
// tree (and its periodic replicas) that emits ChaNGa's GPU wire lists.
//
// CPU semantics followed: ListCompute::doWork / stateReady (Compute.cpp:
// 690-884, 1543-1922) with openCriterionBucket (gravity.h:597-635): a node
// with at most nMinParticleNode = 6 particles is always opened; otherwise its
// opening sphere, radius max(2/sqrt(3) * r / theta, r) (TreeNode.h:35), is
// tested against the bucket's bounding box; a well-separated node whose
// softening overlaps the bucket (openSoftening, gravity.h:251-260) is
// re-tested with the monopole radius 2/sqrt(3) * r / theta^4 (theta_mono,
// TreePiece.cpp:5036). An accepted node becomes an ILCell{node,
// encodeOffset(bucket, x, y, z)} (TreePiece.cpp:3631-3645), flagged in bit 31
// when softened; an opened bucket becomes an ILPart{first particle,
// encodeOffset(0, x, y, z), count}. Per bucket: hexadecapole cells, then
// softened cells, then particle runs (the layout ListCompute::stateReady
// serializes, Compute.cpp:1735-1741, 1806).
//
// One lane per target bucket walks depth first with a private stack, in the
// host walk's order (replicas from (-R, -R, -R) to (R, R, R), z fastest;
// the first child first), so the lists are the host walk's entry for entry.
// The geometry tests run in FP64 with contraction off, so every opening
// decision is the host's on the same node geometry. Two passes: count (per
// bucket: hexadecapole cells, softened cells, runs, source particles), then,
// after the caller allocated the lists, fill.
#include <rocprim/device/device_scan.hpp>

#include "common.h"

namespace cgpu {

constexpr int kWalkStack = 72;  // depth-first stack: tree depth + 1 (Morton trees: <= 64)
constexpr int kWalkBlock = 64;

struct WalkArgs {
  int nb;
  int* nCell;  // count pass: cells (hexadecapole + softened) and runs per bucket
  int* nPart;
  const int* bucketNode;
  const int* first;
  const int* count;
  const int* child;
  const double* node64;
  double kgeom, theta, thetaMono, period;
  int R;
  int* cellMarks;
  int* partMarks;
  int* softStart;
  int* sources;
  ILCell* cells;
  ILPart* parts;
  int* err;
};

#pragma clang fp contract(off)

__device__ __forceinline__ bool box_sphere(const double* lo, const double* hi, const double c[3],
                                           double r) {
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

// node record (node64): cm.xyz, radius, soft, mass, lo.xyz, hi.xyz, moments
template <bool FILL>
__global__ __launch_bounds__(kWalkBlock) void walk_kernel(WalkArgs a) {
  const int b = blockIdx.x * kWalkBlock + threadIdx.x;
  if (b >= a.nb) return;
  double bg[12];
  {
    const double* g = a.node64 + (size_t)a.bucketNode[b] * 34;
#pragma unroll
    for (int k = 0; k < 12; ++k) bg[k] = g[k];
  }
  const double rb = 2.0 * bg[4];
  int nhex = 0, nsoft = 0, nruns = 0, nsrc = 0;
  int hpos = 0, spos = 0, ppos = 0;
  if (FILL) {
    hpos = a.cellMarks[b];
    spos = a.softStart[b];
    ppos = a.partMarks[b];
  }
  int stack[kWalkStack];
  bool overflow = false;
  for (int x = -a.R; x <= a.R && !overflow; ++x)
    for (int y = -a.R; y <= a.R && !overflow; ++y)
      for (int z = -a.R; z <= a.R && !overflow; ++z) {
        const int oid = ((x + 3) | ((y + 3) << 3) | ((z + 3) << 6)) << 22;
        const double off[3] = {x * a.period, y * a.period, z * a.period};
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
          const int nd = stack[--sp];
          const double* g = a.node64 + (size_t)nd * 34;
          const int cnt = a.count[nd];
          bool open = true, softened = false;
          if (cnt > 6) {
            const double c3[3] = {g[0] + off[0], g[1] + off[1], g[2] + off[2]};
            const double r = g[3];
            double rad = a.kgeom * r / a.theta;
            if (rad < r) rad = r;
            if (!box_sphere(bg + 6, bg + 9, c3, rad)) {
              // openSoftening: the node's softening sphere against the bucket's
              const double rs = 2.0 * g[4];
              const double dx = c3[0] - bg[0], dy = c3[1] - bg[1], dz = c3[2] - bg[2];
              const bool soft_overlap = dx * dx + dy * dy + dz * dz <= (rs + rb) * (rs + rb) ||
                                        box_sphere(bg + 6, bg + 9, c3, rs);
              if (!soft_overlap) {
                open = false;
              } else {
                const double rm = a.kgeom * r / a.thetaMono;
                if (!box_sphere(bg + 6, bg + 9, c3, rm)) {
                  open = false;
                  softened = true;
                }
              }
            }
          }
          if (!open) {
            if (softened) {
              if (FILL) a.cells[spos + nsoft] = {nd, oid | b | (int)0x80000000u};
              ++nsoft;
            } else {
              if (FILL) a.cells[hpos + nhex] = {nd, oid | b};
              ++nhex;
            }
          } else {
            const int c0 = a.child[2 * nd];
            if (c0 < 0) {
              if (FILL) a.parts[ppos + nruns] = {a.first[nd], oid, cnt};
              ++nruns;
              nsrc += cnt;
            } else {
              if (sp + 2 > kWalkStack) {
                overflow = true;
                break;
              }
              stack[sp++] = a.child[2 * nd + 1];
              stack[sp++] = c0;
            }
          }
        }
      }
  if (overflow) {
    atomicExch(a.err, 1);
    return;
  }
  if (!FILL) {
    a.nCell[b] = nhex + nsoft;
    a.nPart[b] = nruns;
    a.softStart[b] = nhex;
    if (a.sources) a.sources[b] = nsrc;
    // 64-bit totals, to reject lists beyond the int32 marker range
    unsigned long long* tot = reinterpret_cast<unsigned long long*>(a.err + 2);
    atomicAdd(tot, (unsigned long long)(nhex + nsoft));
    atomicAdd(tot + 1, (unsigned long long)nruns);
  }
}

#pragma clang fp contract(on)

// after the scans: softStart = first softened cell of the bucket
__global__ void walk_soft_start_kernel(int nb, const int* cellMarks, int* softStart) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) softStart[b] += cellMarks[b];
}

__global__ void walk_init_kernel(int nb, int* nCell, int* nPart, int* err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    nCell[nb] = 0;
    nPart[nb] = 0;
    err[0] = 0;
    unsigned long long* tot = reinterpret_cast<unsigned long long*>(err + 2);
    tot[0] = 0;
    tot[1] = 0;
  }
}

// work buffer: error flag and two 64-bit totals (256 B), per-bucket counts
// (2 x (nb + 1) ints, 256-byte aligned), rocPRIM scan storage
struct WalkWork {
  size_t counts, temp, total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t scan_bytes(int nb) {
  size_t bytes = 0;
  if (rocprim::exclusive_scan(nullptr, bytes, (const int*)nullptr, (int*)nullptr, 0,
                              (size_t)nb + 1, rocprim::plus<int>()) != hipSuccess)
    return 0;
  return bytes;
}

static int check_args(const ChangaListBuild* L, bool fill) {
  if (!L || L->nBuckets < 0 || (L->nBuckets > 0 && (!L->bucketNode || !L->first || !L->count ||
                                                      !L->child || !L->node64 ||
                                                      !L->cellMarks || !L->partMarks ||
                                                      !L->softStart)))
    return fail("changaBuildLists: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  if (L->nReplicas < 0 || L->nReplicas > 3 || !(L->theta > 0.0) ||
      (L->nReplicas > 0 && !(L->period > 0.0)))
    return fail("changaBuildLists: theta > 0, 0 <= nReplicas <= 3 (period > 0 when periodic)",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (fill && L->nBuckets > 0 && (!L->cells || !L->parts))
    return fail("changaBuildListsFill: cells and parts", hipErrorInvalidValue, __FILE__,
                __LINE__);
  return 0;
}

static WalkWork work_layout(int nb, size_t scan) {
  WalkWork w;
  w.counts = 256;
  w.temp = w.counts + align256(2 * sizeof(int) * ((size_t)nb + 1));
  w.total = w.temp + align256(scan);
  return w;
}

static WalkArgs walk_args(const ChangaListBuild* L, int* err) {
  WalkArgs a;
  a.nb = L->nBuckets;
  char* base = static_cast<char*>(L->work);
  a.nCell = reinterpret_cast<int*>(base + 256);
  a.nPart = a.nCell + (L->nBuckets + 1);
  a.bucketNode = L->bucketNode;
  a.first = L->first;
  a.count = L->count;
  a.child = L->child;
  a.node64 = L->node64;
  a.kgeom = 2.0 / sqrt(3.0);
  a.theta = L->theta;
  a.thetaMono = L->theta * L->theta * L->theta * L->theta;  // the host walk's order
  a.period = L->nReplicas > 0 ? L->period : 0.0;
  a.R = L->nReplicas;
  a.cellMarks = L->cellMarks;
  a.partMarks = L->partMarks;
  a.softStart = L->softStart;
  a.sources = L->sourceCounts;
  a.cells = L->cells;
  a.parts = L->parts;
  a.err = err;
  return a;
}

}  // namespace cgpu

using namespace cgpu;

extern "C" size_t changaBuildListsWorkBytes(int nBuckets) {
  if (nBuckets < 0) return 0;
  const size_t s = scan_bytes(nBuckets);
  return s == 0 ? 0 : work_layout(nBuckets, s).total;
}

extern "C" int changaBuildListsCount(const ChangaListBuild* L, long long* numCells,
                                     long long* numParts, void* stream) {
  if (int e = check_args(L, false)) return e;
  hipStream_t s = (hipStream_t)stream;
  const int nb = L->nBuckets;
  if (!L->work || L->workBytes < changaBuildListsWorkBytes(nb))
    return fail("changaBuildListsCount: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  int* err = static_cast<int*>(L->work);
  const WalkWork W = work_layout(nb, scan_bytes(nb));
  void* temp = static_cast<char*>(L->work) + W.temp;
  const WalkArgs args = walk_args(L, err);
  hipLaunchKernelGGL(walk_init_kernel, dim3(1), dim3(1), 0, s, nb, args.nCell, args.nPart, err);
  CGPU_CHECK(hipGetLastError());
  if (nb > 0) {
    hipLaunchKernelGGL(walk_kernel<false>, dim3((nb + kWalkBlock - 1) / kWalkBlock),
                       dim3(kWalkBlock), 0, s, args);
    CGPU_CHECK(hipGetLastError());
  }
  size_t tb = L->workBytes - W.temp;
  CGPU_CHECK(rocprim::exclusive_scan(temp, tb, args.nCell, L->cellMarks, 0, (size_t)nb + 1,
                                     rocprim::plus<int>(), s));
  tb = L->workBytes - W.temp;
  CGPU_CHECK(rocprim::exclusive_scan(temp, tb, args.nPart, L->partMarks, 0, (size_t)nb + 1,
                                     rocprim::plus<int>(), s));
  if (nb > 0) {
    hipLaunchKernelGGL(walk_soft_start_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, nb,
                       L->cellMarks, L->softStart);
    CGPU_CHECK(hipGetLastError());
  }
  int flag = 0;
  unsigned long long tot[2] = {0, 0};
  CGPU_CHECK(hipMemcpyAsync(&flag, err, sizeof(int), hipMemcpyDeviceToHost, s));
  CGPU_CHECK(hipMemcpyAsync(tot, err + 2, sizeof(tot), hipMemcpyDeviceToHost, s));
  CGPU_CHECK(hipStreamSynchronize(s));
  if (flag)
    return fail("changaBuildListsCount: tree deeper than the walk stack", hipErrorInvalidValue,
                __FILE__, __LINE__);
  if (tot[0] > 0x7fffffffull || tot[1] > 0x7fffffffull)
    return fail("changaBuildListsCount: list exceeds the int32 marker range",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (numCells) *numCells = (long long)tot[0];
  if (numParts) *numParts = (long long)tot[1];
  return 0;
}

extern "C" int changaBuildListsFill(const ChangaListBuild* L, void* stream) {
  if (int e = check_args(L, true)) return e;
  hipStream_t s = (hipStream_t)stream;
  const int nb = L->nBuckets;
  if (!L->work || L->workBytes < changaBuildListsWorkBytes(nb))
    return fail("changaBuildListsFill: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (nb > 0) {
    hipLaunchKernelGGL(walk_kernel<true>, dim3((nb + kWalkBlock - 1) / kWalkBlock),
                       dim3(kWalkBlock), 0, s, walk_args(L, static_cast<int*>(L->work)));
    CGPU_CHECK(hipGetLastError());
  }
  return 0;
}
