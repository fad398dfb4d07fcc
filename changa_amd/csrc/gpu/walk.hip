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
// One lane per target bucket walks depth first, in the
// host walk's order (replicas from (-R, -R, -R) to (R, R, R), z fastest;
// the first child first), so the lists are the host walk's entry for entry.
// The nodes are read as 64-byte walk records and tested in FP64 with
// contraction off (walk_rec.h), so every opening decision is the host's on
// the same node geometry. Two passes: count (per bucket: hexadecapole cells,
// softened cells, runs, source particles; it also writes the walk records and
// the escape nodes that let a lane walk depth first without a stack),
// then, after the caller allocated the lists, fill. (A wavefront-cooperative
// count walk, one wavefront per bucket, measured slower: a lane per bucket
// keeps 64 times more independent node loads in flight.)
#include <rocprim/device/device_scan.hpp>

#include "common.h"
#include "walk_rec.h"

namespace cgpu {

constexpr int kWalkBlock = 64;
constexpr int kEscPasses = 8;  // pointer-jumping passes (even): chains up to 2^8 > the tree depth

struct WalkArgs {
  int nb;
  int* nCell;  // count pass: cells (hexadecapole + softened) and runs per bucket
  int* nPart;
  const int* bucketNode;
  const double* node64;  // the target buckets' geometry
  const WalkRec* rec;    // walk records of every node (link: the escape node)
  const int* esc;        // escape node of every node (depth-first successor of its subtree)
  double period;
  int R;
  int* cellMarks;
  int* partMarks;
  int* softStart;
  int* sources;
  ILCell* cells;
  ILPart* parts;
  int* err;
};

// node record (node64): cm.xyz, radius, soft, mass, lo.xyz, hi.xyz, moments
template <bool FILL>
__global__ __launch_bounds__(kWalkBlock) void walk_kernel(WalkArgs a) {
  // blocks of 64 consecutive buckets, each XCD taking a contiguous range of
  // blocks so that its L2 holds the nodes near its buckets
  const int b = xcd_bucket(blockIdx.x, gridDim.x) * kWalkBlock + threadIdx.x;
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
  // depth first without a stack: an opened internal node continues with its
  // first child; anything else continues with its escape node (the next node
  // of the depth-first order after its subtree; -1 after the last)
  for (int x = -a.R; x <= a.R; ++x)
    for (int y = -a.R; y <= a.R; ++y)
      for (int z = -a.R; z <= a.R; ++z) {
        const int oid = ((x + 3) | ((y + 3) << 3) | ((z + 3) << 6)) << 22;
        const double ox = x * a.period, oy = y * a.period, oz = z * a.period;
        int nd = 0;
        while (nd >= 0) {
          const WalkRec v = a.rec[nd];
          const int next = v.link;  // the escape node
          const int kind = walk_decide(v, bg, rb, ox, oy, oz);
          if (kind == 2) {
            if (FILL) a.cells[spos + nsoft] = {nd, oid | b | (int)0x80000000u};
            ++nsoft;
            nd = next;
          } else if (kind == 1) {
            if (FILL) a.cells[hpos + nhex] = {nd, oid | b};
            ++nhex;
            nd = next;
          } else if (v.c0 < 0) {
            if (FILL) a.parts[ppos + nruns] = {v.first, oid, v.count};
            ++nruns;
            nsrc += v.count;
            nd = next;
          } else {
            nd = v.c0;
          }
        }
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

#pragma clang fp contract(off)

__global__ void walk_records_kernel(int nn, const double* node64, const int* first,
                                    const int* count, const int* child, const int* esc,
                                    double kgeom, double theta, double thetaMono, WalkRec* rec) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  const double* g = node64 + (size_t)nd * 34;
  WalkRec v;
  v.cx = g[0];
  v.cy = g[1];
  v.cz = g[2];
  const double r = g[3];
  double rad = kgeom * r / theta;
  if (rad < r) rad = r;
  v.rad = rad;
  v.rs = 2.0 * g[4];
  v.rm = kgeom * r / thetaMono;
  v.count = count[nd];
  v.c0 = child[2 * nd];
  v.link = esc ? esc[nd] : child[2 * nd + 1];
  v.first = first[nd];
  rec[nd] = v;
}

#pragma clang fp contract(on)

// Escape nodes. A first child escapes to its sibling; a second child to its
// parent's escape; the root to -1. Chains of second children are resolved
// by pointer jumping: jump[n] = the ancestor whose escape n inherits (-1 once
// esc[n] is final), halving every chain per pass.
__global__ void esc_init_kernel(int nn, const int* child, int* esc, int* jump) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  if (nd == 0) {
    esc[0] = -1;
    jump[0] = -1;
  }
  const int c0 = child[2 * nd], c1 = child[2 * nd + 1];
  if (c0 >= 0) {
    esc[c0] = c1;
    jump[c0] = -1;
    jump[c1] = nd;
  }
}

__global__ void esc_jump_kernel(int nn, const int* esc_in, const int* jump_in, int* esc_out,
                                int* jump_out) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  const int q = jump_in[nd];
  if (q < 0) {
    esc_out[nd] = esc_in[nd];
    jump_out[nd] = -1;
  } else if (jump_in[q] < 0) {
    esc_out[nd] = esc_in[q];
    jump_out[nd] = -1;
  } else {
    esc_out[nd] = esc_in[nd];
    jump_out[nd] = jump_in[q];
  }
}

__global__ void esc_check_kernel(int nn, const int* jump, int* err) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd < nn && jump[nd] >= 0) atomicExch(err, 1);
}

// esc (and its pass buffers esc2, jump, jump2: nn ints each) from the topology
static int walk_escapes(int nn, const int* child, int* esc, int* esc2, int* jump, int* jump2,
                        int* err, hipStream_t s) {
  const dim3 grid((nn + 255) / 256), block(256);
  hipLaunchKernelGGL(esc_init_kernel, grid, block, 0, s, nn, child, esc, jump);
  CGPU_CHECK(hipGetLastError());
  for (int k = 0; k < kEscPasses; ++k) {
    const bool even = (k & 1) == 0;
    hipLaunchKernelGGL(esc_jump_kernel, grid, block, 0, s, nn, even ? esc : esc2,
                       even ? jump : jump2, even ? esc2 : esc, even ? jump2 : jump);
    CGPU_CHECK(hipGetLastError());
  }
  static_assert(kEscPasses % 2 == 0, "the last pass writes esc / jump");
  hipLaunchKernelGGL(esc_check_kernel, grid, block, 0, s, nn, jump, err);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

hipError_t walk_records(int nn, const double* node64, const int* first, const int* count,
                        const int* child, const int* esc, double theta, WalkRec* rec,
                        hipStream_t s) {
  if (nn <= 0) return hipSuccess;
  // theta^4 in the host walk's order
  const double thetaMono = theta * theta * theta * theta;
  hipLaunchKernelGGL(walk_records_kernel, dim3((nn + 255) / 256), dim3(256), 0, s, nn, node64,
                     first, count, child, esc, 2.0 / sqrt(3.0), theta, thetaMono, rec);
  return hipGetLastError();
}

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
// (2 x (nb + 1) ints, 256-byte aligned), the walk records, the escape nodes
// and their pass buffers (4 x nn ints), rocPRIM scan storage
struct WalkWork {
  size_t counts, rec, esc, temp, total;
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
  if (!L || L->nBuckets < 0 || L->numNodes <= 0 ||
      (L->nBuckets > 0 && (!L->bucketNode || !L->first || !L->count ||
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

static WalkWork work_layout(int nb, int nn, size_t scan) {
  WalkWork w;
  w.counts = 256;
  w.rec = w.counts + align256(2 * sizeof(int) * ((size_t)nb + 1));
  w.esc = w.rec + align256(sizeof(WalkRec) * (size_t)nn);
  w.temp = w.esc + align256(4 * sizeof(int) * (size_t)nn);
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
  a.node64 = L->node64;
  const WalkWork W = work_layout(L->nBuckets, L->numNodes, 0);
  a.rec = reinterpret_cast<const WalkRec*>(base + W.rec);
  a.esc = reinterpret_cast<const int*>(base + W.esc);
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

extern "C" size_t changaBuildListsWorkBytes(int nBuckets, int numNodes) {
  if (nBuckets < 0 || numNodes <= 0) return 0;
  const size_t s = scan_bytes(nBuckets);
  return s == 0 ? 0 : work_layout(nBuckets, numNodes, s).total;
}

extern "C" int changaBuildListsCount(const ChangaListBuild* L, long long* numCells,
                                     long long* numParts, void* stream) {
  if (int e = check_args(L, false)) return e;
  hipStream_t s = (hipStream_t)stream;
  const int nb = L->nBuckets;
  if (!L->work || L->workBytes < changaBuildListsWorkBytes(nb, L->numNodes))
    return fail("changaBuildListsCount: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  int* err = static_cast<int*>(L->work);
  const WalkWork W = work_layout(nb, L->numNodes, scan_bytes(nb));
  void* temp = static_cast<char*>(L->work) + W.temp;
  const WalkArgs args = walk_args(L, err);
  hipLaunchKernelGGL(walk_init_kernel, dim3(1), dim3(1), 0, s, nb, args.nCell, args.nPart, err);
  CGPU_CHECK(hipGetLastError());
  {
    // escape nodes first: the records carry them
    const int nn = L->numNodes;
    int* esc = const_cast<int*>(args.esc);
    if (int e = walk_escapes(nn, L->child, esc, esc + nn, esc + 2 * (size_t)nn,
                             esc + 3 * (size_t)nn, err, s))
      return e;
    CGPU_CHECK(walk_records(nn, L->node64, L->first, L->count, L->child, esc, L->theta,
                            const_cast<WalkRec*>(args.rec), s));
  }
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
    return fail("changaBuildListsCount: tree deeper than 256 levels",
                hipErrorInvalidValue, __FILE__, __LINE__);
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
  if (!L->work || L->workBytes < changaBuildListsWorkBytes(nb, L->numNodes))
    return fail("changaBuildListsFill: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (nb > 0) {
    hipLaunchKernelGGL(walk_kernel<true>, dim3((nb + kWalkBlock - 1) / kWalkBlock),
                       dim3(kWalkBlock), 0, s, walk_args(L, static_cast<int*>(L->work)));
    CGPU_CHECK(hipGetLastError());
  }
  return 0;
}
