// Fused walk and force: one persistent launch in which each wavefront walks
// the tree for a target bucket and evaluates what the walk accepts, with no
// interaction list in HBM (the counterpart of the reference's
// gpuLocalTreeWalk, HostCUDA.cu:724-942, with the list semantics of
// ListCompute::doWork, Compute.cpp:690-884).
//
// Walk: wavefront-cooperative over an LDS stack of (node, replica) entries
// packed in 32 bits; each iteration pops up to 64 entries, one per lane, and
// applies the list builder's opening test (openCriterionBucket with the
// softening re-test, walk_rec.h) to their walk records. Accepted cells, softened cells and
// opened buckets (particle runs) are appended, through wavefront prefix
// sums, to the wavefront's own scratch lists (L2-resident); opened internal
// nodes push their two children. The interaction set is the list builder's;
// only its order differs (the walk is breadth-first in chunks of 64).
//
// Force: the bucket's targets then run the cell and particle loops of the
// list path (gravity_body.h) over the scratch lists, precision form, into
// FP64 accumulators and one update of the result array per chunk.
#include <algorithm>

#include "gravity_body.h"
#include "walk_rec.h"

namespace cgpu {

constexpr int kFusedPcTile = 64;
constexpr int kFusedRecTile = 32;  // cell tile with the 128-byte records (their prefetch fits
                                   // the four-wavefront register budget)
constexpr int kFusedPpTile = 256;
constexpr int kFusedRanges = kNumXcd;
constexpr int kFusedCounterStride = 32;

struct FusedArgs {
  ChangaGravityLaunch pc, pp;  // lists: the bases of the wavefronts' scratch
  const int* bucketNode;
  const double* node64;        // the target bucket's geometry
  const WalkRec* rec;          // walk records of every node
  double period;
  int R;
  ILCell* cells;  // per wave: maxCells entries
  ILPart* runs;   // per wave: maxRuns entries
  int maxCells, maxRuns;
  int* counters;  // kFusedRanges claim counters, kFusedCounterStride ints apart; then err
  int* err;
};

template <int PCT, int PPT>
struct FusedLds {
  static constexpr size_t stack = kCoopStack * sizeof(unsigned);
  static constexpr size_t pc = (size_t)(PCT + 1) * kFwdRec * sizeof(float);
  static constexpr size_t pp = sizeof(PpFwdLds<PPT>);
  static constexpr size_t red = reduce_bytes<1>();
  static constexpr size_t m1 = stack > pc ? stack : pc;
  static constexpr size_t m2 = pp > red ? pp : red;
  static constexpr size_t bytes = m1 > m2 ? m1 : m2;
  static constexpr size_t f4 = (bytes + 15) / 16;
};

// Walks bucket b into the wave's scratch: hexadecapole cells at [0, nhex),
// softened cells at [maxCells - nsoft, maxCells), runs at [0, nruns).
// Returns false (and raises err) when a capacity is exceeded. Softened cells
// count twice against maxCells, so that the two regions never overlap and
// the move behind the hexadecapole cells reads no slot it has written.
__device__ bool fused_walk(const FusedArgs& a, int b, int lane, unsigned* stack, ILCell* cells,
                           ILPart* runs, int& nhex, int& nsoft, int& nruns) {
  double bg[12];
  {
    const double* g = a.node64 + (size_t)a.bucketNode[b] * 34;
#pragma unroll
    for (int k = 0; k < 12; ++k) bg[k] = g[k];
  }
  const int maxCells = a.maxCells, maxRuns = a.maxRuns;
  const bool ok = coop_walk(
      a.rec, bg, 2.0 * bg[4], a.R, a.period, lane, stack, nhex, nsoft, nruns,
      [&](int h, int s, int r) { return h + 2 * s <= maxCells && r <= maxRuns; },
      [&](int kind, int nd, int code, const WalkRec& v, int slot) {
        const int oid = code << 22;
        if (kind == 1)
          cells[slot] = {nd, oid | b};
        else if (kind == 2)
          cells[maxCells - 1 - slot] = {nd, oid | b | (int)0x80000000u};
        else
          runs[slot] = {v.first, oid, v.count};
      });
  if (!ok && lane == 0) atomicExch(a.err, 1);
  return ok;
}

// Four wavefronts per SIMD (128 VGPRs, a few spills outside the force loops)
// measured 11-12% faster than the three that 141 VGPRs allow: the walk's load
// latency is hidden behind more resident wavefronts.
// REC: cells from the 128-byte cell records (a.pc.cellRecords), whole-line
// fetches in tiles of PCT.
template <int PCT, int PPT, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void fused_kernel(
    FusedArgs a) {
  __shared__ float4 lds[FusedLds<PCT, PPT>::f4];
  unsigned* stack = reinterpret_cast<unsigned*>(lds);
  PpFwdLds<PPT>& ps = *reinterpret_cast<PpFwdLds<PPT>*>(lds);
  const int lane = threadIdx.x;
  // this wavefront's scratch: a.pc.list / a.pp.list are the scratch bases
  const int cbase = blockIdx.x * a.maxCells, rbase = blockIdx.x * a.maxRuns;
  ILCell* cells = a.cells + cbase;
  ILPart* runs = a.runs + rbase;
  const int nb = a.pc.numBuckets;
  int q = blockIdx.x % kFusedRanges;
  for (int tried = 0; tried < kFusedRanges;) {
    const int r0 = (int)((long long)nb * q / kFusedRanges);
    const int r1 = (int)((long long)nb * (q + 1) / kFusedRanges);
    int k = 0;
    if (lane == 0) k = atomicAdd(a.counters + q * kFusedCounterStride, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    if (r0 + k >= r1) {
      ++tried;
      q = q + 1 == kFusedRanges ? 0 : q + 1;
      continue;
    }
    const int b = r0 + k;
    int nhex, nsoft, nruns;
    if (!fused_walk(a, b, lane, stack, cells, runs, nhex, nsoft, nruns)) return;
    // softened cells (filled from the end of the region) moved behind the
    // hexadecapole cells, so one cell loop covers both
    if (nsoft > 0) {
      for (int i = lane; i < nsoft; i += kWave) cells[nhex + i] = cells[a.maxCells - 1 - i];
      __syncthreads();
    }
    const int bstart = a.pc.starts[b], n = a.pc.sizes[b];
    double cx0, cy0, cz0;
    bucket_centre<true>(a.pc.tpos64, bstart, n, cx0, cy0, cz0);
    for (int t0 = 0; t0 < n; t0 += kWave) {
      // one target per lane: the walk's stack and scratch leave no registers
      // for more
      const Lanes l = lanes_for<1>(lane, min(kWave, n - t0));
      Target t[1];
      load_targets<true, 1>(a.pc, bstart + t0, l, cx0, cy0, cz0, t);
      Acc64 acc[1] = {{0.0, 0.0, 0.0, 0.0, 0.0f}};
      // the scratch is re-read per chunk and rewritten per bucket: plain loads
      // keep it in L2
      pc_cells<true, PCT, false, 1, REC>(a.pc, cbase, cbase + nhex + nsoft, cx0, cy0, cz0, t, l,
                                         lds, acc);
      pp_sources<true, 1, PPT, false, 1>(a.pp, rbase, rbase + nruns, cx0, cy0, cz0, t, l, ps,
                                         acc);
      reduce_store<1>(lds, l, acc, a.pc.vars + bstart + t0);
    }
  }
}

__global__ void fused_reset_kernel(int* counters) {
  const int i = threadIdx.x;
  if (i < kFusedRanges * kFusedCounterStride + 1) counters[i] = 0;  // counters, then err
}

using FusedFn = void (*)(FusedArgs);

static FusedFn fused_fn(bool rec) {
  return rec ? fused_kernel<kFusedRecTile, kFusedPpTile, true>
             : fused_kernel<kFusedPcTile, kFusedPpTile, false>;
}

static int fused_grid(int* grid, bool rec) {
  int dev = 0, ncu = 0, per = 0;
  CGPU_CHECK(hipGetDevice(&dev));
  CGPU_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  CGPU_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)fused_fn(rec),
                                                          kWave, 0));
  *grid = ncu * (per > 0 ? per : 1);
  return 0;
}

// work buffer: claim counters and error flag (256-byte aligned head), the
// walk records, then the per-wavefront cell and run scratch
struct FusedWork {
  size_t rec, cells, runs, total;
};

static FusedWork fused_layout(int grid, int numNodes, int maxCells, int maxRuns) {
  FusedWork w;
  w.rec = 256 * ((sizeof(int) * (kFusedRanges * kFusedCounterStride + 1) + 255) / 256);
  w.cells = w.rec + 256 * (((size_t)numNodes * sizeof(WalkRec) + 255) / 256);
  w.runs = w.cells + 256 * (((size_t)grid * maxCells * sizeof(ILCell) + 255) / 256);
  w.total = w.runs + (size_t)grid * maxRuns * sizeof(ILPart);
  return w;
}

}  // namespace cgpu

using namespace cgpu;

extern "C" size_t changaFusedGravityWorkBytes(int numNodes, int maxCells, int maxRuns) {
  // enough for either cell source (the work buffer does not know which runs)
  int grid = 0, grid_rec = 0;
  if (numNodes <= 0 || maxCells <= 0 || maxRuns <= 0 || fused_grid(&grid, false) ||
      fused_grid(&grid_rec, true))
    return 0;
  return fused_layout(std::max(grid, grid_rec), numNodes, maxCells, maxRuns).total;
}

extern "C" int changaFusedGravity(const ChangaFusedLaunch* f, void* stream) {
  if (f && f->cells.numBuckets == 0) return 0;  // an empty rung
  if (!f || !f->bucketNode || !f->first || !f->count || !f->child || !f->node64 || !f->work)
    return fail("changaFusedGravity: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  const ChangaGravityLaunch& pc = f->cells;
  const ChangaGravityLaunch& pp = f->parts;
  if (!pc.targets || !pc.vars || !pc.moments || !pc.starts || !pc.sizes || !pc.tpos64 ||
      !pc.cm64 || !pp.sources || !pp.spos64 || pp.targets != pc.targets || pp.vars != pc.vars ||
      pp.starts != pc.starts || pp.sizes != pc.sizes || pp.numBuckets != pc.numBuckets)
    return fail("changaFusedGravity: cell/particle launches (precision form, same buckets)",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (f->nReplicas < 0 || f->nReplicas > 3 || !(f->theta > 0.0) ||
      (f->nReplicas > 0 && !(f->period > 0.0)) || f->maxCells <= 0 || f->maxRuns <= 0)
    return fail("changaFusedGravity: theta, nReplicas (0..3), period, capacities",
                hipErrorInvalidValue, __FILE__, __LINE__);
  const int W = 2 * f->nReplicas + 1;
  if (W * W * W > kCoopStack)
    return fail("changaFusedGravity: replicas exceed the walk stack", hipErrorInvalidValue,
                __FILE__, __LINE__);
  if (f->numNodes <= 0 || f->numNodes > (1 << kCoopNodeBits))
    return fail("changaFusedGravity: numNodes in 1..2^23", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (reinterpret_cast<uintptr_t>(pc.cellRecords) % CHANGA_CELL_RECORD_BYTES != 0)
    return fail("changaFusedGravity: cellRecords not 128-byte aligned", hipErrorInvalidValue,
                __FILE__, __LINE__);
  const bool rec = pc.cellRecords != nullptr;
  int grid = 0;
  if (int e = fused_grid(&grid, rec)) return e;
  if ((long long)grid * f->maxCells > 0x7fffffffll || (long long)grid * f->maxRuns > 0x7fffffffll)
    return fail("changaFusedGravity: scratch capacities exceed the int32 index range",
                hipErrorInvalidValue, __FILE__, __LINE__);
  const FusedWork L = fused_layout(grid, f->numNodes, f->maxCells, f->maxRuns);
  if (f->workBytes < L.total)
    return fail("changaFusedGravity: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (pc.numBuckets <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  char* wb = static_cast<char*>(f->work);
  FusedArgs a;
  a.pc = pc;
  a.pp = pp;
  a.bucketNode = f->bucketNode;
  a.node64 = f->node64;
  a.rec = reinterpret_cast<const WalkRec*>(wb + L.rec);
  a.period = f->nReplicas > 0 ? f->period : 0.0;
  a.R = f->nReplicas;
  a.counters = reinterpret_cast<int*>(wb);
  a.err = a.counters + kFusedRanges * kFusedCounterStride;
  a.cells = reinterpret_cast<ILCell*>(wb + L.cells);
  a.runs = reinterpret_cast<ILPart*>(wb + L.runs);
  a.maxCells = f->maxCells;
  a.maxRuns = f->maxRuns;
  a.pc.list = a.cells;
  a.pc.listKind = 0;
  a.pp.list = a.runs;
  a.pp.listKind = 1;
  hipLaunchKernelGGL(fused_reset_kernel, dim3(1), dim3(kFusedRanges * kFusedCounterStride + 1), 0,
                     s, a.counters);
  CGPU_CHECK(hipGetLastError());
  CGPU_CHECK(walk_records(f->numNodes, f->node64, f->first, f->count, f->child, nullptr,
                          f->theta, const_cast<WalkRec*>(a.rec), s));
  hipLaunchKernelGGL(fused_fn(rec), dim3(grid), dim3(kWave), 0, s, a);
  CGPU_CHECK(hipGetLastError());
  if (f->errorFlag) {
    CGPU_CHECK(hipMemcpyAsync(f->errorFlag, a.err, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  return 0;
}
