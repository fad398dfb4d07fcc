// Batched form: every bucket's particle-cell and particle-particle lists in
// one persistent launch, replayed as a hipGraph.
//
// The per-bucket kernels (gravity.hip) follow ChaNGa's one-launch-per-list
// pattern (nodeGravityComputation / particleGravityComputation,
// HostCUDA.cu:1008-1205, 1565-1751). Here one wavefront takes a bucket from a
// work queue, runs its cell list and then its particle list into the same
// FP64 accumulators (the chunk loops of gravity_body.h), and updates the
// bucket's VariablePartData once.
//
// Work queues. Buckets are cut into 8 contiguous ranges of equal estimated
// cost, one per XCD, so that the waves of one XCD (workgroup id mod 8) share
// source cells in their L2. A wave claims `grab` buckets of its XCD's range
// per atomic; when the range is exhausted it moves on to the next XCD's, so
// no XCD idles while another has work. Within a range buckets run either in
// tree order or longest list first.
//
// Plan. changaBatchPlanCreate estimates each bucket's cost on the GPU, builds
// the queues, times the candidate configurations (cell tile, particle tile,
// buckets per claim, order) on the caller's data with a scratch result array,
// and captures the chosen one (queue reset, optional zeroing, persistent
// kernel) into a graph. Every node is a kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "gravity_body.h"

struct ChangaBatchPlan;

namespace cgpu {

constexpr int kQueueHdr = 16;       // ints before the bucket ids: range starts [0..8]
constexpr int kCounterStride = 32;  // queue counters 128 B apart
constexpr int kMaxGrab = 64;

struct BatchArgs {
  ChangaGravityLaunch pc, pp;
  const int* queue;  // [0..kNumXcd]: range starts, then bucket ids from kQueueHdr
  int* counters;     // kNumXcd counters, kCounterStride ints apart
  int grab;          // buckets claimed per atomic
};

template <int PCT, int PPT>
struct BatchLds {
  static constexpr size_t pc = (size_t)(PCT + 1) * kFwdRec * sizeof(float);
  static constexpr size_t pp = sizeof(PpFwdLds<PPT>);
  static constexpr size_t red = reduce_bytes<3>();
  static constexpr size_t bytes = pc > pp ? (pc > red ? pc : red) : (pp > red ? pp : red);
  static constexpr size_t f4 = (bytes + 15) / 16;
};

// One bucket: cells, then particles, then one update of its results. Every
// phase that uses LDS ends with a barrier, so the phases share one buffer.
// KMAX 1: one target per lane; 3: one to three, chosen per chunk by its size
// (targets_per_lane). REC: cells from the 128-byte cell records.
template <bool PRECISE, int KIND, int PCT, int PPT, int KMAX, bool REC>
__device__ __forceinline__ void batch_bucket(const BatchArgs& a, int b, int lane, float4* lds) {
  const int pcs = a.pc.marks[b], pce = a.pc.marks[b + 1];
  const int pps = a.pp.marks[b], ppe = a.pp.marks[b + 1];
  const int bstart = a.pc.starts[b], n = a.pc.sizes[b];
  double cx0, cy0, cz0;
  bucket_centre<PRECISE>(a.pc.tpos64, bstart, n, cx0, cy0, cz0);
  PpFwdLds<PPT>& s = *reinterpret_cast<PpFwdLds<PPT>*>(lds);
  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    auto chunk = [&](auto kt) {
      constexpr int K = decltype(kt)::value;
      const Lanes l = lanes_for<K>(lane, nt);
      Target t[K];
      load_targets<PRECISE, K>(a.pc, bstart + t0, l, cx0, cy0, cz0, t);
      Acc64 acc[K];
#pragma unroll
      for (int j = 0; j < K; ++j) acc[j] = {0.0, 0.0, 0.0, 0.0, 0.0f};
      pc_cells<PRECISE, PCT, true, K, REC>(a.pc, pcs, pce, cx0, cy0, cz0, t, l, lds, acc);
      pp_sources<PRECISE, KIND, PPT, true, K>(a.pp, pps, ppe, cx0, cy0, cz0, t, l, s, acc);
      reduce_store<K>(lds, l, acc, a.pc.vars + bstart + t0);
    };
    if constexpr (KMAX == 1) chunk(KTag<1>());
    else with_targets_per_lane(nt, chunk);
  }
}

// Persistent kernel: one wavefront per workgroup; runs until every queue is
// exhausted. Each wave visits the 8 queues once, starting at its own XCD's,
// and stays on a queue until a claim comes back past its end, so every wave
// exits after finitely many claims and no wave waits on another.
// Three wavefronts per SIMD: up to 168 VGPRs for the three-targets-per-lane
// loops (measured faster on the 2^22 halos than one target per lane at four
// wavefronts). With KMAX 1 the kernel holds one chunk loop instead of three
// and needs 139-160 VGPRs, without spills; on a sparse rung (one to three
// targets in most buckets) it measured 16% faster than KMAX 3, on the dense
// halos 7% slower. With KMAX 1 and cell records it is built for four
// wavefronts (128 VGPRs: the 32-cell tile spills 2, the 64-cell tile 23),
// which the sparse rung prefers.
template <bool PRECISE, int KIND, int PCT, int PPT, int KMAX, bool REC>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(KMAX == 1 && REC ? 4 : 3))) void batch_kernel(BatchArgs a) {
  __shared__ float4 lds[BatchLds<PCT, PPT>::f4];
  const int lane = threadIdx.x;
  const int* qbeg = a.queue;
  const int* ids = a.queue + kQueueHdr;
  int q = blockIdx.x % kNumXcd;
  for (int tried = 0; tried < kNumXcd;) {
    int k = 0;
    if (lane == 0) k = atomicAdd(a.counters + q * kCounterStride, a.grab);
    k = __builtin_amdgcn_readfirstlane(k);
    const int q0 = qbeg[q], qn = qbeg[q + 1] - q0;
    if (k >= qn) {
      ++tried;
      q = q + 1 == kNumXcd ? 0 : q + 1;
      continue;
    }
    const int kend = min(k + a.grab, qn);
    for (; k < kend; ++k)
      batch_bucket<PRECISE, KIND, PCT, PPT, KMAX, REC>(a, ids[q0 + k], lane, lds);
  }
}

__global__ void batch_reset_kernel(int* counters) {
  const int i = threadIdx.x;
  if (i < kNumXcd * kCounterStride) counters[i] = 0;
}

// Estimated wavefront time of each bucket (VALU instructions per slot
// iteration: ~130 for a cell, ~25 for a particle, plus per-chunk overhead).
template <int KIND>
__global__ void bucket_cost_kernel(ChangaGravityLaunch pc, ChangaGravityLaunch pp, float* cost) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= pc.numBuckets) return;
  const int nc = pc.marks[b + 1] - pc.marks[b];
  float np = 0.0f;
  if (KIND == 0) {
    np = (float)(pp.marks[b + 1] - pp.marks[b]);
  } else {
    const ILPart* r = static_cast<const ILPart*>(pp.list);
    for (int i = pp.marks[b]; i < pp.marks[b + 1]; ++i) np += (float)r[i].num;
  }
  const int n = pc.sizes[b];
  float c = 0.0f;
  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    c += (130.0f * nc + 25.0f * np) / (float)(kWave / nt) + 400.0f;
  }
  cost[b] = c;
}

using BatchFn = void (*)(BatchArgs);

template <bool PR, int KIND, int KM, bool REC>
static BatchFn pick_tiles(int pct, int ppt) {
  if (pct == 32)
    return ppt == 128 ? batch_kernel<PR, KIND, 32, 128, KM, REC>
                      : batch_kernel<PR, KIND, 32, 256, KM, REC>;
  return ppt == 128 ? batch_kernel<PR, KIND, 64, 128, KM, REC>
                    : batch_kernel<PR, KIND, 64, 256, KM, REC>;
}

template <bool PR, int KIND, bool REC>
static BatchFn pick_kmax(int pct, int ppt, int kmax) {
  return kmax == 1 ? pick_tiles<PR, KIND, 1, REC>(pct, ppt) : pick_tiles<PR, KIND, 3, REC>(pct, ppt);
}

template <bool REC>
static BatchFn pick_rec(bool pr, int kind, const ChangaBatchOptions& o) {
  const int pct = o.pcTile, ppt = o.ppTile, km = o.targetsPerLane;
  if (pr)
    return kind ? pick_kmax<true, 1, REC>(pct, ppt, km) : pick_kmax<true, 0, REC>(pct, ppt, km);
  return kind ? pick_kmax<false, 1, REC>(pct, ppt, km) : pick_kmax<false, 0, REC>(pct, ppt, km);
}

// the kernel for these options; with cell records when the launch has them
static BatchFn pick(bool pr, int kind, const ChangaBatchOptions& o, bool rec) {
  return rec ? pick_rec<true>(pr, kind, o) : pick_rec<false>(pr, kind, o);
}

}  // namespace cgpu

using namespace cgpu;

struct ChangaBatchPlan {
  BatchArgs args;
  bool precise;
  int kind;
  ChangaBatchOptions opt;
  int grid;
  int nParticles;
  int* d_queue;
  int* d_counters;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace {

// Resident workgroups of the persistent kernel on the whole device; -1 on error.
int grid_for(BatchFn fn) {
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)fn, kWave, 0) != hipSuccess)
    return -1;
  return ncu * std::max(per, 1);
}

// Enqueue one batched evaluation with the plan's chosen configuration.
int enqueue(const ChangaBatchPlan* p, const BatchArgs& args, const ChangaBatchOptions& o, int grid,
            hipStream_t s) {
  if (p->args.pc.numBuckets == 0) return 0;
  if (o.zeroVars) {
    if (int e = changaZeroVars(args.pc.vars, p->nParticles, s)) return e;
  }
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(kNumXcd * kCounterStride), 0, s,
                     args.counters);
  CGPU_CHECK(hipGetLastError());
  BatchFn fn = pick(p->precise, p->kind, o, o.cellSource == 1);
  BatchArgs a = args;
  a.grab = o.listsPerGrab;
  void* kargs[] = {&a};
  CGPU_CHECK(hipLaunchKernel((const void*)fn, dim3(grid), dim3(kWave), kargs, 0, s));
  return 0;
}

bool same_buckets(const ChangaGravityLaunch& a, const ChangaGravityLaunch& b) {
  return a.targets == b.targets && a.vars == b.vars && a.starts == b.starts &&
         a.sizes == b.sizes && a.numBuckets == b.numBuckets && a.fperiod == b.fperiod &&
         a.tpos64 == b.tpos64;
}

// Queue layout for one order: range starts, then bucket ids.
std::vector<int> make_queue(const std::vector<float>& cost, bool longest_first) {
  const int nb = (int)cost.size();
  std::vector<int> q(kQueueHdr + nb, 0);
  double total = 0.0;
  for (float c : cost) total += c;
  int x = 1;
  double run = 0.0;
  for (int b = 0; b < nb; ++b) {
    while (x < kNumXcd && run >= total * x / kNumXcd) q[x++] = b;
    run += cost[b];
  }
  while (x < kNumXcd) q[x++] = nb;
  q[kNumXcd] = nb;
  int* ids = q.data() + kQueueHdr;
  std::iota(ids, ids + nb, 0);
  if (longest_first)
    for (int r = 0; r < kNumXcd; ++r)
      std::stable_sort(ids + q[r], ids + q[r + 1],
                       [&](int u, int v) { return cost[u] > cost[v]; });
  return q;
}

void release(ChangaBatchPlan* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_queue) (void)hipFree(p->d_queue);
  if (p->d_counters) (void)hipFree(p->d_counters);
  delete p;
}

}  // namespace

extern "C" int changaBatchPlanCreate(const ChangaGravityLaunch* cells,
                                     const ChangaGravityLaunch* parts, ChangaBatchOptions* opts,
                                     ChangaBatchPlan** plan, void* stream) {
  if (!cells || !parts || !opts || !plan)
    return fail("changaBatchPlanCreate: missing argument", hipErrorInvalidValue, __FILE__,
                __LINE__);
  *plan = nullptr;
  const ChangaGravityLaunch& pc = *cells;
  const ChangaGravityLaunch& pp = *parts;
  // a plan over no buckets (an empty rung) needs no arrays and runs nothing
  if (pc.numBuckets < 0 ||
      (pc.numBuckets > 0 && (!pc.marks || !pc.starts || !pc.sizes || !pc.vars || !pc.targets ||
                             !pc.moments || !pc.list || !pp.marks || !pp.list || !pp.sources)))
    return fail("changaBatchPlanCreate: missing array", hipErrorInvalidValue, __FILE__, __LINE__);
  if (!same_buckets(pc, pp))
    return fail("changaBatchPlanCreate: cell and particle launches differ in their buckets",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (pc.tpos64 && (!pc.cm64 || !pp.spos64))
    return fail("changaBatchPlanCreate: precision form needs cm64 and spos64",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (pp.listKind != 0 && pp.listKind != 1)
    return fail("changaBatchPlanCreate: listKind", hipErrorInvalidValue, __FILE__, __LINE__);
  if (reinterpret_cast<uintptr_t>(pc.cellRecords) % CHANGA_CELL_RECORD_BYTES != 0)
    return fail("changaBatchPlanCreate: cellRecords not 128-byte aligned", hipErrorInvalidValue,
                __FILE__, __LINE__);
  ChangaBatchOptions o = *opts;
  if ((o.pcTile != 0 && o.pcTile != 32 && o.pcTile != 64) ||
      (o.ppTile != 0 && o.ppTile != 128 && o.ppTile != 256) || o.listsPerGrab < 0 ||
      o.listsPerGrab > kMaxGrab || o.order < 0 || o.order > 2 ||
      (o.targetsPerLane != 0 && o.targetsPerLane != 1 && o.targetsPerLane != 3) ||
      o.cellSource < 0 || o.cellSource > 2)
    return fail("changaBatchPlanCreate: option out of range", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (o.cellSource == 1 && !pc.cellRecords)
    return fail("changaBatchPlanCreate: cellSource 1 needs cellRecords", hipErrorInvalidValue,
                __FILE__, __LINE__);
  // without records only the wire moments can serve
  if (!pc.cellRecords) o.cellSource = 2;
  hipStream_t s = (hipStream_t)stream;
  ChangaBatchPlan* p = new ChangaBatchPlan();
  std::memset(p, 0, sizeof(*p));
  p->precise = pc.tpos64 != nullptr;
  p->kind = pp.listKind;
  p->args.pc = pc;
  p->args.pp = pp;
  const int nb = pc.numBuckets;

  auto bail = [&](int e) {
    release(p);
    return e;
  };
  // bucket extents -> particle count; per-bucket cost
  std::vector<int> starts(nb), sizes(nb);
  std::vector<float> cost(nb);
  float* d_cost = nullptr;
  if (nb > 0) {
    hipError_t e = hipMalloc(&d_cost, sizeof(float) * nb);
    if (e != hipSuccess) return bail(fail("hipMalloc cost", e, __FILE__, __LINE__));
    if (p->kind)
      hipLaunchKernelGGL(bucket_cost_kernel<1>, dim3((nb + 255) / 256), dim3(256), 0, s, pc, pp,
                         d_cost);
    else
      hipLaunchKernelGGL(bucket_cost_kernel<0>, dim3((nb + 255) / 256), dim3(256), 0, s, pc, pp,
                         d_cost);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cost.data(), d_cost, sizeof(float) * nb,
                                            hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(starts.data(), pc.starts, sizeof(int) * nb,
                                            hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(sizes.data(), pc.sizes, sizeof(int) * nb,
                                            hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_cost);
    if (e != hipSuccess) return bail(fail("batch plan: cost estimate", e, __FILE__, __LINE__));
  }
  for (int b = 0; b < nb; ++b) p->nParticles = std::max(p->nParticles, starts[b] + sizes[b]);

  // queues for both orders (index 0: tree order, 1: longest first)
  const size_t qbytes = sizeof(int) * (kQueueHdr + (size_t)nb);
  int* d_q[2] = {nullptr, nullptr};
  for (int k = 0; k < 2; ++k) {
    std::vector<int> q = make_queue(cost, k == 1);
    hipError_t e = hipMalloc(&d_q[k], qbytes);
    if (e == hipSuccess) e = hipMemcpy(d_q[k], q.data(), qbytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      for (int j = 0; j < 2; ++j)
        if (d_q[j]) (void)hipFree(d_q[j]);
      return bail(fail("batch plan: queue upload", e, __FILE__, __LINE__));
    }
  }
  {
    hipError_t e = hipMalloc(&p->d_counters, sizeof(int) * kNumXcd * kCounterStride);
    if (e != hipSuccess) {
      (void)hipFree(d_q[0]);
      (void)hipFree(d_q[1]);
      return bail(fail("hipMalloc counters", e, __FILE__, __LINE__));
    }
  }
  p->args.counters = p->d_counters;

  // autotune whatever the caller left at 0, on a scratch result array
  const bool tune = o.pcTile == 0 || o.ppTile == 0 || o.listsPerGrab == 0 || o.order == 0 ||
                    o.targetsPerLane == 0 || o.cellSource == 0;
  o.tunedMs = 0.0f;
  if (tune && nb > 0) {
    VariablePartData* scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&scratch, sizeof(VariablePartData) * (size_t)std::max(p->nParticles, 1));
    if (e == hipSuccess)
      e = hipMemsetAsync(scratch, 0, sizeof(VariablePartData) * (size_t)std::max(p->nParticles, 1), s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // CHANGA_BATCH_TUNE_LOG set: each candidate's time on stderr
    const bool tune_log = std::getenv("CHANGA_BATCH_TUNE_LOG") != nullptr;
    // mean time of `reps` runs of candidate c (after one untimed run), or -1
    auto time_config = [&](const ChangaBatchOptions& c, int reps) -> float {
      BatchArgs a = p->args;
      a.pc.vars = scratch;
      a.pp.vars = scratch;
      a.queue = d_q[c.order - 1];
      const int grid = grid_for(pick(p->precise, p->kind, c, c.cellSource == 1));
      if (grid <= 0) {
        e = hipErrorInvalidValue;
        return -1.0f;
      }
      if (enqueue(p, a, c, grid, s)) {
        e = hipErrorLaunchFailure;
        return -1.0f;
      }
      e = hipEventRecord(e0, s);
      for (int r = 0; r < reps && e == hipSuccess; ++r)
        if (enqueue(p, a, c, grid, s)) e = hipErrorLaunchFailure;
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) return -1.0f;
      ms /= (float)reps;
      if (tune_log)
        std::fprintf(stderr,
                     "[batch plan] cell tile %d, particle tile %d, %d per claim, %s, targets "
                     "per lane %d, %s, %d runs: %.4f ms\n",
                     c.pcTile, c.ppTile, c.listsPerGrab,
                     c.order == 1 ? "tree order" : "costliest first", c.targetsPerLane,
                     c.cellSource == 1 ? "cell records" : "wire moments", reps, ms);
      return ms;
    };
    // Every candidate is timed over 2 runs; the 4 fastest and the default
    // configuration (cell tile 64, particle tile 256, one bucket per claim,
    // tree order, up to three targets per lane, wire moments) are timed again
    // over 6 runs, and the default is kept unless another is more than 2%
    // faster, so that variants of near-equal time do not flip from one plan
    // to the next. The default runs twice, untimed, before the sweep.
    const int pcts[2] = {64, 32}, ppts[2] = {256, 128}, grabs[2] = {1, 4}, orders[2] = {1, 2},
              kmaxs[2] = {3, 1}, sources[2] = {2, 1};
    std::vector<std::pair<float, ChangaBatchOptions>> swept;
    ChangaBatchOptions dflt = o;
    for (int i = 0; i < 64 && e == hipSuccess; ++i) {
      const int i0 = i & 1, i1 = (i >> 1) & 1, i2 = (i >> 2) & 1, i3 = (i >> 3) & 1,
                i4 = (i >> 4) & 1, i5 = i >> 5;
      // skip duplicates of a fixed option
      if ((o.pcTile && i0) || (o.ppTile && i1) || (o.listsPerGrab && i2) || (o.order && i3) ||
          (o.targetsPerLane && i4) || (o.cellSource && i5))
        continue;
      ChangaBatchOptions c = o;
      c.pcTile = o.pcTile ? o.pcTile : pcts[i0];
      c.ppTile = o.ppTile ? o.ppTile : ppts[i1];
      c.listsPerGrab = o.listsPerGrab ? o.listsPerGrab : grabs[i2];
      c.order = o.order ? o.order : orders[i3];
      c.targetsPerLane = o.targetsPerLane ? o.targetsPerLane : kmaxs[i4];
      c.cellSource = o.cellSource ? o.cellSource : sources[i5];
      c.zeroVars = 0;
      if (i == 0 || swept.empty()) {
        dflt = c;
        time_config(c, 1);  // warm-up
      }
      const float ms = time_config(c, 2);
      if (ms >= 0.0f) swept.emplace_back(ms, c);
    }
    float best = -1.0f;
    ChangaBatchOptions bo = dflt;
    if (e == hipSuccess && !swept.empty()) {
      std::stable_sort(swept.begin(), swept.end(),
                       [](const auto& u, const auto& v) { return u.first < v.first; });
      const float dms = time_config(dflt, 6);
      best = dms;
      for (size_t k = 0; k < swept.size() && k < 4 && e == hipSuccess; ++k) {
        const ChangaBatchOptions& c = swept[k].second;
        if (std::memcmp(&c, &dflt, sizeof(c)) == 0) continue;
        const float ms = time_config(c, 6);
        if (ms >= 0.0f && ms < best && ms < 0.98f * dms) {
          best = ms;
          bo = c;
        }
      }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (scratch) (void)hipFree(scratch);
    if (e != hipSuccess) {
      (void)hipFree(d_q[0]);
      (void)hipFree(d_q[1]);
      return bail(fail("batch plan: autotune", e, __FILE__, __LINE__));
    }
    bo.zeroVars = o.zeroVars;
    bo.tunedMs = best;
    o = bo;
  } else {
    if (!o.pcTile) o.pcTile = 64;
    if (!o.ppTile) o.ppTile = 256;
    if (!o.listsPerGrab) o.listsPerGrab = 1;
    if (!o.order) o.order = 1;
    if (!o.targetsPerLane) o.targetsPerLane = 3;
    if (!o.cellSource) o.cellSource = 2;
  }
  p->opt = o;
  p->d_queue = d_q[o.order - 1];
  (void)hipFree(d_q[2 - o.order]);
  p->args.queue = p->d_queue;
  p->args.grab = o.listsPerGrab;
  if (nb > 0) {
    p->grid = grid_for(pick(p->precise, p->kind, o, o.cellSource == 1));
    if (p->grid <= 0) return bail(fail("batch plan: occupancy", hipErrorInvalidValue, __FILE__,
                                       __LINE__));
    // capture reset (+ zeroing) + persistent kernel; every node is a kernel
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      int le = enqueue(p, p->args, p->opt, p->grid, cs);
      hipError_t ee = hipStreamEndCapture(cs, &p->graph);
      e = le ? hipErrorLaunchFailure : ee;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return bail(fail("batch plan: graph capture", e, __FILE__, __LINE__));
  }
  *opts = o;
  *plan = p;
  return 0;
}

extern "C" int changaBatchPlanLaunch(ChangaBatchPlan* plan, void* stream) {
  if (!plan) return fail("changaBatchPlanLaunch: no plan", hipErrorInvalidValue, __FILE__, __LINE__);
  return enqueue(plan, plan->args, plan->opt, plan->grid, (hipStream_t)stream);
}

extern "C" int changaBatchPlanRun(ChangaBatchPlan* plan, void* stream) {
  if (!plan) return fail("changaBatchPlanRun: no plan", hipErrorInvalidValue, __FILE__, __LINE__);
  if (!plan->exec) return 0;
  CGPU_CHECK(hipGraphLaunch(plan->exec, (hipStream_t)stream));
  return 0;
}

extern "C" int changaBatchPlanDestroy(ChangaBatchPlan* plan) {
  release(plan);
  return 0;
}

// ---------------------------------------------------------------------------
// Batched evaluation without allocation. The queue (range starts and bucket
// ids in tree order), its counters and the partial sums of the range split
// live in one file-scope device array; host code reaches it only through the
// address hipGetSymbolAddress returns for the current device. Every call on
// a GPU rewrites what it reads of the array before reading it.
// ---------------------------------------------------------------------------
namespace cgpu {

constexpr int kNaChunk = 4096;  // buckets per partial cost sum
constexpr int kNaMaxChunks = (CHANGA_NOALLOC_MAX_BUCKETS + kNaChunk - 1) / kNaChunk;
// [counters][partial sums (float)][queue header][bucket ids]
constexpr size_t kNaCounters = 0;
constexpr size_t kNaPartials = kNumXcd * kCounterStride;
constexpr size_t kNaQueue = kNaPartials + kNaMaxChunks;
constexpr size_t kNaInts = kNaQueue + kQueueHdr + CHANGA_NOALLOC_MAX_BUCKETS;
static_assert(kNaQueue % 4 == 0, "queue aligned to 16 bytes");

static __device__ __align__(256) int scratch[kNaInts];

// Per-bucket cost (the estimate of bucket_cost_kernel), stored as float bits
// in the bucket-id region until the range split has read it.
template <int KIND>
__global__ void noalloc_cost_kernel(ChangaGravityLaunch pc, ChangaGravityLaunch pp, int* ws) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= pc.numBuckets) return;
  const int nc = pc.marks[b + 1] - pc.marks[b];
  float np = 0.0f;
  if (KIND == 0) {
    np = (float)(pp.marks[b + 1] - pp.marks[b]);
  } else {
    const ILPart* r = static_cast<const ILPart*>(pp.list);
    for (int i = pp.marks[b]; i < pp.marks[b + 1]; ++i) np += (float)r[i].num;
  }
  const int n = pc.sizes[b];
  float c = 0.0f;
  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    c += (130.0f * nc + 25.0f * np) / (float)(kWave / nt) + 400.0f;
  }
  reinterpret_cast<float*>(ws + kNaQueue + kQueueHdr)[b] = c;
}

// Sum of each chunk of kNaChunk costs (one 256-thread block per chunk).
__global__ __launch_bounds__(256) void noalloc_chunk_sum_kernel(int nb, int* ws) {
  __shared__ float part[256 / kWave];
  const float* cost = reinterpret_cast<const float*>(ws + kNaQueue + kQueueHdr);
  const int c0 = blockIdx.x * kNaChunk, c1 = min(nb, c0 + kNaChunk);
  float s = 0.0f;
  for (int b = c0 + (int)threadIdx.x; b < c1; b += 256) s += cost[b];
  for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
  if (threadIdx.x % kWave == 0) part[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    reinterpret_cast<float*>(ws + kNaPartials)[blockIdx.x] =
        part[0] + part[1] + part[2] + part[3];
}

// Range starts: range x begins at the first bucket whose preceding cost
// reaches x/8 of the total (make_queue's rule). Thread x < 8 searches the
// chunk sums, then the costs inside its chunk.
__global__ void noalloc_split_kernel(int nb, int nchunks, int* ws) {
  const float* part = reinterpret_cast<const float*>(ws + kNaPartials);
  const float* cost = reinterpret_cast<const float*>(ws + kNaQueue + kQueueHdr);
  int* hdr = ws + kNaQueue;
  const int x = threadIdx.x;
  if (x > kNumXcd) return;
  if (x == 0) {
    hdr[0] = 0;
    return;
  }
  if (x == kNumXcd) {
    hdr[kNumXcd] = nb;
    return;
  }
  double total = 0.0;
  for (int c = 0; c < nchunks; ++c) total += part[c];
  const double goal = total * x / kNumXcd;
  double run = 0.0;
  int c = 0;
  while (c < nchunks && run + part[c] < goal) run += part[c++];
  int b = c * kNaChunk;
  const int bend = min(nb, b + kNaChunk);
  while (b < bend && run < goal) run += cost[b++];
  hdr[x] = min(b, nb);
}

__global__ void noalloc_ids_kernel(int nb, int* ws) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) ws[kNaQueue + kQueueHdr + b] = b;
}

// Zeroes the results of the particles the buckets cover.
__global__ void zero_buckets_kernel(const int* starts, const int* sizes, int nb,
                                    VariablePartData* vars) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const int s = starts[b], n = sizes[b];
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    vars[s + i] = VariablePartData{{0.f, 0.f, 0.f}, 0.f, 0.f};
}

}  // namespace cgpu

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
int* g_na_ptr[kMaxDevices] = {};

// Address of this device's copy of the scratch array, looked up once per
// device with that device current.
int na_scratch(int* dev_out, int** out) {
  int dev = 0;
  CGPU_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail("changaBatchedGravityNoAlloc: device ordinal", hipErrorInvalidDevice, __FILE__,
                __LINE__);
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    CGPU_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(cgpu::scratch)));
    g_na_ptr[dev] = static_cast<int*>(p);
  }
  *dev_out = dev;
  *out = g_na_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int changaBatchedGravityNoAlloc(const ChangaGravityLaunch* cells,
                                           const ChangaGravityLaunch* parts,
                                           const ChangaBatchOptions* opts, void* stream) {
  if (!cells || !parts)
    return fail("changaBatchedGravityNoAlloc: missing argument", hipErrorInvalidValue, __FILE__,
                __LINE__);
  const ChangaGravityLaunch& pc = *cells;
  const ChangaGravityLaunch& pp = *parts;
  const int nb = pc.numBuckets;
  if (nb < 0 || nb > CHANGA_NOALLOC_MAX_BUCKETS)
    return fail("changaBatchedGravityNoAlloc: numBuckets outside 0..CHANGA_NOALLOC_MAX_BUCKETS "
                "(the size of the static work queue)",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (nb > 0 && (!pc.marks || !pc.starts || !pc.sizes || !pc.vars || !pc.targets ||
                 !pc.moments || !pc.list || !pp.marks || !pp.list || !pp.sources))
    return fail("changaBatchedGravityNoAlloc: missing array", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (!same_buckets(pc, pp))
    return fail("changaBatchedGravityNoAlloc: cell and particle launches differ in their buckets",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (pc.tpos64 && (!pc.cm64 || !pp.spos64))
    return fail("changaBatchedGravityNoAlloc: precision form needs cm64 and spos64",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (pp.listKind != 0 && pp.listKind != 1)
    return fail("changaBatchedGravityNoAlloc: listKind", hipErrorInvalidValue, __FILE__, __LINE__);
  if (reinterpret_cast<uintptr_t>(pc.cellRecords) % CHANGA_CELL_RECORD_BYTES != 0)
    return fail("changaBatchedGravityNoAlloc: cellRecords not 128-byte aligned",
                hipErrorInvalidValue, __FILE__, __LINE__);
  ChangaBatchOptions o = {64, 256, 1, 1, 0, 0.0f, 3, 0};
  if (opts) {
    o = *opts;
    if (!o.pcTile) o.pcTile = 64;
    if (!o.ppTile) o.ppTile = 256;
    if (!o.listsPerGrab) o.listsPerGrab = 1;
    if (!o.order) o.order = 1;
    if (!o.targetsPerLane) o.targetsPerLane = 3;
  }
  if (o.cellSource == 0) o.cellSource = pc.cellRecords ? 1 : 2;
  if (o.cellSource < 1 || o.cellSource > 2 || (o.cellSource == 1 && !pc.cellRecords))
    return fail("changaBatchedGravityNoAlloc: cellSource (1 needs cellRecords)",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if ((o.pcTile != 32 && o.pcTile != 64) || (o.ppTile != 128 && o.ppTile != 256) ||
      o.listsPerGrab < 1 || o.listsPerGrab > kMaxGrab ||
      (o.targetsPerLane != 1 && o.targetsPerLane != 3))
    return fail("changaBatchedGravityNoAlloc: option out of range", hipErrorInvalidValue,
                __FILE__, __LINE__);
  if (o.order != 1)
    return fail("changaBatchedGravityNoAlloc: only tree order (order 1) runs without allocation",
                hipErrorInvalidValue, __FILE__, __LINE__);
  int dev = 0;
  int* ws = nullptr;
  if (int e = na_scratch(&dev, &ws)) return e;
  if (nb == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const BatchFn fn = pick(pc.tpos64 != nullptr, pp.listKind, o, o.cellSource == 1);
  const int grid = grid_for(fn);
  if (grid <= 0)
    return fail("changaBatchedGravityNoAlloc: occupancy", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (o.zeroVars) {
    hipLaunchKernelGGL(zero_buckets_kernel, dim3(nb), dim3(kWave), 0, s, pc.starts, pc.sizes, nb,
                       pc.vars);
    CGPU_CHECK(hipGetLastError());
  }
  const int nchunks = (nb + kNaChunk - 1) / kNaChunk;
  if (pp.listKind)
    hipLaunchKernelGGL(noalloc_cost_kernel<1>, dim3((nb + 255) / 256), dim3(256), 0, s, pc, pp, ws);
  else
    hipLaunchKernelGGL(noalloc_cost_kernel<0>, dim3((nb + 255) / 256), dim3(256), 0, s, pc, pp, ws);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(noalloc_chunk_sum_kernel, dim3(nchunks), dim3(256), 0, s, nb, ws);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(noalloc_split_kernel, dim3(1), dim3(kWave), 0, s, nb, nchunks, ws);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(noalloc_ids_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, nb, ws);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(kNumXcd * kCounterStride), 0, s,
                     ws + kNaCounters);
  CGPU_CHECK(hipGetLastError());
  BatchArgs a;
  a.pc = pc;
  a.pp = pp;
  a.queue = ws + kNaQueue;
  a.counters = ws + kNaCounters;
  a.grab = o.listsPerGrab;
  void* kargs[] = {&a};
  CGPU_CHECK(hipLaunchKernel((const void*)fn, dim3(grid), dim3(kWave), kargs, 0, s));
  return 0;
}
