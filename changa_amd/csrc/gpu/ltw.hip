// The local-tree-walk request route of ChaNGa's GPU_LOCAL_TREE_WALK build:
// TreePieceCellListDataTransferLocal runs gpuLocalTreeWalk
// (HostCUDA.cu:190-224, 724-942) on the particles [firstParticle,
// lastParticle] of the serialized local tree (DataManager.cpp:469-953) with
// no interaction list.
//
// Semantics kept from the reference walk:
//   - every particle is a target, whatever its bucket's rung;
//   - the tree is walked from rootIdx once per periodic replica (offsets
//     x fperiod, y fperiodY, z fperiodZ added to the node's centre of mass,
//     and to the particles of an opened bucket);
//   - a node is opened (CUDA_openCriterionNode, CUDAMoments.cu:197-242) when
//     it holds at most 6 particles, or when its opening sphere (radius
//     max(2/sqrt(3) r / theta, r)) meets the bounding box of the target
//     particle's bucket node, or, when its softening overlaps that bucket
//     (CUDA_openSoftening, :180-195), when the monopole sphere (2/sqrt(3) r /
//     thetaMono) does; all of it in FP32 with the reference's operation order;
//   - the node type maps the decision to an action (CUDA_OptAction,
//     :323-352): a kept node is a hexadecapole cell, or a spline-softened
//     monopole where the softening overlaps; an opened internal or boundary
//     node pushes its children, an opened bucket its particles
//     [bucketStart, bucketStart + bucketSize); non-local, cached and empty
//     nodes are skipped, as are top and invalid nodes (the reference's ERROR
//     action computes nothing either);
//   - a += and potential += per particle, dtGrav = max (:930-935).
// The test depends only on the particle's bucket node, so this kernel walks
// once per bucket: one wavefront takes a run of consecutive particles with
// the same nodeId, walks the tree cooperatively over an LDS stack (as the
// fused kernel does, with this route's test and node types) into its own
// scratch lists, then runs the plain form's cell and particle loops
// (gravity_body.h) over them.
//
// Unlike the reference (a 64-entry stack per thread, unchecked), capacities
// and node indices are checked: a walk beyond the scratch, a child or bucket
// node outside the node array, or a particle range outside the particle
// array sets a flag (the run's forces are then not added) that the launcher
// turns into an error.
#ifndef GPU_LOCAL_TREE_WALK
#error "ltw.hip belongs to the GPU_LOCAL_TREE_WALK build of the library"
#endif

#include <cmath>
#include <cstring>

#include "gravity_body.h"
#include "walk_rec.h"

namespace cgpu {

constexpr int kLtwPcTile = 64;
constexpr int kLtwPpTile = 256;
constexpr int kLtwRanges = kNumXcd;
constexpr int kLtwCounterStride = 32;
constexpr int kLtwErrCapacity = 1;  // a walk exceeded maxCells / maxRuns or the stack
constexpr int kLtwErrIndex = 2;     // a node or particle index outside its array

struct LtwArgs {
  ChangaGravityLaunch pc, pp;  // lists: the bases of the wavefronts' scratch
  const CudaMultipoleMoments* nodes;
  const CompactPartData* parts;
  int numNodes, numParts, first, last, root, R;
  float theta, thetaMono;
  float px, py, pz;  // replica periods per axis
  ILCell* cells;  // per wave: maxCells entries
  ILPart* runs;   // per wave: maxRuns entries
  int maxCells, maxRuns;
  int* counters;  // kLtwRanges claim counters, kLtwCounterStride ints apart; then err
  int* err;
};

template <int PCT, int PPT>
struct LtwLds {
  static constexpr size_t stack = kCoopStack * sizeof(unsigned);
  static constexpr size_t pc = (size_t)(PCT + 1) * kFwdRec * sizeof(float);
  static constexpr size_t pp = sizeof(PpFwdLds<PPT>);
  static constexpr size_t red = reduce_bytes<1>();
  static constexpr size_t m1 = stack > pc ? stack : pc;
  static constexpr size_t m2 = pp > red ? pp : red;
  static constexpr size_t bytes = m1 > m2 ? m1 : m2;
  static constexpr size_t f4 = (bytes + 15) / 16;
};

// ---------------------------------------------------------------------------
// The test, FP32 as the reference computes it (contraction off: no fused
// multiply-adds that the reference's expressions do not spell out).
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

struct LtwBucket {
  float cx, cy, cz, rb;  // centre of mass, twice the softening
  float lo[3], hi[3];
};

struct LtwNode {
  float radius, soft, cx, cy, cz;
  int count, c0, c1, type, bstart, bsize;
};

__device__ __forceinline__ LtwNode ltw_node(const CudaMultipoleMoments* nodes, int nd) {
  const CudaMultipoleMoments& m = nodes[nd];
  LtwNode v;
  v.radius = m.radius;
  v.soft = m.soft;
  v.cx = m.cm.x;
  v.cy = m.cm.y;
  v.cz = m.cm.z;
  v.count = m.particleCount;
  v.c0 = m.children[0];
  v.c1 = m.children[1];
  v.type = m.type;
  v.bstart = m.bucketStart;
  v.bsize = m.bucketSize;
  return v;
}

// cuda_intersect (CUDAMoments.cu:137-159): squared distance from the sphere
// centre to the box, axes in order, against r^2 (the reference's early exits
// after x and y cannot change the answer: the partial sums only grow)
__device__ __forceinline__ bool ltw_box_sphere(const LtwBucket& b, float x, float y, float z,
                                               float r) {
  const float c[3] = {x, y, z};
  float dsq = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float d = b.lo[k] - c[k];
    if (d > 0.0f) {
      dsq += d * d;
    } else {
      d = c[k] - b.hi[k];
      if (d > 0.0f) dsq += d * d;
    }
  }
  return dsq <= r * r;
}

// CUDA_openSoftening: the bucket's softening sphere against the node's, then
// the bucket's box against the node's softening sphere (rs = 2 soft)
__device__ __forceinline__ bool ltw_soft(const LtwBucket& b, float x, float y, float z,
                                         float rs) {
  const float dx = b.cx - x, dy = b.cy - y, dz = b.cz - z;
  const float dist = dx * dx + dy * dy + dz * dz;
  if (dist <= (b.rb + rs) * (b.rb + rs)) return true;
  return ltw_box_sphere(b, x, y, z, rs);
}

// -1 nothing, 0 push the children, 1 hexadecapole cell, 2 softened cell,
// 3 the node's particles
__device__ __forceinline__ int ltw_decide(const LtwNode& v, const LtwBucket& b, float ox,
                                          float oy, float oz, float theta, float thetaMono) {
  const double kGeom = 2.0 / sqrt(3.0);  // openingGeometryFactor
  const float x = v.cx + ox, y = v.cy + oy, z = v.cz + oz;
  const float rs = 2.0f * v.soft;
  bool open;
  if (v.count <= 6) {
    open = true;
  } else {
    float r = (float)(kGeom * (double)v.radius / (double)theta);
    if (r < v.radius) r = v.radius;
    if (ltw_box_sphere(b, x, y, z, r))
      open = true;
    else if (!ltw_soft(b, x, y, z, rs))
      open = false;
    else
      open = ltw_box_sphere(b, x, y, z, (float)(kGeom * (double)v.radius / (double)thetaMono));
  }
  const int t = v.type;
  if (!open) {
    if (t == CudaInternal || t == CudaBucket || t == CudaBoundary || t == CudaNonLocalBucket)
      return ltw_soft(b, x, y, z, rs) ? 2 : 1;
    return -1;
  }
  if (t == CudaInternal || t == CudaBoundary) return 0;
  if (t == CudaBucket) return 3;
  return -1;
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Cooperative walk of one bucket (node bn) over the LDS stack: entries are
// node indices (low kCoopNodeBits bits) with the replica code above. Cells
// go to cells[0, nhex), softened cells to cells[maxCells - nsoft, maxCells),
// particle runs to runs[0, nruns). Returns false (and raises err) when a
// capacity or an index check fails.
// ---------------------------------------------------------------------------
__device__ bool ltw_walk(const LtwArgs& a, const LtwBucket& b, int lane, unsigned* stack,
                         ILCell* cells, ILPart* runs, int& nhex, int& nsoft, int& nruns) {
  nhex = nsoft = nruns = 0;
  const int R = a.R, W = 2 * R + 1, nrep = W * W * W;
  for (int k = lane; k < nrep; k += kWave) {
    const int x = k / (W * W) - R, y = (k / W) % W - R, z = k % W - R;
    const unsigned code = (unsigned)((x + 3) | ((y + 3) << 3) | ((z + 3) << 6));
    stack[k] = (code << kCoopNodeBits) | (unsigned)a.root;
  }
  int sp = nrep;
  int bad = 0;
  __syncthreads();
  while (sp > 0) {
    const int n = min(min(kWave, sp), max(1, kCoopStack - sp));
    unsigned e = 0;
    if (lane < n) e = stack[sp - n + lane];
    __syncthreads();
    sp -= n;
    int kind = -1, nd = 0, code = 0;
    LtwNode v = {};
    if (lane < n) {
      nd = (int)(e & ((1u << kCoopNodeBits) - 1u));
      code = (int)(e >> kCoopNodeBits);
      v = ltw_node(a.nodes, nd);
      const int x = (code & 7) - 3, y = ((code >> 3) & 7) - 3, z = ((code >> 6) & 7) - 3;
      kind = ltw_decide(v, b, (float)x * a.px, (float)y * a.py, (float)z * a.pz, a.theta,
                        a.thetaMono);
      // children and particle ranges must lie in their arrays
      if (kind == 0 && (v.c0 < -1 || v.c0 >= a.numNodes || v.c1 < -1 || v.c1 >= a.numNodes)) {
        bad = 1;
        kind = -1;
      }
      if (kind == 3 && (v.bsize < 0 || (v.bsize > 0 && (v.bstart < 0 ||
                        (long long)v.bstart + v.bsize > (long long)a.numParts)))) {
        bad = 1;
        kind = -1;
      }
    }
    const int p0 = kind == 0 && v.c0 >= 0, p1 = kind == 0 && v.c1 >= 0;
    const unsigned long long mh = __ballot(kind == 1), ms = __ballot(kind == 2),
                             mr = __ballot(kind == 3 && v.bsize > 0), m0 = __ballot(p0),
                             m1 = __ballot(p1);
    const int ch = __popcll(mh), cs = __popcll(ms), cr = __popcll(mr);
    const int cp = __popcll(m0) + __popcll(m1);
    if (__ballot(bad) != 0ull) {
      if (lane == 0) atomicOr(a.err, kLtwErrIndex);
      __syncthreads();
      return false;
    }
    if (nhex + ch + 2 * (nsoft + cs) > a.maxCells || nruns + cr > a.maxRuns ||
        sp + cp > kCoopStack) {
      if (lane == 0) atomicOr(a.err, kLtwErrCapacity);
      __syncthreads();
      return false;
    }
    const int oid = code << 22;
    if (kind == 1) cells[nhex + lanes_below(mh, lane)] = {nd, oid};
    if (kind == 2)
      cells[a.maxCells - 1 - (nsoft + lanes_below(ms, lane))] = {nd, oid | CHANGA_ILCELL_SOFTENED};
    if (kind == 3 && v.bsize > 0) runs[nruns + lanes_below(mr, lane)] = {v.bstart, oid, v.bsize};
    // pushes: lane order, the first child above the second (as the reference)
    const int at = sp + lanes_below(m0, lane) + lanes_below(m1, lane);
    const unsigned hi = (unsigned)code << kCoopNodeBits;
    if (p1) stack[at] = hi | (unsigned)v.c1;
    if (p0) stack[at + p1] = hi | (unsigned)v.c0;
    nhex += ch;
    nsoft += cs;
    nruns += cr;
    sp += cp;
    __syncthreads();
  }
  __syncthreads();
  return true;
}

// Four wavefronts per SIMD, as the fused kernel, and one target per lane:
// one to three targets per lane at three wavefronts (168 VGPRs) measured
// 5-20% slower on the 2^22 halos, the 128^3 box and the 2^20 sphere (the
// walk's load latency wants the fourth wavefront).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void ltw_kernel(
    LtwArgs a) {
  __shared__ float4 lds[LtwLds<kLtwPcTile, kLtwPpTile>::f4];
  unsigned* stack = reinterpret_cast<unsigned*>(lds);
  PpFwdLds<kLtwPpTile>& ps = *reinterpret_cast<PpFwdLds<kLtwPpTile>*>(lds);
  const int lane = threadIdx.x;
  const int cbase = blockIdx.x * a.maxCells, rbase = blockIdx.x * a.maxRuns;
  ILCell* cells = a.cells + cbase;
  ILPart* runs = a.runs + rbase;
  // chunks of 64 consecutive particles in kLtwRanges contiguous ranges
  const int nchunk = (a.last - a.first) / kWave + 1;
  int q = blockIdx.x % kLtwRanges;
  for (int tried = 0; tried < kLtwRanges;) {
    const int r0 = (int)((long long)nchunk * q / kLtwRanges);
    const int r1 = (int)((long long)nchunk * (q + 1) / kLtwRanges);
    int k = 0;
    if (lane == 0) k = atomicAdd(a.counters + q * kLtwCounterStride, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    if (r0 + k >= r1) {
      ++tried;
      q = q + 1 == kLtwRanges ? 0 : q + 1;
      continue;
    }
    const int p0 = a.first + (r0 + k) * kWave;
    const int np = min(kWave, a.last - p0 + 1);
    // the runs (same nodeId) that start in this chunk
    const int nid = lane < np ? a.parts[p0 + lane].nodeId : -1;
    int prev = __shfl_up(nid, 1, kWave);
    if (lane == 0) prev = p0 > a.first ? a.parts[p0 - 1].nodeId : ~nid;
    unsigned long long starts = __ballot(lane < np && nid != prev);
    while (starts) {
      const int s = __builtin_ctzll(starts);
      starts &= starts - 1ull;
      const int bn = __shfl(nid, s, kWave);
      const int rs = p0 + s;
      // the run's end: the first particle past it with another nodeId
      int re = a.last + 1;
      for (int q0 = rs + 1; q0 <= a.last; q0 += kWave) {
        const int p = q0 + lane;
        const unsigned long long m = __ballot(p <= a.last && a.parts[p].nodeId != bn);
        if (m) {
          re = q0 + __builtin_ctzll(m);
          break;
        }
      }
      if (bn < 0 || bn >= a.numNodes) {
        if (lane == 0) atomicOr(a.err, kLtwErrIndex);
        continue;
      }
      LtwBucket b;
      {
        const CudaMultipoleMoments& m = a.nodes[bn];
        b.cx = m.cm.x;
        b.cy = m.cm.y;
        b.cz = m.cm.z;
        b.rb = 2.0f * m.soft;
        b.lo[0] = m.lesser_corner.x;
        b.lo[1] = m.lesser_corner.y;
        b.lo[2] = m.lesser_corner.z;
        b.hi[0] = m.greater_corner.x;
        b.hi[1] = m.greater_corner.y;
        b.hi[2] = m.greater_corner.z;
      }
      int nhex, nsoft, nruns;
      if (!ltw_walk(a, b, lane, stack, cells, runs, nhex, nsoft, nruns)) continue;
      // softened cells (filled from the end of the region) moved behind the
      // hexadecapole cells, so one cell loop covers both
      if (nsoft > 0) {
        for (int i = lane; i < nsoft; i += kWave) cells[nhex + i] = cells[a.maxCells - 1 - i];
        __syncthreads();
      }
      for (int t0 = rs; t0 < re; t0 += kWave) {
        const Lanes l = lanes_for<1>(lane, min(kWave, re - t0));
        Target t[1];
        load_targets<false, 1>(a.pc, t0, l, 0.0, 0.0, 0.0, t);
        Acc64 acc[1] = {{0.0, 0.0, 0.0, 0.0, 0.0f}};
        pc_cells<false, kLtwPcTile, false, 1, false>(a.pc, cbase, cbase + nhex + nsoft, 0.0,
                                                      0.0, 0.0, t, l, lds, acc);
        pp_sources<false, 1, kLtwPpTile, false, 1>(a.pp, rbase, rbase + nruns, 0.0, 0.0, 0.0, t,
                                                   l, ps, acc);
        reduce_store<1>(lds, l, acc, a.pc.vars + t0);
      }
    }
  }
}

__global__ void ltw_reset_kernel(int* counters) {
  const int i = threadIdx.x;
  if (i < kLtwRanges * kLtwCounterStride + 1) counters[i] = 0;  // counters, then err
}

// Wavefronts of a request for n particles: the resident maximum, but no
// more than its 64-particle chunks (each wavefront holds its own scratch).
static int ltw_grid(int n, int* grid) {
  int dev = 0, ncu = 0, per = 0;
  CGPU_CHECK(hipGetDevice(&dev));
  CGPU_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  CGPU_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)ltw_kernel, kWave,
                                                          0));
  const int full = ncu * (per > 0 ? per : 1);
  const int nchunk = n > 0 ? (n - 1) / kWave + 1 : 1;
  *grid = full < nchunk ? full : nchunk;
  return 0;
}

// work buffer: claim counters and error flag (256-byte aligned head), then
// the per-wavefront cell and run scratch
struct LtwWork {
  size_t cells, runs, total;
};

static LtwWork ltw_layout(int grid, int maxCells, int maxRuns) {
  LtwWork w;
  w.cells = 256 * ((sizeof(int) * (kLtwRanges * kLtwCounterStride + 1) + 255) / 256);
  w.runs = w.cells + 256 * (((size_t)grid * maxCells * sizeof(ILCell) + 255) / 256);
  w.total = w.runs + (size_t)grid * maxRuns * sizeof(ILPart);
  return w;
}

// Work bytes of a request for n particles.
int ltw_work_bytes(int n, int maxCells, int maxRuns, size_t* bytes) {
  int grid = 0;
  if (int e = ltw_grid(n, &grid)) return e;
  if ((long long)grid * maxCells > 0x7fffffffll || (long long)grid * maxRuns > 0x7fffffffll)
    return fail("local tree walk: scratch capacities exceed the int32 index range",
                hipErrorInvalidValue, __FILE__, __LINE__);
  *bytes = ltw_layout(grid, maxCells, maxRuns).total;
  return 0;
}

// One local-tree-walk request: the checks of the arguments, the counter
// reset, the walk-and-force launch and the copy of its error flag to errOut
// (device or pinned host memory).
int ltw_launch(const CudaMultipoleMoments* nodes, int numNodes, const CompactPartData* parts,
               int numParts, VariablePartData* vars, int first, int last, int root, float theta,
               float thetaMono, int nReplicas, float fperiod, float fperiodY, float fperiodZ,
               void* work, size_t workBytes, int maxCells, int maxRuns, int* errOut,
               hipStream_t s) {
  if (last < first) return 0;  // no particles
  if (!nodes || !parts || !vars || !work)
    return fail("local tree walk: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  if (first < 0 || last >= numParts || numNodes <= 0 || numNodes > (1 << kCoopNodeBits) ||
      root < 0 || root >= numNodes)
    return fail("local tree walk: particle range, node count (1..2^23) or root index",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (nReplicas < 0 || nReplicas > 3 || !(theta > 0.0f) || !(thetaMono > 0.0f))
    return fail("local tree walk: theta, thetaMono, nReplicas (0..3)", hipErrorInvalidValue,
                __FILE__, __LINE__);
  if (nReplicas > 0 && (!(fperiod > 0.0f) || !(fperiodY > 0.0f) || !(fperiodZ > 0.0f)))
    return fail("local tree walk: periodic replicas need fperiod, fperiodY, fperiodZ > 0",
                hipErrorInvalidValue, __FILE__, __LINE__);
  const int n = last - first + 1;
  int grid = 0;
  if (int e = ltw_grid(n, &grid)) return e;
  size_t need = 0;
  if (int e = ltw_work_bytes(n, maxCells, maxRuns, &need)) return e;
  if (workBytes < need)
    return fail("local tree walk: work buffer too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  const LtwWork L = ltw_layout(grid, maxCells, maxRuns);
  char* wb = static_cast<char*>(work);
  LtwArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pc.targets = parts;
  a.pc.vars = vars;
  a.pc.moments = nodes;
  a.pc.fperiod = nReplicas > 0 ? fperiod : 0.0f;
  a.pc.fperiodY = nReplicas > 0 ? fperiodY : 0.0f;
  a.pc.fperiodZ = nReplicas > 0 ? fperiodZ : 0.0f;
  a.pc.listKind = 0;
  a.pp = a.pc;
  a.pp.sources = parts;
  a.pp.listKind = 1;
  a.nodes = nodes;
  a.parts = parts;
  a.numNodes = numNodes;
  a.numParts = numParts;
  a.first = first;
  a.last = last;
  a.root = root;
  a.R = nReplicas;
  a.theta = theta;
  a.thetaMono = thetaMono;
  a.px = a.pc.fperiod;
  a.py = a.pc.fperiodY;
  a.pz = a.pc.fperiodZ;
  a.counters = reinterpret_cast<int*>(wb);
  a.err = a.counters + kLtwRanges * kLtwCounterStride;
  a.cells = reinterpret_cast<ILCell*>(wb + L.cells);
  a.runs = reinterpret_cast<ILPart*>(wb + L.runs);
  a.maxCells = maxCells;
  a.maxRuns = maxRuns;
  a.pc.list = a.cells;
  a.pp.list = a.runs;
  hipLaunchKernelGGL(ltw_reset_kernel, dim3(1), dim3(kLtwRanges * kLtwCounterStride + 1), 0, s,
                     a.counters);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(ltw_kernel, dim3(grid), dim3(kWave), 0, s, a);
  CGPU_CHECK(hipGetLastError());
  if (errOut) CGPU_CHECK(hipMemcpyAsync(errOut, a.err, sizeof(int), hipMemcpyDefault, s));
  return 0;
}

}  // namespace cgpu
