// Per-bucket particle-cell and particle-particle kernels and their C-ABI
// launchers (changaNodeGravity, changaPartGravity, changaZeroVars,
// changaAddVars). The
// interaction bodies and the per-chunk loops are in gravity_body.h, shared
// with the persistent batched kernel of batch.hip.
#include "gravity_body.h"

namespace cgpu {

// ---------------------------------------------------------------------------
// Per-bucket particle-cell kernel: one 64-lane workgroup per bucket, buckets
// mapped to XCDs in contiguous ranges (xcd_bucket).
// ---------------------------------------------------------------------------
// REC: cells from the 128-byte cell records in tiles of 32 (their prefetch
// then fits three wavefronts per SIMD), else from the wire moments in tiles
// of kPcTile.
template <bool PRECISE, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void pc_kernel(
    ChangaGravityLaunch a) {
  constexpr int TILE = REC ? 32 : kPcTile;
  // staged tile; reused for the slot reduction at the end of each chunk
  __shared__ float4 rec4[(kPcTile + 1) * (kFwdRec / 4)];
  static_assert(sizeof(rec4) >= reduce_bytes<3>(), "reduction");

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  double cx0, cy0, cz0;
  bucket_centre<PRECISE>(a.tpos64, bstart, n, cx0, cy0, cz0);

  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    with_targets_per_lane(nt, [&](auto kt) {
      constexpr int K = decltype(kt)::value;
      const Lanes l = lanes_for<K>(lane, nt);
      Target t[K];
      load_targets<PRECISE, K>(a, bstart + t0, l, cx0, cy0, cz0, t);
      Acc64 acc[K];
#pragma unroll
      for (int j = 0; j < K; ++j) acc[j] = {0.0, 0.0, 0.0, 0.0, 0.0f};
      pc_cells<PRECISE, TILE, true, K, REC>(a, start, end, cx0, cy0, cz0, t, l, rec4, acc);
      reduce_store<K>(rec4, l, acc, a.vars + bstart + t0);
    });
  }
}

// ---------------------------------------------------------------------------
// Per-bucket particle-particle kernel (KIND: see pp_sources).
// ---------------------------------------------------------------------------
template <bool PRECISE, int KIND>
__global__ __launch_bounds__(64) void pp_kernel(ChangaGravityLaunch a) {
  // staged tiles; reused for the slot reduction at the end of each chunk
  __shared__ PpFwdLds<kPpTile> s;
  static_assert(sizeof(s) >= reduce_bytes<3>(), "reduction");

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  double cx0, cy0, cz0;
  bucket_centre<PRECISE>(a.tpos64, bstart, n, cx0, cy0, cz0);

  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    with_targets_per_lane(nt, [&](auto kt) {
      constexpr int K = decltype(kt)::value;
      const Lanes l = lanes_for<K>(lane, nt);
      Target t[K];
      load_targets<PRECISE, K>(a, bstart + t0, l, cx0, cy0, cz0, t);
      Acc64 acc[K];
#pragma unroll
      for (int j = 0; j < K; ++j) acc[j] = {0.0, 0.0, 0.0, 0.0, 0.0f};
      pp_sources<PRECISE, KIND, kPpTile, true, K>(a, start, end, cx0, cy0, cz0, t, l, s, acc);
      reduce_store<K>(&s, l, acc, a.vars + bstart + t0);
    });
  }
}

// 128-byte cell records (kCellRecFloats): centre of mass as 3 doubles, mass,
// softening, then the moments scaled as stage_cell_fwd scales them (same
// FP32 operations, so both staging paths give identical records).
__global__ void cell_records_kernel(const CudaMultipoleMoments* mom, const double* cm64, int n,
                                    float4* out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* w = reinterpret_cast<const float*>(mom) + (size_t)i * kWireStride;
  float v[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) v[k] = w[wire_float(k)];
  double c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = cm64 ? cm64[3 * (size_t)i + k] : (double)v[3 + k];
  float r[kCellRecFloats];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const long long b = __double_as_longlong(c[k]);
    r[2 * k] = __int_as_float((int)(unsigned)(b & 0xffffffffll));
    r[2 * k + 1] = __int_as_float((int)(unsigned)((unsigned long long)b >> 32));
  }
  r[F_M] = v[2];
  r[F_SOFT] = v[1];
  const float u = v[0], u2 = u * u;
  const float s2 = 3.0f * u2, s3 = 7.5f * u2 * u, s4 = 17.5f * u2 * u2;
#pragma unroll
  for (int k = 0; k < 5; ++k) r[F_Q2 + k] = s2 * v[6 + k];
  r[F_Q2 + 5] = -(r[F_Q2 + 0] + r[F_Q2 + 3]);
#pragma unroll
  for (int k = 0; k < 7; ++k) r[F_Q3 + k] = s3 * v[11 + k];
  r[F_Q3 + 7] = -(r[F_Q3 + 0] + r[F_Q3 + 1]);
  r[F_Q3 + 8] = -(r[F_Q3 + 2] + r[F_Q3 + 3]);
#pragma unroll
  for (int k = 0; k < 9; ++k) r[F_Q4 + k] = s4 * v[18 + k];
  float4* o = out + (size_t)i * (kCellRecFloats / 4);
#pragma unroll
  for (int k = 0; k < kCellRecFloats / 4; ++k)
    o[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

__global__ void zero_vars_kernel(VariablePartData* v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = VariablePartData{{0.f, 0.f, 0.f}, 0.f, 0.f};
}

// dst += src with the result array's semantics (+= for a and the potential,
// max for dtGrav)
__global__ void add_vars_kernel(VariablePartData* dst, const VariablePartData* src, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  VariablePartData d = dst[i];
  const VariablePartData s = src[i];
  d.a.x += s.a.x;
  d.a.y += s.a.y;
  d.a.z += s.a.z;
  d.potential += s.potential;
  d.dtGrav = fmaxf(d.dtGrav, s.dtGrav);
  dst[i] = d;
}

// A launch over no buckets (an empty rung) needs no arrays.
static int check_launch(const ChangaGravityLaunch* a) {
  if (!a || a->numBuckets < 0 ||
      (a->numBuckets > 0 && (!a->marks || !a->starts || !a->sizes || !a->vars || !a->targets)))
    return fail("gravity launch: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  return 0;
}

}  // namespace cgpu

using namespace cgpu;

extern "C" int changaNodeGravity(const ChangaGravityLaunch* a, void* stream) {
  if (int e = check_launch(a)) return e;
  if (a->numBuckets == 0) return 0;
  if (!a->moments || !a->list)
    return fail("changaNodeGravity: moments/list", hipErrorInvalidValue, __FILE__, __LINE__);
  if (a->tpos64 && !a->cm64)
    return fail("changaNodeGravity: precision form needs cm64", hipErrorInvalidValue, __FILE__,
                __LINE__);
  if (a->variant != 0)
    return fail("changaNodeGravity: variant must be 0", hipErrorInvalidValue, __FILE__, __LINE__);
  if (reinterpret_cast<uintptr_t>(a->cellRecords) % CHANGA_CELL_RECORD_BYTES != 0)
    return fail("changaNodeGravity: cellRecords not 128-byte aligned", hipErrorInvalidValue,
                __FILE__, __LINE__);
  hipStream_t s = (hipStream_t)stream;
  const dim3 g(a->numBuckets), blk(kWave);
  if (a->cellRecords) {
    if (a->tpos64) hipLaunchKernelGGL((pc_kernel<true, true>), g, blk, 0, s, *a);
    else hipLaunchKernelGGL((pc_kernel<false, true>), g, blk, 0, s, *a);
  } else {
    if (a->tpos64) hipLaunchKernelGGL((pc_kernel<true, false>), g, blk, 0, s, *a);
    else hipLaunchKernelGGL((pc_kernel<false, false>), g, blk, 0, s, *a);
  }
  CGPU_CHECK(hipGetLastError());
  return 0;
}

extern "C" int changaCellRecords(const CudaMultipoleMoments* moments, const double* cm64,
                                 int numCells, void* records, void* stream) {
  if (numCells < 0 || (numCells > 0 && (!moments || !records)))
    return fail("changaCellRecords: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  if (numCells > CHANGA_CELL_RECORD_MAX)
    return fail("changaCellRecords: more than CHANGA_CELL_RECORD_MAX cells", hipErrorInvalidValue,
                __FILE__, __LINE__);
  if (reinterpret_cast<uintptr_t>(records) % CHANGA_CELL_RECORD_BYTES != 0)
    return fail("changaCellRecords: records not 128-byte aligned", hipErrorInvalidValue,
                __FILE__, __LINE__);
  if (numCells == 0) return 0;
  hipLaunchKernelGGL(cell_records_kernel, dim3((numCells + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, moments, cm64, numCells, static_cast<float4*>(records));
  CGPU_CHECK(hipGetLastError());
  return 0;
}

extern "C" int changaPartGravity(const ChangaGravityLaunch* a, void* stream) {
  if (int e = check_launch(a)) return e;
  if (a->numBuckets == 0) return 0;
  if (!a->sources || !a->list)
    return fail("changaPartGravity: sources/list", hipErrorInvalidValue, __FILE__, __LINE__);
  if (a->tpos64 && !a->spos64)
    return fail("changaPartGravity: precision form needs spos64", hipErrorInvalidValue, __FILE__,
                __LINE__);
  hipStream_t s = (hipStream_t)stream;
  dim3 g(a->numBuckets), blk(kWave);
  bool pr = a->tpos64 != nullptr;
  if (a->listKind == 0) {
    if (pr) hipLaunchKernelGGL((pp_kernel<true, 0>), g, blk, 0, s, *a);
    else hipLaunchKernelGGL((pp_kernel<false, 0>), g, blk, 0, s, *a);
  } else {
    if (pr) hipLaunchKernelGGL((pp_kernel<true, 1>), g, blk, 0, s, *a);
    else hipLaunchKernelGGL((pp_kernel<false, 1>), g, blk, 0, s, *a);
  }
  CGPU_CHECK(hipGetLastError());
  return 0;
}

extern "C" int changaAddVars(VariablePartData* dst, const VariablePartData* src, int n,
                             void* stream) {
  if (n <= 0) return 0;
  if (!dst || !src)
    return fail("changaAddVars: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  hipLaunchKernelGGL(add_vars_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     dst, src, n);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

extern "C" int changaZeroVars(VariablePartData* vars, int n, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(zero_vars_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     vars, n);
  CGPU_CHECK(hipGetLastError());
  return 0;
}
