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
template <bool PRECISE>
__global__ __launch_bounds__(64) void pc_kernel(ChangaGravityLaunch a) {
  // staged tile; reused for the slot reduction at the end of each chunk
  __shared__ float4 rec4[kPcTile * (kPcRec / 4)];
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
      pc_cells<PRECISE, kPcTile, true, K>(a, start, end, cx0, cy0, cz0, t, l, rec4, acc);
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
  hipStream_t s = (hipStream_t)stream;
  if (a->tpos64) hipLaunchKernelGGL(pc_kernel<true>, dim3(a->numBuckets), dim3(kWave), 0, s, *a);
  else hipLaunchKernelGGL(pc_kernel<false>, dim3(a->numBuckets), dim3(kWave), 0, s, *a);
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
