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
  static_assert(sizeof(rec4) >= 4 * kWave * sizeof(double) + kWave * sizeof(float), "reduction");
  double(*red)[kWave] = reinterpret_cast<double(*)[kWave]>(rec4);
  float* redi = reinterpret_cast<float*>(reinterpret_cast<double*>(rec4) + 4 * kWave);

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  double cx0, cy0, cz0;
  bucket_centre<PRECISE>(a.tpos64, bstart, n, cx0, cy0, cz0);

  for (int t0 = 0; t0 < n; t0 += kWave) {
    const Lanes l = lanes_for(lane, min(kWave, n - t0));
    const Target t = load_target<PRECISE>(a, bstart + t0 + (l.live ? l.lane - l.slot * l.nt : 0),
                                          cx0, cy0, cz0);
    Acc64 acc = {0.0, 0.0, 0.0, 0.0, 0.0f};
    pc_cells<PRECISE, kPcTile>(a, start, end, cx0, cy0, cz0, t, l, rec4, acc);
    reduce_store(red, redi, lane, l.nt, l.nslot, l.slot, acc.ax, acc.ay, acc.az, acc.pot, acc.idt,
                 a.vars + bstart + t0 + (lane < l.nt ? lane : 0));
  }
}

// ---------------------------------------------------------------------------
// Particle-cell kernel, flat-staging variant (variant 1): cells are fetched
// with coalesced flat loads (27 floats per cell spread over the wavefront)
// into a 28-float record, then pre-scaled in place; per-cell softened test.
// ---------------------------------------------------------------------------
constexpr int kFlatRec = 28;
enum { F_CX = 0, F_CY, F_CZ, F_M, F_Q2 = 4, F_Q3 = 9, F_Q4 = 16, F_SOFT = 25, F_U = 26, F_FLAG = 27 };
__device__ __forceinline__ int flat_slot(int f) {
  return f == 0 ? F_U : f == 1 ? F_SOFT : f == 2 ? F_M : f <= 5 ? f - 3 : f - 2;
}

// Hexadecapole cell at r = tgt - cm. The moments in `q` are radius-scaled
// reduced STF moments pre-multiplied by n c_n radius^n (c_n = (2n-1)!!/n!).
// With n = r/|r| and V_n = Q_n . n^(n-1) (from harmonic polynomials of n),
// P_n = V_n . n:
//   phi -= dir (m + dir^2 (P2/2 + dir (P3/3 + dir P4/4)))
//   a   += dir^2 (dir^2 (V2 + dir (V3 + dir V4)) - S n),
//   S = m + dir^2 (5/2 P2 + dir (7/3 P3 + dir 9/4 P4)).
__device__ __forceinline__ void pc_interact_wire(float rx, float ry, float rz, const float4* rec,
                                            float mt, Acc32& A) {
  float r2 = rx * rx + ry * ry + rz * rz;
  if (r2 == 0.0f) return;
  float dir = rsq(r2);
  float4 c0 = rec[0], c1 = rec[1], c2 = rec[2], c3 = rec[3], c4 = rec[4], c5 = rec[5];
  float m = c0.w;
  float qxx = c1.x, qxy = c1.y, qxz = c1.z, qyy = c1.w, qyz = c2.x;
  float qxxx = c2.y, qxyy = c2.z, qxxy = c2.w, qyyy = c3.x, qxxz = c3.y, qyyz = c3.z, qxyz = c3.w;
  float qxxxx = c4.x, qxyyy = c4.y, qxxxy = c4.z, qyyyy = c4.w, qxxxz = c5.x, qyyyz = c5.y,
        qxxyy = c5.z, qxxyz = c5.w;
  float qxyyz = rec[6].x;
  float x = rx * dir, y = ry * dir, z = rz * dir;
  float x2 = x * x, y2 = y * y, z2 = z * z;
  // rank-2 harmonic polynomials
  float fxx = x2 - z2, fyy = y2 - z2;
  float x_2 = x + x;
  float fxy = x_2 * y, fxz = x_2 * z, fyz = 2.0f * y * z;
  // rank-3 harmonic polynomials
  float z3 = 3.0f * z2;
  float fxxx = x * (x2 - z3), fyyy = y * (y2 - z3);
  float fxyy = 3.0f * x * fyy, fxxy = 3.0f * y * fxx;
  float fxxz = z * (3.0f * x2 - z2), fyyz = z * (3.0f * y2 - z2);
  float fxyz = 3.0f * fxy * z;
  // V2 = Q2 . n
  float v2x = qxx * x + qxy * y + qxz * z;
  float v2y = qxy * x + qyy * y + qyz * z;
  float v2z = qxz * x + qyz * y - (qxx + qyy) * z;
  // V3 = Q3 . n n
  float v3x = qxxx * fxx + qxyy * fyy + qxxy * fxy + qxxz * fxz + qxyz * fyz;
  float v3y = qxyy * fxy + qxxy * fxx + qyyy * fyy + qyyz * fyz + qxyz * fxz;
  float v3z = qxxz * fxx + qyyz * fyy + qxyz * fxy - (qxxx + qxyy) * fxz - (qxxy + qyyy) * fyz;
  // V4 = Q4 . n n n
  float v4x = qxxxx * fxxx + qxxyy * fxyy + qxxxy * fxxy + qxyyy * fyyy + qxxxz * fxxz +
              qxyyz * fyyz + qxxyz * fxyz;
  float v4y = qyyyy * fyyy + qxxyy * fxxy + qxxxy * fxxx + qxyyy * fxyy + qxyyz * fxyz +
              qxxyz * fxxz + qyyyz * fyyz;
  float v4z = qxxxz * fxxx + qxyyz * fxyy + qxxyz * fxxy + qyyyz * fyyy - qxxxx * fxxz -
              qyyyy * fyyz - qxxyy * (fxxz + fyyz) - (qxxxy + qxyyy) * fxyz;
  float p2 = v2x * x + v2y * y + v2z * z;
  float p3 = v3x * x + v3y * y + v3z * z;
  float p4 = v4x * x + v4y * y + v4z * z;
  float d2 = dir * dir;
  float pin = m + d2 * (0.5f * p2 + dir * ((1.0f / 3.0f) * p3 + dir * (0.25f * p4)));
  float S = m + d2 * (2.5f * p2 + dir * ((7.0f / 3.0f) * p3 + dir * (2.25f * p4)));
  float tx = v2x + dir * (v3x + dir * v4x);
  float ty = v2y + dir * (v3y + dir * v4y);
  float tz = v2z + dir * (v3z + dir * v4z);
  A.pot -= dir * pin;
  A.ax += d2 * (d2 * tx - S * x);
  A.ay += d2 * (d2 * ty - S * y);
  A.az += d2 * (d2 * tz - S * z);
  A.idt = fmaxf(A.idt, (mt + m) * d2 * dir);
}

template <bool PRECISE>
__global__ __launch_bounds__(64) void pc_kernel_flat(ChangaGravityLaunch a) {
  __shared__ float4 rec4[kPcTile * (kFlatRec / 4)];
  __shared__ int s_idx[kPcTile], s_oid[kPcTile];
  __shared__ double red[4][kWave];
  __shared__ float redi[kWave];
  float* rec = reinterpret_cast<float*>(rec4);

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  const float L = a.fperiod;
  const ILCell* cells = static_cast<const ILCell*>(a.list);
  const float* wire = reinterpret_cast<const float*>(a.moments);

  double cx0 = 0, cy0 = 0, cz0 = 0;  // bucket centre (precision form)
  if (PRECISE && n > 0) {
    cx0 = a.tpos64[3 * (size_t)bstart];
    cy0 = a.tpos64[3 * (size_t)bstart + 1];
    cz0 = a.tpos64[3 * (size_t)bstart + 2];
  }

  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    const int nslot = kWave / nt;
    const int slot = lane / nt, j = lane - slot * nt;
    const bool live = slot < nslot;
    const int tp = bstart + t0 + (live ? j : 0);
    float tx, ty, tz, tm, tsoft;
    {
      CompactPartData p = a.targets[tp];
      tm = p.mass;
      tsoft = p.soft;
      if (PRECISE) {
        tx = (float)(a.tpos64[3 * (size_t)tp] - cx0);
        ty = (float)(a.tpos64[3 * (size_t)tp + 1] - cy0);
        tz = (float)(a.tpos64[3 * (size_t)tp + 2] - cz0);
      } else {
        tx = p.position.x;
        ty = p.position.y;
        tz = p.position.z;
      }
    }
    double dax = 0, day = 0, daz = 0, dpot = 0;
    float idt = 0.0f;

    for (int base = start; base < end; base += kPcTile) {
      const int nc = min(kPcTile, end - base);
      if (lane < nc) {
        const ILCell c = load_stream(cells + base + lane);
        s_idx[lane] = c.index;
        s_oid[lane] = c.offsetID;
      }
      __syncthreads();
      for (int e = lane; e < nc * 27; e += kWave) {
        int c = e / 27, f = e - 27 * c;
        rec[c * kFlatRec + flat_slot(f)] = wire[(size_t)s_idx[c] * 27 + f];
      }
      __syncthreads();
      if (lane < nc) {
        float* r = rec + lane * kFlatRec;
        int oid = s_oid[lane];
        if (PRECISE) {
          size_t ci = (size_t)s_idx[lane] * 3;
          r[F_CX] = (float)(a.cm64[ci] + (double)off_x(oid) * (double)L - cx0);
          r[F_CY] = (float)(a.cm64[ci + 1] + (double)off_y(oid) * (double)L - cy0);
          r[F_CZ] = (float)(a.cm64[ci + 2] + (double)off_z(oid) * (double)L - cz0);
        } else {
          r[F_CX] = (float)off_x(oid) * L + r[F_CX];
          r[F_CY] = (float)off_y(oid) * L + r[F_CY];
          r[F_CZ] = (float)off_z(oid) * L + r[F_CZ];
        }
        float u = r[F_U], u2 = u * u;
        float s2 = 3.0f * u2, s3 = 7.5f * u2 * u, s4 = 17.5f * u2 * u2;
#pragma unroll
        for (int k = 0; k < 5; ++k) r[F_Q2 + k] *= s2;
#pragma unroll
        for (int k = 0; k < 7; ++k) r[F_Q3 + k] *= s3;
#pragma unroll
        for (int k = 0; k < 9; ++k) r[F_Q4 + k] *= s4;
        r[F_FLAG] = (oid & CHANGA_ILCELL_SOFTENED) ? 1.0f : 0.0f;
      }
      __syncthreads();
      if (live) {
        Acc32 A = {0.f, 0.f, 0.f, 0.f, idt};
        for (int c = slot; c < nc; c += nslot) {
          const float4* r4 = rec4 + c * (kFlatRec / 4);
          float4 h = r4[0];
          if (r4[6].w != 0.0f) {
            // softened cell: spline monopole pseudo-particle at the cm
            pp_interact(h.x - tx, h.y - ty, h.z - tz, h.w, r4[6].y, tm, tsoft, A);
          } else {
            pc_interact_wire(tx - h.x, ty - h.y, tz - h.z, r4, tm, A);
          }
        }
        dax += A.ax;
        day += A.ay;
        daz += A.az;
        dpot += A.pot;
        idt = A.idt;
      }
      __syncthreads();
    }
    reduce_store(red, redi, lane, nt, nslot, slot, dax, day, daz, dpot, idt,
                 a.vars + bstart + t0 + (lane < nt ? lane : 0));
  }
}

// ---------------------------------------------------------------------------
// Per-bucket particle-particle kernel (KIND: see pp_sources).
// ---------------------------------------------------------------------------
template <bool PRECISE, int KIND>
__global__ __launch_bounds__(64) void pp_kernel(ChangaGravityLaunch a) {
  __shared__ PpLds<kPpTile> s;
  __shared__ double red[4][kWave];
  __shared__ float redi[kWave];

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  double cx0, cy0, cz0;
  bucket_centre<PRECISE>(a.tpos64, bstart, n, cx0, cy0, cz0);

  for (int t0 = 0; t0 < n; t0 += kWave) {
    const Lanes l = lanes_for(lane, min(kWave, n - t0));
    const Target t = load_target<PRECISE>(a, bstart + t0 + (l.live ? l.lane - l.slot * l.nt : 0),
                                          cx0, cy0, cz0);
    Acc64 acc = {0.0, 0.0, 0.0, 0.0, 0.0f};
    pp_sources<PRECISE, KIND, kPpTile>(a, start, end, cx0, cy0, cz0, t, l, s, acc);
    reduce_store(red, redi, lane, l.nt, l.nslot, l.slot, acc.ax, acc.ay, acc.az, acc.pot, acc.idt,
                 a.vars + bstart + t0 + (lane < l.nt ? lane : 0));
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
  hipStream_t s = (hipStream_t)stream;
  const bool pr = a->tpos64 != nullptr;
  if (a->variant == 1) {
    if (pr) hipLaunchKernelGGL(pc_kernel_flat<true>, dim3(a->numBuckets), dim3(kWave), 0, s, *a);
    else hipLaunchKernelGGL(pc_kernel_flat<false>, dim3(a->numBuckets), dim3(kWave), 0, s, *a);
  } else {
    if (pr) hipLaunchKernelGGL(pc_kernel<true>, dim3(a->numBuckets), dim3(kWave), 0, s, *a);
    else hipLaunchKernelGGL(pc_kernel<false>, dim3(a->numBuckets), dim3(kWave), 0, s, *a);
  }
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
