// Ewald correction kernel: real-space replica sum plus k-space h-loop of the
// root multipole, per active particle.
//
// CPU semantics followed: TreePiece::BucketEwald (Ewald.cpp:72-281). The
// GPU work split of the reference is EwaldKernel (HostCUDA.cu:1958-2192).
//
//  - d = x - cm_root. Real space: lattice vectors n in [-nEwReps, nEwReps]^3
//    (nEwReps = max(ceil(fEwCut), nReps)); a term is skipped when
//    |d + nL|^2 > (fEwCut L)^2 unless n is in the "hole" |n_i| <= nReps, whose
//    terms use -erf (the tree already summed those replicas) instead of erfc.
//  - Radial kernels g_0..g_5 by the recursion g_(n+1) = (2n+1) g_n / r^2 +
//    (2 alpha^2)^n ka e^(-alpha^2 r^2) / r^2, or by the two-term series about
//    r = 0 when r^2 < 1.2e-3 L^2 (the CPU switch, Ewald.cpp:119).
//  - Multipole contraction with the root's STF moments at unit scale:
//    phi -= sum_n (1/n!) g_n Q_n . r^n,
//    a   += sum_n (1/n!) [n g_n Q_n . r^(n-1) - g_(n+1) (Q_n . r^n) r].
//  - k space: phi += hC cos(k.d) + hS sin(k.d), a += k (hC sin - hS cos).
//
// Precision: the central replica (n = 0, the only term that can reach small
// r, where the -erf kernels cancel catastrophically in FP32) is evaluated in
// FP64 with the CPU's series switch; every other term in FP32. All per-particle
// sums run in FP64 and are added into the FP32 result array at the end.
//
// Work split: one lane per active particle in tree order, so a wavefront
// holds spatially close particles. Each wavefront bounds its particles' d,
// keeps only the lattice vectors any of them can need (compacted into LDS),
// then every lane applies the exact per-particle test. The h-loop table is
// staged in LDS (dynamic size; no 80-entry limit).
#include <cmath>

#include "common.h"

namespace cgpu {

struct EwaldParams {
  double cmx, cmy, cmz;  // root centre of mass
  double mk1;            // M * k1, the k = 0 term
  double q[22];          // root reduced STF moments at unit scale (FMOMR order)
  double L, alpha, alpha2, ka, cut2, inner2;
  int nReps, nEwReps, nH, pad;
};

template <typename T>
struct Mom22 {
  T m, xx, yy, xy, xz, yz, xxx, xyy, xxy, yyy, xxz, yyz, xyz, xxxx, xyyy, xxxy, yyyy, xxxz, yyyz,
      xxyy, xxyz, xyyz;
};

// Ewald contribution of one replica vector r with radial kernels g[0..5].
template <typename T>
__device__ __forceinline__ void ewald_term(const Mom22<T>& Q, T x, T y, T z, const T* g, T& pot,
                                           T& ax, T& ay, T& az) {
  T x2 = x * x, y2 = y * y, z2 = z * z;
  T fxx = x2 - z2, fyy = y2 - z2;
  T fxy = T(2) * x * y, fxz = T(2) * x * z, fyz = T(2) * y * z;
  T z3 = T(3) * z2;
  T fxxx = x * (x2 - z3), fyyy = y * (y2 - z3);
  T fxyy = T(3) * x * fyy, fxxy = T(3) * y * fxx;
  T fxxz = z * (T(3) * x2 - z2), fyyz = z * (T(3) * y2 - z2);
  T fxyz = T(3) * fxy * z;
  T v2x = Q.xx * x + Q.xy * y + Q.xz * z;
  T v2y = Q.xy * x + Q.yy * y + Q.yz * z;
  T v2z = Q.xz * x + Q.yz * y - (Q.xx + Q.yy) * z;
  T v3x = Q.xxx * fxx + Q.xyy * fyy + Q.xxy * fxy + Q.xxz * fxz + Q.xyz * fyz;
  T v3y = Q.xyy * fxy + Q.xxy * fxx + Q.yyy * fyy + Q.yyz * fyz + Q.xyz * fxz;
  T v3z = Q.xxz * fxx + Q.yyz * fyy + Q.xyz * fxy - (Q.xxx + Q.xyy) * fxz - (Q.xxy + Q.yyy) * fyz;
  T v4x = Q.xxxx * fxxx + Q.xxyy * fxyy + Q.xxxy * fxxy + Q.xyyy * fyyy + Q.xxxz * fxxz +
          Q.xyyz * fyyz + Q.xxyz * fxyz;
  T v4y = Q.yyyy * fyyy + Q.xxyy * fxxy + Q.xxxy * fxxx + Q.xyyy * fxyy + Q.xyyz * fxyz +
          Q.xxyz * fxxz + Q.yyyz * fyyz;
  T v4z = Q.xxxz * fxxx + Q.xyyz * fxyy + Q.xxyz * fxxy + Q.yyyz * fyyy - Q.xxxx * fxxz -
          Q.yyyy * fyyz - Q.xxyy * (fxxz + fyyz) - (Q.xxxy + Q.xyyy) * fxyz;
  T p2 = v2x * x + v2y * y + v2z * z;
  T p3 = v3x * x + v3y * y + v3z * z;
  T p4 = v4x * x + v4y * y + v4z * z;
  const T h = T(0.5), s = T(1) / T(6), f = T(1) / T(24);
  pot -= g[0] * Q.m + g[2] * h * p2 + g[3] * s * p3 + g[4] * f * p4;
  T ta = g[1] * Q.m + g[3] * h * p2 + g[4] * s * p3 + g[5] * f * p4;
  T c3 = g[3] * h, c4 = g[4] * s;
  ax += g[2] * v2x + c3 * v3x + c4 * v4x - ta * x;
  ay += g[2] * v2y + c3 * v3y + c4 * v4y - ta * y;
  az += g[2] * v2z + c3 * v3z + c4 * v4z - ta * z;
}

// g_0..g_5: series about r = 0 or closed form (hole: -erf, else erfc)
__device__ __forceinline__ void radial64(double r2, bool hole, const EwaldParams& p, double* g) {
  if (r2 < p.inner2) {
    double s = p.ka, x = p.alpha2 * r2;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      g[k] = s * (x / (2 * k + 3) - 1.0 / (2 * k + 1));
      s *= 2 * p.alpha2;
    }
  } else {
    double r = sqrt(r2), ir2 = 1.0 / r2;
    double e = p.ka * exp(-p.alpha2 * r2) * ir2;
    g[0] = (hole ? -erf(p.alpha * r) : erfc(p.alpha * r)) / r;
    double s = 1.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      g[k + 1] = (2 * k + 1) * g[k] * ir2 + s * e;
      s *= 2 * p.alpha2;
    }
  }
}

__device__ __forceinline__ void radial32(float r2, bool hole, float alpha, float alpha2, float ka,
                                         float inner2, float* g) {
  if (r2 < inner2) {
    const float inv_odd[7] = {1.0f, 1.0f / 3.0f, 1.0f / 5.0f, 1.0f / 7.0f,
                              1.0f / 9.0f, 1.0f / 11.0f, 1.0f / 13.0f};
    float s = ka, x = alpha2 * r2;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      g[k] = s * (x * inv_odd[k + 1] - inv_odd[k]);
      s *= 2 * alpha2;
    }
  } else {
    float dir = rsqrtf(r2), r = r2 * dir, ir2 = dir * dir;
    float e = ka * expf(-alpha2 * r2) * ir2;
    g[0] = (hole ? -erff(alpha * r) : erfcf(alpha * r)) * dir;
    float s = 1.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      g[k + 1] = (2 * k + 1) * g[k] * ir2 + s * e;
      s *= 2 * alpha2;
    }
  }
}

constexpr int kEwBlock = 256;
constexpr int kEwMaxCand = 7 * 7 * 7;  // nEwReps <= 3 (encodeOffset replica range)

__global__ __launch_bounds__(kEwBlock) void ewald_kernel(const CompactPartData* parts,
                                                         VariablePartData* vars,
                                                         const int* markers, int first, int n,
                                                         const EwtData* ewt, const double* pos64,
                                                         EwaldParams p) {
  // h-loop table: per entry (k, hC) and (h as integers, hS); dynamic size
  extern __shared__ float4 sh[];
  // lattice offsets (n L, hole flag) that some particle of the wavefront needs
  __shared__ float4 cand[kEwBlock / kWave][kEwMaxCand];
  __shared__ int ncand[kEwBlock / kWave];

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const float kscale = (float)(p.L / (2.0 * M_PI));  // k -> integer h
  for (int i = tid; i < p.nH; i += kEwBlock) {
    EwtData e = ewt[i];
    sh[2 * i] = make_float4(e.hx, e.hy, e.hz, e.hCfac);
    sh[2 * i + 1] = make_float4(rintf(e.hx * kscale), rintf(e.hy * kscale), rintf(e.hz * kscale),
                                e.hSfac);
  }

  const int gi = blockIdx.x * kEwBlock + tid;
  const bool live = gi < n;
  const int id = live ? (markers ? markers[gi] : first + gi) : (markers ? markers[n - 1] : first + n - 1);
  double dx, dy, dz;
  if (pos64) {
    dx = pos64[3 * (size_t)id] - p.cmx;
    dy = pos64[3 * (size_t)id + 1] - p.cmy;
    dz = pos64[3 * (size_t)id + 2] - p.cmz;
  } else {
    CompactPartData q = parts[id];
    dx = (double)q.position.x - p.cmx;
    dy = (double)q.position.y - p.cmy;
    dz = (double)q.position.z - p.cmz;
  }
  // wavefront bounding box of d (dead lanes hold a live lane's particle)
  float lo[3] = {(float)dx, (float)dy, (float)dz}, hi[3] = {lo[0], lo[1], lo[2]};
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(lo[k], __shfl_xor(lo[k], m, kWave));
      hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], m, kWave));
    }
  const int W = 2 * p.nEwReps + 1, NC = W * W * W;
  const float Lf = (float)p.L, cut2f = (float)p.cut2;
  if (lane == 0) ncand[wv] = 0;
  __syncthreads();
  for (int c0 = 0; c0 < NC; c0 += kWave) {
    int c = c0 + lane;
    bool need = false;
    if (c < NC) {
      int ix = c / (W * W) - p.nEwReps, iy = (c / W) % W - p.nEwReps, iz = c % W - p.nEwReps;
      bool hole = abs(ix) <= p.nReps && abs(iy) <= p.nReps && abs(iz) <= p.nReps;
      bool centre = ix == 0 && iy == 0 && iz == 0;
      int iv[3] = {ix, iy, iz};
      float d2 = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float l = lo[k] + iv[k] * Lf, h = hi[k] + iv[k] * Lf;
        float e = l > 0.f ? l : (h < 0.f ? -h : 0.f);
        d2 += e * e;
      }
      // slack for FP32 rounding of the bound; the per-particle test is exact
      need = !centre && (hole || d2 <= cut2f * 1.0001f + 1e-6f * Lf * Lf);
    }
    unsigned long long mask = __ballot(need);
    int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (need) {
      int ix = c / (W * W) - p.nEwReps, iy = (c / W) % W - p.nEwReps, iz = c % W - p.nEwReps;
      bool hole = abs(ix) <= p.nReps && abs(iy) <= p.nReps && abs(iz) <= p.nReps;
      cand[wv][ncand[wv] + before] = make_float4(ix * Lf, iy * Lf, iz * Lf, hole ? 1.0f : 0.0f);
    }
    __syncthreads();
    if (lane == 0) ncand[wv] += __popcll(mask);
    __syncthreads();
  }
  const int nc = ncand[wv];

  Mom22<double> Qd;
  Mom22<float> Qf;
  {
    double* qd = &Qd.m;
    float* qf = &Qf.m;
#pragma unroll
    for (int k = 0; k < 22; ++k) {
      qd[k] = p.q[k];
      qf[k] = (float)p.q[k];
    }
  }
  double pot = p.mk1, ax = 0, ay = 0, az = 0;
  // central replica in FP64
  {
    double g[6];
    radial64(dx * dx + dy * dy + dz * dz, p.nReps >= 0, p, g);
    ewald_term<double>(Qd, dx, dy, dz, g, pot, ax, ay, az);
  }
  const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
  const float alpha = (float)p.alpha, alpha2 = (float)p.alpha2, ka = (float)p.ka,
              inner2 = (float)p.inner2, cut2 = (float)p.cut2;
  for (int k = 0; k < nc; ++k) {
    const float4 cv = cand[wv][k];
    const bool hole = cv.w != 0.0f;
    float x = fx + cv.x, y = fy + cv.y, z = fz + cv.z;
    float r2 = x * x + y * y + z * z;
    if (r2 > cut2 && !hole) continue;
    float g[6];
    radial32(r2, hole, alpha, alpha2, ka, inner2, g);
    float tp = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
    ewald_term<float>(Qf, x, y, z, g, tp, tx, ty, tz);
    pot += tp;
    ax += tx;
    ay += ty;
    az += tz;
  }
  // k space: phase h.d / L in revolutions (h integer), reduced exactly, then
  // sin / cos of pi times twice that
  const float iL = (float)(1.0 / p.L);
  const float ux = fx * iL, uy = fy * iL, uz = fz * iL;
  for (int i = 0; i < p.nH; ++i) {
    const float4 e0 = sh[2 * i];
    const float4 e1 = sh[2 * i + 1];
    const float hs = e1.w;
    float ph = e1.x * ux + e1.y * uy + e1.z * uz;
    ph -= rintf(ph);
    float s, c;
    sincospif(2.0f * ph, &s, &c);
    pot += (double)(e0.w * c + hs * s);
    float t = e0.w * s - hs * c;
    ax += (double)(e0.x * t);
    ay += (double)(e0.y * t);
    az += (double)(e0.z * t);
  }
  if (live) {
    VariablePartData v = vars[id];
    v.a.x += (float)ax;
    v.a.y += (float)ay;
    v.a.z += (float)az;
    v.potential += (float)pot;
    vars[id] = v;
  }
}

// Builds the kernel parameters from the reference's read-only record.
static EwaldParams make_params(const EwaldReadOnlyData* ro, const double* rootCm64) {
  EwaldParams p;
  const MomcData& mc = ro->momcRoot;
  const float* m = &mc.m;
  for (int k = 0; k < 22; ++k) p.q[k] = m[k];
  p.cmx = rootCm64 ? rootCm64[0] : (double)ro->mm.cmx;
  p.cmy = rootCm64 ? rootCm64[1] : (double)ro->mm.cmy;
  p.cmz = rootCm64 ? rootCm64[2] : (double)ro->mm.cmz;
  double L = ro->L;
  double alpha = 2.0 / L;
  p.L = L;
  p.alpha = alpha;
  p.alpha2 = alpha * alpha;
  p.ka = 2.0 * alpha / sqrt(M_PI);
  p.mk1 = (double)ro->mm.totalMass * M_PI / (alpha * alpha * L * L * L);
  p.cut2 = (double)ro->fEwCut * (double)ro->fEwCut * L * L;
  p.inner2 = 1.2e-3 * L * L;
  p.nReps = ro->nReps;
  p.nEwReps = ro->nEwReps > ro->nReps ? ro->nEwReps : ro->nReps;
  p.nH = ro->nEwhLoop;
  p.pad = 0;
  return p;
}

int launch_ewald(const CompactPartData* parts, VariablePartData* vars, const int* markers,
                 int first, int n, const EwaldReadOnlyData* ro, const EwtData* d_ewt,
                 const double* pos64, const double* rootCm64, hipStream_t s) {
  if (n <= 0) return 0;
  EwaldParams p = make_params(ro, rootCm64);
  if (p.nEwReps > 3)
    return fail("Ewald: nEwReps > 3", hipErrorInvalidValue, __FILE__, __LINE__);
  size_t shm = sizeof(float4) * 2 * (size_t)(p.nH > 0 ? p.nH : 1);
  int blocks = (n + kEwBlock - 1) / kEwBlock;
  hipLaunchKernelGGL(ewald_kernel, dim3(blocks), dim3(kEwBlock), shm, s, parts, vars, markers,
                     first, n, d_ewt, pos64, p);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

}  // namespace cgpu
