// Device code of the Ewald correction shared by the forward kernel
// (ewald.hip) and its adjoint (adjoint.hip): parameters, the radial kernels,
// the multipole contraction of one replica term, the k-space product form.
#pragma once
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

// V_n = Q_n . r^(n-1) (vectors) and P_n = Q_n . r^n of the root's traceless
// moments for n = 2, 3, 4. S: the type of the moments; T: the type of r.
template <typename T>
struct EwaldVP {
  T v2x, v2y, v2z, v3x, v3y, v3z, v4x, v4y, v4z, p2, p3, p4;
};

template <typename S, typename T>
__device__ __forceinline__ EwaldVP<T> ewald_vp(const Mom22<S>& Q, T x, T y, T z) {
  EwaldVP<T> o;
  T x2 = x * x, y2 = y * y, z2 = z * z;
  T fxx = x2 - z2, fyy = y2 - z2;
  T fxy = S(2) * x * y, fxz = S(2) * x * z, fyz = S(2) * y * z;
  T z3 = S(3) * z2;
  T fxxx = x * (x2 - z3), fyyy = y * (y2 - z3);
  T fxyy = S(3) * x * fyy, fxxy = S(3) * y * fxx;
  T fxxz = z * (S(3) * x2 - z2), fyyz = z * (S(3) * y2 - z2);
  T fxyz = S(3) * fxy * z;
  o.v2x = Q.xx * x + Q.xy * y + Q.xz * z;
  o.v2y = Q.xy * x + Q.yy * y + Q.yz * z;
  o.v2z = Q.xz * x + Q.yz * y - (Q.xx + Q.yy) * z;
  o.v3x = Q.xxx * fxx + Q.xyy * fyy + Q.xxy * fxy + Q.xxz * fxz + Q.xyz * fyz;
  o.v3y = Q.xyy * fxy + Q.xxy * fxx + Q.yyy * fyy + Q.yyz * fyz + Q.xyz * fxz;
  o.v3z = Q.xxz * fxx + Q.yyz * fyy + Q.xyz * fxy - (Q.xxx + Q.xyy) * fxz - (Q.xxy + Q.yyy) * fyz;
  o.v4x = Q.xxxx * fxxx + Q.xxyy * fxyy + Q.xxxy * fxxy + Q.xyyy * fyyy + Q.xxxz * fxxz +
          Q.xyyz * fyyz + Q.xxyz * fxyz;
  o.v4y = Q.yyyy * fyyy + Q.xxyy * fxxy + Q.xxxy * fxxx + Q.xyyy * fxyy + Q.xyyz * fxyz +
          Q.xxyz * fxxz + Q.yyyz * fyyz;
  o.v4z = Q.xxxz * fxxx + Q.xyyz * fxyy + Q.xxyz * fxxy + Q.yyyz * fyyy - Q.xxxx * fxxz -
          Q.yyyy * fyyz - Q.xxyy * (fxxz + fyyz) - (Q.xxxy + Q.xyyy) * fxyz;
  o.p2 = o.v2x * x + o.v2y * y + o.v2z * z;
  o.p3 = o.v3x * x + o.v3y * y + o.v3z * z;
  o.p4 = o.v4x * x + o.v4y * y + o.v4z * z;
  return o;
}

// Ewald contribution of one replica vector r with radial kernels g[0..5]:
//   pot -= sum_n (1/n!) g_n P_n,
//   a   += sum_n (1/n!) [n g_n V_n - g_(n+1) P_n r]   (n = 0, 2, 3, 4).
// S: the type of the moments and constants; T: the type of r, g and the
// results (S itself, or a dual number carrying a derivative).
template <typename S, typename T>
__device__ __forceinline__ void ewald_term(const Mom22<S>& Q, T x, T y, T z, const T* g, T& pot,
                                           T& ax, T& ay, T& az) {
  const EwaldVP<T> v = ewald_vp<S, T>(Q, x, y, z);
  const S h = S(0.5), s = S(1) / S(6), f = S(1) / S(24);
  pot = pot - (g[0] * Q.m + g[2] * h * v.p2 + g[3] * s * v.p3 + g[4] * f * v.p4);
  T ta = g[1] * Q.m + g[3] * h * v.p2 + g[4] * s * v.p3 + g[5] * f * v.p4;
  T c3 = g[3] * h, c4 = g[4] * s;
  ax = ax + (g[2] * v.v2x + c3 * v.v3x + c4 * v.v4x - ta * x);
  ay = ay + (g[2] * v.v2y + c3 * v.v3y + c4 * v.v4y - ta * y);
  az = az + (g[2] * v.v2z + c3 * v.v3z + c4 * v.v4z - ta * z);
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

constexpr int kEwCube = 3;  // k-space table entries with |h_i| <= 3 use the product form
constexpr int kEwCubeW = 2 * kEwCube + 1;
constexpr int kEwCubeN = kEwCubeW * kEwCubeW * kEwCubeW;

// Compensated (Kahan) FP32 sum: the rounding error of each addition is
// carried in c and fed back into the next.
struct KahanF {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// e^(2 pi i m u) for m = -kEwCube..kEwCube into e[m + kEwCube]
__device__ __forceinline__ void axis_powers(float u, float2* e) {
  float2 e1;
  sincospif(2.0f * u, &e1.y, &e1.x);
  e[kEwCube] = make_float2(1.0f, 0.0f);
#pragma unroll
  for (int m = 1; m <= kEwCube; ++m) {
    const float2 v = cmul(e[kEwCube + m - 1], e1);
    e[kEwCube + m] = v;
    e[kEwCube - m] = make_float2(v.x, -v.y);
  }
}

// e1^(-kEwCube)
__device__ __forceinline__ float2 cpow_neg(float2 e1) {
  float2 v = make_float2(1.0f, 0.0f);
#pragma unroll
  for (int m = 0; m < kEwCube; ++m) v = cmul(v, e1);
  return make_float2(v.x, -v.y);
}

static inline EwaldParams make_params(const EwaldReadOnlyData* ro, const double* rootCm64) {
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

}  // namespace cgpu
