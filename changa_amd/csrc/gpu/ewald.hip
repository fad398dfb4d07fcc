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
#include "ewald_body.h"

namespace cgpu {

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

  // the same table as weights on the cube |h_i| <= kEwCube (zero off the table)
  __shared__ float2 cube[kEwCubeN];
  __shared__ int n_outside;

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const float kscale = (float)(p.L / (2.0 * M_PI));  // k -> integer h
  for (int i = tid; i < kEwCubeN; i += kEwBlock) cube[i] = make_float2(0.0f, 0.0f);
  if (tid == 0) n_outside = 0;
  __syncthreads();
  for (int i = tid; i < p.nH; i += kEwBlock) {
    EwtData e = ewt[i];
    const float hx = rintf(e.hx * kscale), hy = rintf(e.hy * kscale), hz = rintf(e.hz * kscale);
    sh[2 * i] = make_float4(e.hx, e.hy, e.hz, e.hCfac);
    sh[2 * i + 1] = make_float4(hx, hy, hz, e.hSfac);
    if (fabsf(hx) <= kEwCube && fabsf(hy) <= kEwCube && fabsf(hz) <= kEwCube) {
      const int ci = ((int)hx + kEwCube) * kEwCubeW * kEwCubeW + ((int)hy + kEwCube) * kEwCubeW +
                     (int)hz + kEwCube;
      cube[ci] = make_float2(e.hCfac, e.hSfac);
    } else {
      atomicAdd(&n_outside, 1);
    }
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
  // k space. On the cube |h_i| <= kEwCube: e^(2 pi i h.u) (u = d / L) as the
  // product of per-axis powers of e^(2 pi i u_x) etc., which are formed once
  // per particle; the plane hx is advanced by one complex multiplication.
  // Compensated FP32 sums within a plane (the residual c is subtracted when
  // the plane's sum is carried into FP64), FP64 across planes.
  const float iL = (float)(1.0 / p.L);
  const float ux = fx * iL, uy = fy * iL, uz = fz * iL;
  {
    float2 ey[kEwCubeW], ez[kEwCubeW];
    float2 e1x;
    sincospif(2.0f * ux, &e1x.y, &e1x.x);
    axis_powers(uy, ey);
    axis_powers(uz, ez);
    float2 ex = cpow_neg(e1x);  // e^(-2 pi i kEwCube u_x)
    const float kf = (float)(2.0 * M_PI / p.L);
    for (int hx = -kEwCube; hx <= kEwCube; ++hx) {
      KahanF sp, sx, sy, sz;
      const float2* cw = cube + (hx + kEwCube) * kEwCubeW * kEwCubeW;
#pragma unroll
      for (int hy = -kEwCube; hy <= kEwCube; ++hy) {
        const float2 exy = cmul(ex, ey[hy + kEwCube]);
#pragma unroll
        for (int hz = -kEwCube; hz <= kEwCube; ++hz) {
          const float2 wgt = cw[(hy + kEwCube) * kEwCubeW + hz + kEwCube];
          if (wgt.x != 0.0f || wgt.y != 0.0f) {
            const float2 e = cmul(exy, ez[hz + kEwCube]);
            sp.add(wgt.x * e.x + wgt.y * e.y);
            const float t = wgt.x * e.y - wgt.y * e.x;
            sx.add((float)hx * t);
            sy.add((float)hy * t);
            sz.add((float)hz * t);
          }
        }
      }
      pot += (double)sp.s - (double)sp.c;
      ax += (double)kf * ((double)sx.s - (double)sx.c);
      ay += (double)kf * ((double)sy.s - (double)sy.c);
      az += (double)kf * ((double)sz.s - (double)sz.c);
      ex = cmul(ex, e1x);
    }
  }
  // entries off the cube (a table cut beyond |h_i| = kEwCube): phase h.d / L
  // in revolutions (h integer), reduced exactly, then sin / cos of 2 pi that
  if (n_outside > 0) {
    for (int i = 0; i < p.nH; ++i) {
      const float4 e0 = sh[2 * i];
      const float4 e1 = sh[2 * i + 1];
      if (fabsf(e1.x) <= kEwCube && fabsf(e1.y) <= kEwCube && fabsf(e1.z) <= kEwCube) continue;
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
