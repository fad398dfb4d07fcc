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
                                                         const double* root, EwaldParams p) {
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

  // root centre of mass, k = 0 term and moments: from the device root record
  // (changaEwaldRoot) when there is one, else from the parameters
  const double cmx = root ? root[0] : p.cmx, cmy = root ? root[1] : p.cmy,
               cmz = root ? root[2] : p.cmz;
  const double mk1 = root ? root[3] * M_PI / (p.alpha2 * p.L * p.L * p.L) : p.mk1;
  const int gi = blockIdx.x * kEwBlock + tid;
  const bool live = gi < n;
  const int id = live ? (markers ? markers[gi] : first + gi) : (markers ? markers[n - 1] : first + n - 1);
  double dx, dy, dz;
  if (pos64) {
    dx = pos64[3 * (size_t)id] - cmx;
    dy = pos64[3 * (size_t)id + 1] - cmy;
    dz = pos64[3 * (size_t)id + 2] - cmz;
  } else {
    CompactPartData q = parts[id];
    dx = (double)q.position.x - cmx;
    dy = (double)q.position.y - cmy;
    dz = (double)q.position.z - cmz;
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
      qd[k] = root ? root[4 + k] : p.q[k];
      qf[k] = (float)qd[k];
    }
  }
  double pot = mk1, ax = 0, ay = 0, az = 0;
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

// Builds the kernel parameters from the reference's read-only record; with
// d_root (a device record of changaEwaldRoot) the root's centre of mass, mass
// and moments come from there instead of ro and rootCm64.
int launch_ewald(const CompactPartData* parts, VariablePartData* vars, const int* markers,
                 int first, int n, const EwaldReadOnlyData* ro, const EwtData* d_ewt,
                 const double* pos64, const double* rootCm64, const double* d_root,
                 hipStream_t s) {
  if (n <= 0) return 0;
  EwaldParams p = make_params(ro, rootCm64);
  if (p.nEwReps > 3)
    return fail("Ewald: nEwReps > 3", hipErrorInvalidValue, __FILE__, __LINE__);
  size_t shm = sizeof(float4) * 2 * (size_t)(p.nH > 0 ? p.nH : 1);
  int blocks = (n + kEwBlock - 1) / kEwBlock;
  hipLaunchKernelGGL(ewald_kernel, dim3(blocks), dim3(kEwBlock), shm, s, parts, vars, markers,
                     first, n, d_ewt, pos64, d_root, p);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// EwaldInit on the device (Ewald.cpp:285-383): the root record (centre of
// mass, mass, the 22 reduced moments brought to unit scale, the first 22
// MOMC components of momRescaleFmomr + momFmomr2Momc) and the h-loop table,
// from the root's node record node64[0] (cm.xyz, radius, soft, mass,
// lo.xyz, hi.xyz, 22 radius-scaled reduced moments). One block: each thread
// takes one lattice vector h of the cube |h_i| <= hreps, the valid ones
// (0 < |h|^2 <= hcut^2) are numbered in the host table's order (hx
// outermost, hz innermost) by a block scan, and each writes its entry:
//   k = 2 pi h / L,  w = exp(-k^2 / 4 alpha^2) / (pi h^2 L),  alpha = 2 / L,
//   hC = -w (M - (Q2:kk)/2 + (Q4:kkkk)/24),  hS = w (Q3:kkk)/6.
// ---------------------------------------------------------------------------
constexpr int kRootBlock = 512;

__global__ __launch_bounds__(kRootBlock) void ewald_root_kernel(const double* node64, double L,
                                                                double hcut2, int hreps,
                                                                double* root, EwtData* ewt) {
  __shared__ double q[22];
  __shared__ int wave_count[kRootBlock / kWave];
  __shared__ int base;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  if (tid < 22) {
    const int order = tid == 0 ? 0 : (tid < 6 ? 2 : (tid < 13 ? 3 : 4));
    const double u = node64[3];
    double s = 1.0;
    for (int k = 0; k < order; ++k) s *= u;
    q[tid] = node64[12 + tid] * s;
    root[4 + tid] = q[tid];
  }
  if (tid < 3) root[tid] = node64[tid];
  if (tid == 3) root[3] = node64[5];
  if (tid == 0) base = 0;
  __syncthreads();
  const double m = q[0], xx = q[1], yy = q[2], xy = q[3], xz = q[4], yz = q[5];
  const double zz = -(xx + yy);
  const double xxx = q[6], xyy = q[7], xxy = q[8], yyy = q[9], xxz = q[10], yyz = q[11],
               xyz = q[12];
  const double xzz = -(xxx + xyy), yzz = -(xxy + yyy), zzz = -(xxz + yyz);
  const double xxxx = q[13], xyyy = q[14], xxxy = q[15], yyyy = q[16], xxxz = q[17],
               yyyz = q[18], xxyy = q[19], xxyz = q[20], xyyz = q[21];
  const double xxzz = -(xxxx + xxyy), xyzz = -(xxxy + xyyy), xzzz = -(xxxz + xyyz),
               yyzz = -(xxyy + yyyy), yzzz = -(xxyz + yyyz), zzzz = -(xxzz + yyzz);
  const double alpha = 2.0 / L;
  const double k4 = M_PI * M_PI / (alpha * alpha * L * L);
  const double f = 2.0 * M_PI / L;
  const int W = 2 * hreps + 1, NC = W * W * W;
  for (int c0 = 0; c0 < NC; c0 += kRootBlock) {
    const int c = c0 + tid;
    const int hx = c / (W * W) - hreps, hy = (c / W) % W - hreps, hz = c % W - hreps;
    const int h2 = hx * hx + hy * hy + hz * hz;
    const bool valid = c < NC && h2 != 0 && (double)h2 <= hcut2;
    const unsigned long long mask = __ballot(valid);
    if (lane == 0) wave_count[wv] = __popcll(mask);
    __syncthreads();
    int idx = base + __popcll(mask & ((1ull << lane) - 1ull));
    for (int k = 0; k < wv; ++k) idx += wave_count[k];
    if (valid) {
      const double x = f * hx, y = f * hy, z = f * hz;
      const double w = exp(-k4 * h2) / (M_PI * h2 * L);
      const double q2 = xx * x * x + yy * y * y + zz * z * z +
                        2 * (xy * x * y + xz * x * z + yz * y * z);
      const double q3 = xxx * x * x * x + yyy * y * y * y + zzz * z * z * z +
                        3 * (xxy * x * x * y + xxz * x * x * z + xyy * x * y * y +
                             yyz * y * y * z + xzz * x * z * z + yzz * y * z * z) +
                        6 * xyz * x * y * z;
      const double q4 = xxxx * x * x * x * x + yyyy * y * y * y * y + zzzz * z * z * z * z +
                        4 * (xxxy * x * x * x * y + xxxz * x * x * x * z + xyyy * x * y * y * y +
                             yyyz * y * y * y * z + xzzz * x * z * z * z + yzzz * y * z * z * z) +
                        6 * (xxyy * x * x * y * y + xxzz * x * x * z * z + yyzz * y * y * z * z) +
                        12 * (xxyz * x * x * y * z + xyyz * x * y * y * z + xyzz * x * y * z * z);
      EwtData e;
      e.hx = (float)x;
      e.hy = (float)y;
      e.hz = (float)z;
      e.hCfac = (float)(-w * (m - 0.5 * q2 + q4 / 24.0));
      e.hSfac = (float)(w * q3 / 6.0);
      ewt[idx] = e;
    }
    __syncthreads();
    if (tid == 0)
      for (int k = 0; k < kRootBlock / kWave; ++k) base += wave_count[k];
    __syncthreads();
  }
}

int launch_ewald_root(const double* node64, const EwaldReadOnlyData* ro, double hcut,
                      double* root, EwtData* ewt, hipStream_t s) {
  if (!node64 || !ro || !root || (!ewt && ro->nEwhLoop > 0) || !(hcut >= 0.0) || ro->L <= 0.0f)
    return fail("changaEwaldRoot: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  const int hreps = (int)ceil(hcut);
  const double hcut2 = hcut * hcut;
  int count = 0;
  for (int hx = -hreps; hx <= hreps; ++hx)
    for (int hy = -hreps; hy <= hreps; ++hy)
      for (int hz = -hreps; hz <= hreps; ++hz) {
        const int h2 = hx * hx + hy * hy + hz * hz;
        count += h2 != 0 && (double)h2 <= hcut2;
      }
  if (count != ro->nEwhLoop)
    return fail("changaEwaldRoot: ro->nEwhLoop is not the number of h with 0 < |h|^2 <= hCut^2",
                hipErrorInvalidValue, __FILE__, __LINE__);
  hipLaunchKernelGGL(ewald_root_kernel, dim3(1), dim3(kRootBlock), 0, s, node64, (double)ro->L,
                     hcut2, hreps, root, ewt);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

}  // namespace cgpu
