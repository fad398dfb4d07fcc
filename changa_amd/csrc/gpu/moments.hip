// GPU multipole build: bucket P2M and level-synchronous M2M, in FP64.
//
// CPU semantics followed: GenericTreeNode::makeBucket (GenericTreeNode.h:
// 219-257) with MultipoleMoments::operator+=(particle) (MultipoleMoments.h:
// 276-352) for buckets, LocalTreeBuilder::doneChildren (Compute.cpp:
// 2534-2580) with operator+=(MultipoleMoments) (MultipoleMoments.h:207-272)
// for internal nodes, and calculateRadiusFarthestParticle /
// FarthestCorner (MultipoleMoments.h:460-505) for the expansion radius. The
// arithmetic kernels of the reference are momMakeFmomr, momShiftFmomr,
// momScaledAddFmomr and momRescaleFmomr (moments.c:202-268, 772-846,
// 1002-1058).
//
// Formulation. Moments are kept as "complete" symmetric components
// C(a,b,c) (a x's, b y's, c z's; orders 0 and 2..4, 35 slots with order 1
// unused). A bucket accumulates the raw moments sum_i m_i dx^a dy^b dz^c
// about its centre of mass and projects them once onto the symmetric
// trace-free (STF) subspace; the incremental per-particle shifting of the CPU
// build is the same sum. An internal node shifts each child's STF moments by
// the binomial rule for raw moments, R(a,b,c) = sum C(a,a1) C(b,b1) C(c,c1)
// Q(a1,b1,c1) d^(a-a1, b-b1, c-c1) (the dipole about a centre of mass is
// zero), adds the two children and projects onto the STF subspace again
// (trace parts of lower orders vanish under the projection). The shift runs
// axis by axis (moment_shift_axis), 105 FMAs per child. Stored moments
// are divided by radius^order (radius-scaled reduced moments).
#include "stf.h"

namespace cgpu {

// reduced (FMOMR order) <-> complete
__device__ __forceinline__ void complete_from_reduced(const double* q, double* C) {
  C[0] = q[0];
  C[1] = C[2] = C[3] = 0.0;
  C[cidx(2, 0, 0)] = q[1]; C[cidx(0, 2, 0)] = q[2]; C[cidx(1, 1, 0)] = q[3];
  C[cidx(1, 0, 1)] = q[4]; C[cidx(0, 1, 1)] = q[5];
  C[cidx(3, 0, 0)] = q[6]; C[cidx(1, 2, 0)] = q[7]; C[cidx(2, 1, 0)] = q[8];
  C[cidx(0, 3, 0)] = q[9]; C[cidx(2, 0, 1)] = q[10]; C[cidx(0, 2, 1)] = q[11];
  C[cidx(1, 1, 1)] = q[12];
  C[cidx(4, 0, 0)] = q[13]; C[cidx(1, 3, 0)] = q[14]; C[cidx(3, 1, 0)] = q[15];
  C[cidx(0, 4, 0)] = q[16]; C[cidx(3, 0, 1)] = q[17]; C[cidx(0, 3, 1)] = q[18];
  C[cidx(2, 2, 0)] = q[19]; C[cidx(2, 1, 1)] = q[20]; C[cidx(1, 2, 1)] = q[21];
  // z-heavy components from tracelessness, in increasing c
#pragma unroll
  for (int n = 2; n <= 4; ++n)
#pragma unroll
    for (int c = 2; c <= n; ++c)
#pragma unroll
      for (int a = 0; a <= n - c; ++a) {
        const int b = n - c - a;
        C[cidx(a, b, c)] = -(C[cidx(a + 2, b, c - 2)] + C[cidx(a, b + 2, c - 2)]);
      }
}

__device__ __forceinline__ void reduced_from_complete(const double* C, double* q) {
  q[0] = C[0];
  q[1] = C[cidx(2, 0, 0)]; q[2] = C[cidx(0, 2, 0)]; q[3] = C[cidx(1, 1, 0)];
  q[4] = C[cidx(1, 0, 1)]; q[5] = C[cidx(0, 1, 1)];
  q[6] = C[cidx(3, 0, 0)]; q[7] = C[cidx(1, 2, 0)]; q[8] = C[cidx(2, 1, 0)];
  q[9] = C[cidx(0, 3, 0)]; q[10] = C[cidx(2, 0, 1)]; q[11] = C[cidx(0, 2, 1)];
  q[12] = C[cidx(1, 1, 1)];
  q[13] = C[cidx(4, 0, 0)]; q[14] = C[cidx(1, 3, 0)]; q[15] = C[cidx(3, 1, 0)];
  q[16] = C[cidx(0, 4, 0)]; q[17] = C[cidx(3, 0, 1)]; q[18] = C[cidx(0, 3, 1)];
  q[19] = C[cidx(2, 2, 0)]; q[20] = C[cidx(2, 1, 1)]; q[21] = C[cidx(1, 2, 1)];
}

// Raw moments about a centre displaced by -d along one axis, in place:
// M'(a) = sum_{a1 <= a} binom(a, a1) d^(a - a1) M(a1) on every line of fixed
// exponents of the other two axes. This is the transpose of the Taylor
// shift's repeated synthetic division: n (n + 1) / 2 FMAs on a line of
// degree n, 35 per axis (the direct binomial sum over all three axes takes
// ~1,500 operations). Order-1 slots take part: they are zero about the
// child's centre of mass but not after the first axis.
template <int AXIS>
__device__ __forceinline__ void moment_shift_axis(double* P, double d) {
#pragma unroll
  for (int u = 0; u <= 4; ++u)
#pragma unroll
    for (int v = 0; v <= 4 - u; ++v) {
      const int n = 4 - u - v;  // degree along AXIS on this line
#pragma unroll
      for (int k = n - 1; k >= 0; --k)
#pragma unroll
        for (int e = k; e < n; ++e) {
          const int lo = AXIS == 0 ? cidx(e, u, v) : AXIS == 1 ? cidx(u, e, v) : cidx(u, v, e);
          const int hi = AXIS == 0 ? cidx(e + 1, u, v)
                                   : AXIS == 1 ? cidx(u, e + 1, v) : cidx(u, v, e + 1);
          P[hi] = fma(d, P[lo], P[hi]);
        }
    }
}

// write node outputs: node64 record, wire moments, cm64
__device__ __forceinline__ void emit(int nd, const double* g, const double* C, double u,
                                     CudaMultipoleMoments* out, double* node64, double* cm64) {
  double q[22];
  reduced_from_complete(C, q);
  const double iu = 1.0 / u;
  const double s2 = iu * iu, s3 = s2 * iu, s4 = s3 * iu;
#pragma unroll
  for (int k = 1; k < 22; ++k) q[k] *= (k < 6 ? s2 : k < 13 ? s3 : s4);
  double* rec = node64 + (size_t)nd * 34;
#pragma unroll
  for (int k = 0; k < 12; ++k) rec[k] = g[k];
  rec[3] = u;
#pragma unroll
  for (int k = 0; k < 22; ++k) rec[12 + k] = q[k];
  if (cm64) {
    cm64[3 * (size_t)nd] = g[0];
    cm64[3 * (size_t)nd + 1] = g[1];
    cm64[3 * (size_t)nd + 2] = g[2];
  }
  CudaMultipoleMoments m;
  m.radius = (float)u;
  m.soft = (float)g[4];
  m.totalMass = (float)g[5];
  m.cm.x = (float)g[0];
  m.cm.y = (float)g[1];
  m.cm.z = (float)g[2];
  m.xx = (float)q[1]; m.xy = (float)q[3]; m.xz = (float)q[4]; m.yy = (float)q[2]; m.yz = (float)q[5];
  m.xxx = (float)q[6]; m.xyy = (float)q[7]; m.xxy = (float)q[8]; m.yyy = (float)q[9];
  m.xxz = (float)q[10]; m.yyz = (float)q[11]; m.xyz = (float)q[12];
  m.xxxx = (float)q[13]; m.xyyy = (float)q[14]; m.xxxy = (float)q[15]; m.yyyy = (float)q[16];
  m.xxxz = (float)q[17]; m.yyyz = (float)q[18]; m.xxyy = (float)q[19]; m.xxyz = (float)q[20];
  m.xyyz = (float)q[21];
  out[nd] = m;
}

// Sum of v over the 4 lanes of a bucket's group (lanes 4g..4g+3), by an xor
// butterfly: every lane of the group gets the same bits ((v0 + v1) + (v2 + v3)
// up to commutation).
__device__ __forceinline__ double group4_sum(double v) {
  v += __shfl_xor(v, 1);
  return v + __shfl_xor(v, 2);
}
__device__ __forceinline__ double group4_min(double v) {
  v = fmin(v, __shfl_xor(v, 1));
  return fmin(v, __shfl_xor(v, 2));
}
__device__ __forceinline__ double group4_max(double v) {
  v = fmax(v, __shfl_xor(v, 1));
  return fmax(v, __shfl_xor(v, 2));
}

constexpr int kMomBlock = 256;

// Buckets (leaf nodes). A block takes kMomBlock consecutive nodes, compacts
// its leaves into LDS, and gives each leaf a group of 4 lanes: lane j of the
// group takes particles first + j, first + j + 4, ..., so each group reads its
// bucket's particles as contiguous 4-particle runs and every cache line it
// touches is used up within about one step (one lane per bucket, walking its
// own particles, kept ~10^5 partly read lines live at once and fetched each
// line several times over). Pass 1: mass, mass-weighted position and
// softening, bounding box; pass 2 (same particles, now in cache): the raw
// moments about the centre of mass and the farthest particle, summed over
// the group. One lane per bucket then projects and writes the records.
__global__ __launch_bounds__(kMomBlock) void bucket_moments_kernel(
    int nNodes, const int* first, const int* count, const int* child, const double* boxRadius,
    const double* pos, const double* mass, const double* soft, CudaMultipoleMoments* out,
    double* node64, double* cm64) {
  __shared__ int leaves[kMomBlock];
  __shared__ int wave_base[kMomBlock / kWave + 1];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int nd0 = blockIdx.x * kMomBlock + tid;
  const bool leaf = nd0 < nNodes && child[2 * (size_t)nd0] < 0;
  const unsigned long long bal = __ballot(leaf);
  if (lane == 0) wave_base[wv + 1] = __popcll(bal);
  __syncthreads();
  if (tid == 0) {
    wave_base[0] = 0;
    for (int w = 1; w <= kMomBlock / kWave; ++w) wave_base[w] += wave_base[w - 1];
  }
  __syncthreads();
  if (leaf) leaves[wave_base[wv] + __popcll(bal & ((1ull << lane) - 1ull))] = nd0;
  __syncthreads();
  const int nleaf = wave_base[kMomBlock / kWave];
  const int sub = tid & 3;
  // a group's 4 lanes run the same trip counts, so its shuffles always pair
  // active lanes
  for (int j = tid >> 2; j < nleaf; j += kMomBlock / 4) {
    const int nd = leaves[j];
    const int f = first[nd], cnt = count[nd];
    double M = 0, sx = 0, sy = 0, sz = 0, ss = 0;
    double lo0 = HUGE_VAL, lo1 = HUGE_VAL, lo2 = HUGE_VAL;
    double hi0 = -HUGE_VAL, hi1 = -HUGE_VAL, hi2 = -HUGE_VAL;
    for (int i = f + sub; i < f + cnt; i += 4) {
      const double m = mass[i], x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1],
                   z = pos[3 * (size_t)i + 2];
      M += m;
      sx += m * x;
      sy += m * y;
      sz += m * z;
      ss += m * soft[i];
      lo0 = fmin(lo0, x); lo1 = fmin(lo1, y); lo2 = fmin(lo2, z);
      hi0 = fmax(hi0, x); hi1 = fmax(hi1, y); hi2 = fmax(hi2, z);
    }
    double g[12];
    M = group4_sum(M);
    g[0] = group4_sum(sx) / M;
    g[1] = group4_sum(sy) / M;
    g[2] = group4_sum(sz) / M;
    g[4] = group4_sum(ss) / M;
    g[5] = M;
    g[6] = group4_min(lo0); g[7] = group4_min(lo1); g[8] = group4_min(lo2);
    g[9] = group4_max(hi0); g[10] = group4_max(hi1); g[11] = group4_max(hi2);
    double T[35];
#pragma unroll
    for (int k = 0; k < 35; ++k) T[k] = 0.0;
    double r2max = 0.0;
    for (int i = f + sub; i < f + cnt; i += 4) {
      const double m = mass[i];
      const double dx = pos[3 * (size_t)i] - g[0], dy = pos[3 * (size_t)i + 1] - g[1],
                   dz = pos[3 * (size_t)i + 2] - g[2];
      r2max = fmax(r2max, dx * dx + dy * dy + dz * dz);
      double px[5], py[5], pz[5];
      px[0] = m;
      py[0] = pz[0] = 1.0;
#pragma unroll
      for (int k = 1; k < 5; ++k) {
        px[k] = px[k - 1] * dx;
        py[k] = py[k - 1] * dy;
        pz[k] = pz[k - 1] * dz;
      }
#pragma unroll
      for (int n = 2; n <= 4; ++n)
#pragma unroll
        for (int a = 0; a <= n; ++a)
#pragma unroll
          for (int b = 0; b <= n - a; ++b) {
            const int c = n - a - b;
            T[cidx(a, b, c)] += px[a] * py[b] * pz[c];
          }
    }
    T[0] = M;
#pragma unroll
    for (int n = 2; n <= 4; ++n)
#pragma unroll
      for (int a = 0; a <= n; ++a)
#pragma unroll
        for (int b = 0; b <= n - a; ++b) {
          const int k = cidx(a, b, n - a - b);
          T[k] = group4_sum(T[k]);
        }
    r2max = group4_max(r2max);
    if (sub == 0) {
      stf_project(T);
      const double u = (cnt > 1 && r2max > 0.0) ? sqrt(r2max) : boxRadius[nd];
      g[3] = u;
      emit(nd, g, T, u, out, node64, cm64);
    }
  }
}

__global__ __launch_bounds__(128) void internal_moments_kernel(int lo, int hi, const int* child,
                                                               CudaMultipoleMoments* out,
                                                               double* node64, double* cm64) {
  int nd = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= hi) return;
  const int c0 = child[2 * nd], c1 = child[2 * nd + 1];
  if (c0 < 0) return;
  const double* r0 = node64 + (size_t)c0 * 34;
  const double* r1 = node64 + (size_t)c1 * 34;
  double g[12];
  const double M0 = r0[5], M1 = r1[5], M = M0 + M1;
  g[5] = M;
  g[4] = (M0 * r0[4] + M1 * r1[4]) / M;
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = (M0 * r0[k] + M1 * r1[k]) / M;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    g[6 + k] = fmin(r0[6 + k], r1[6 + k]);
    g[9 + k] = fmax(r0[9 + k], r1[9 + k]);
  }
  double d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double e1 = g[k] - g[6 + k], e2 = g[9 + k] - g[k];
    double e = e1 > e2 ? e1 : e2;
    d2 += e * e;
  }
  const double u = sqrt(d2);
  g[3] = u;
  double R[35];
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    const double* r = ch ? r1 : r0;
    double q[22], C[35];
    const double uc = r[3], s2 = uc * uc, s3 = s2 * uc, s4 = s3 * uc;
#pragma unroll
    for (int k = 0; k < 22; ++k) q[k] = r[12 + k] * (k == 0 ? 1.0 : k < 6 ? s2 : k < 13 ? s3 : s4);
    complete_from_reduced(q, C);
    moment_shift_axis<0>(C, r[0] - g[0]);
    moment_shift_axis<1>(C, r[1] - g[1]);
    moment_shift_axis<2>(C, r[2] - g[2]);
#pragma unroll
    for (int k = 0; k < 35; ++k) R[k] = ch ? R[k] + C[k] : C[k];
  }
  stf_project(R);
  emit(nd, g, R, u, out, node64, cm64);
}

}  // namespace cgpu

using namespace cgpu;

extern "C" int changaBuildMoments(int nNodes, int nDepths, const int* levelOffsets,
                                  const int* first, const int* count, const int* child,
                                  const double* boxRadius, const double* pos, const double* mass,
                                  const double* soft, CudaMultipoleMoments* moments,
                                  double* node64, double* cm64, void* stream) {
  if (nNodes <= 0) return 0;
  if (!levelOffsets || !first || !count || !child || !boxRadius || !pos || !mass || !soft ||
      !moments || !node64)
    return fail("changaBuildMoments: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(bucket_moments_kernel, dim3((nNodes + kMomBlock - 1) / kMomBlock),
                     dim3(kMomBlock), 0, s, nNodes,
                     first, count, child, boxRadius, pos, mass, soft, moments, node64, cm64);
  CGPU_CHECK(hipGetLastError());
  for (int d = nDepths - 1; d >= 0; --d) {
    int lo = levelOffsets[d], hi = levelOffsets[d + 1];
    if (hi <= lo) continue;
    hipLaunchKernelGGL(internal_moments_kernel, dim3((hi - lo + 127) / 128), dim3(128), 0, s, lo,
                       hi, child, moments, node64, cm64);
    CGPU_CHECK(hipGetLastError());
  }
  return 0;
}
