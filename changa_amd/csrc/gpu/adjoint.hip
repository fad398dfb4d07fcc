// Adjoint of the particle-cell and particle-particle evaluation: for
// L = sum_i w_i . a_i on fixed interaction lists, dL/dx and dL/dm of every
// particle. No counterpart in the reference.
//
// Terms (s = x_k - cm of a cell, r = x_i - cm - replica offset):
//   - target side: d(w_i . a_ic)/dx_i = (da_ic/dr) w_i, the directional
//     derivative of the hexadecapole acceleration along w_i (forward-mode
//     dual numbers through the formula of the forward kernel), and the
//     pairwise Jacobian (b I - c d d^T) w_i for particles and softened cells;
//   - particle sources: the opposite Jacobian term and w_i . d b;
//   - cell sources: per cell, D_alpha = sum_i w_i . grad d^alpha(1/r)(r_i)
//     over every target i that has the cell in its list, for the stored
//     (reduced) multi-indices of orders 0..5. With h_alpha = (-1)^|alpha| /
//     alpha! D_alpha (orders 0, 2, 3, 4; D completed by tracelessness),
//     H(s) = sum h_alpha s^alpha, the centre-of-mass coupling
//     U = (sum_k m_k grad H(s_k) + sum_i (da_ic/dr) w_i) / M reduces to
//     U_j = D_(e_j) + (sum_|alpha|=4 Q_alpha D_(alpha+e_j) / alpha! + Vs_j) / M
//     (Q the moments about cm, Vs the target-side sum of softened pairs), and
//     with Hhat(s) = H(s) - U . s the cell contributes dL/dm_k = Hhat(s_k),
//     dL/dx_k = m_k grad Hhat(s_k). The polynomials are summed down the tree
//     (re-expanded exactly about each child's centre of mass) and evaluated
//     at the bucket particles. Softened cells add w . d b to D_0.
// Softening lengths are constants (a cell's mass-weighted softening is not
// differentiated).
//
// Passes: (1) the cell list is transposed on the GPU (count, scan, scatter:
// for every cell, the (bucket, replica) pairs that use it); (2) a persistent
// cell-major kernel accumulates D for chunks of each cell's users, one wave
// per chunk, lanes over the users' targets; (3) a bucket-major kernel adds
// the target-side terms of cells (lanes as in the forward kernel: slots of
// targets over LDS-staged cells) and both sides of particle pairs (one source
// per lane, targets broadcast); (4) top-down polynomial pass per depth;
// (5) evaluation at the particles.
#include "gravity_body.h"

namespace cgpu {

constexpr int kAdjStride = 80;   // doubles per node: D[36], Vs[3], pad, P[35] at 40
constexpr int kAdjVs = 36;
constexpr int kAdjPoly = 40;
constexpr int kUsersPerItem = 128;
constexpr int kAdjPcTile = 64;

// ---------------------------------------------------------------------------
// Dual numbers: value and derivative along w.
// ---------------------------------------------------------------------------
struct D2 {
  float v, d;
};
__device__ __forceinline__ D2 operator+(D2 a, D2 b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D2 operator-(D2 a, D2 b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D2 operator*(D2 a, D2 b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
__device__ __forceinline__ D2 operator*(float s, D2 a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ D2 operator+(float s, D2 a) { return {s + a.v, a.d}; }
__device__ __forceinline__ D2 drsq(D2 a) {
  const float v = rsq(a.v);
  return {v, -0.5f * v * v * v * a.d};
}

// (da/dr) w for the hexadecapole acceleration of a staged record (the
// formula of pc_interact, evaluated on dual numbers with tangent w).
__device__ __forceinline__ void pc_tidal(const float4* rec, float rx, float ry, float rz, float wx,
                                         float wy, float wz, float& tx, float& ty, float& tz) {
  const D2 X = {rx, wx}, Y = {ry, wy}, Z = {rz, wz};
  const D2 dir = drsq(X * X + Y * Y + Z * Z);
  const float4 c0 = rec[0], c1 = rec[1], c2 = rec[2], c3 = rec[3];
  const float4 c4 = rec[4], c5 = rec[5], c6 = rec[6], c7 = rec[7];
  const float m = c0.w;
  const float qxx = c1.x, qxy = c1.y, qxz = c1.z, qyy = c1.w, qyz = c2.x, qzz = c2.y;
  const float qxxx = c2.z, qxyy = c2.w, qxxy = c3.x, qyyy = c3.y, qxxz = c3.z, qyyz = c3.w,
              qxyz = c4.x, q3a = c4.y, q3b = c4.z;
  const float qxxxx = c4.w, qxyyy = c5.x, qxxxy = c5.y, qyyyy = c5.z, qxxxz = c5.w, qyyyz = c6.x,
              qxxyy = c6.y, qxxyz = c6.z, qxyyz = c6.w, q4a = c7.x, q4b = c7.y, q4c = c7.z;
  const D2 x = X * dir, y = Y * dir, z = Z * dir;
  const D2 x2 = x * x, y2 = y * y, z2 = z * z;
  const D2 fxx = x2 - z2, fyy = y2 - z2;
  const D2 fxy = 2.0f * (x * y), fxz = 2.0f * (x * z), fyz = 2.0f * (y * z);
  const D2 z3 = 3.0f * z2;
  const D2 fxxx = x * (x2 - z3), fyyy = y * (y2 - z3);
  const D2 fxyy = 3.0f * (x * fyy), fxxy = 3.0f * (y * fxx);
  const D2 fxxz = z * (3.0f * x2 - z2), fyyz = z * (3.0f * y2 - z2);
  const D2 fxyz = 3.0f * (z * fxy);
  const D2 v2x = qxx * x + qxy * y + qxz * z;
  const D2 v2y = qxy * x + qyy * y + qyz * z;
  const D2 v2z = qxz * x + qyz * y + qzz * z;
  const D2 v3x = qxxx * fxx + qxyy * fyy + qxxy * fxy + qxxz * fxz + qxyz * fyz;
  const D2 v3y = qxyy * fxy + qxxy * fxx + qyyy * fyy + qyyz * fyz + qxyz * fxz;
  const D2 v3z = qxxz * fxx + qyyz * fyy + qxyz * fxy + q3a * fxz + q3b * fyz;
  const D2 v4x = qxxxx * fxxx + qxxyy * fxyy + qxxxy * fxxy + qxyyy * fyyy + qxxxz * fxxz +
                 qxyyz * fyyz + qxxyz * fxyz;
  const D2 v4y = qyyyy * fyyy + qxxyy * fxxy + qxxxy * fxxx + qxyyy * fxyy + qxyyz * fxyz +
                 qxxyz * fxxz + qyyyz * fyyz;
  const D2 v4z = qxxxz * fxxx + qxyyz * fxyy + qxxyz * fxxy + qyyyz * fyyy + q4a * fxxz +
                 q4b * fyyz + q4c * fxyz;
  const D2 p2 = v2x * x + v2y * y + v2z * z;
  const D2 p3 = v3x * x + v3y * y + v3z * z;
  const D2 p4 = v4x * x + v4y * y + v4z * z;
  const D2 d2 = dir * dir;
  const D2 S = m + d2 * (2.5f * p2 + dir * ((7.0f / 3.0f) * p3 + 2.25f * (dir * p4)));
  const D2 e = d2 * d2, f = S * d2;
  tx = (e * (v2x + dir * (v3x + dir * v4x)) - f * x).d;
  ty = (e * (v2y + dir * (v3y + dir * v4y)) - f * y).d;
  tz = (e * (v2z + dir * (v3z + dir * v4z)) - f * z).d;
}

// ---------------------------------------------------------------------------
// Derivatives of 1/r: T[a][b][c] = d^(a+b+c)(1/|r|)/dx^a dy^b dz^c for
// a+b+c <= 6, c <= 2 (a set closed under the recursion). From
// r^2 d_x f + x f = 0 differentiated by Leibniz' rule:
//   r^2 T(al+e_x) = -(2a+1) x T(al) - 2b y T(al+e_x-e_y) - 2c z T(al+e_x-e_z)
//                   - a^2 T(al-e_x) - b(b-1) T(al+e_x-2e_y) - c(c-1) T(al+e_x-2e_z),
// and the same with the axes permuted.
// ---------------------------------------------------------------------------
struct RinvDer {
  float t[7][7][3];
};

__device__ __forceinline__ void rinv_derivs6(float x, float y, float z, float ir, float ir2,
                                             RinvDer& T) {
  T.t[0][0][0] = ir;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
#pragma unroll
    for (int a = n + 1; a >= 0; --a) {
#pragma unroll
      for (int b = n + 1 - a; b >= 0; --b) {
        const int c = n + 1 - a - b;
        if (c > 2) continue;
        float v;
        if (a > 0) {
          const int A = a - 1;
          v = -(float)(2 * A + 1) * x * T.t[A][b][c];
          if (b >= 1) v -= (float)(2 * b) * y * T.t[a][b - 1][c];
          if (c >= 1) v -= (float)(2 * c) * z * T.t[a][b][c - 1];
          if (A >= 1) v -= (float)(A * A) * T.t[A - 1][b][c];
          if (b >= 2) v -= (float)(b * (b - 1)) * T.t[a][b - 2][c];
          if (c >= 2) v -= (float)(c * (c - 1)) * T.t[a][b][c - 2];
        } else if (b > 0) {
          const int B = b - 1;
          v = -(float)(2 * B + 1) * y * T.t[0][B][c];
          if (c >= 1) v -= (float)(2 * c) * z * T.t[0][b][c - 1];
          if (B >= 1) v -= (float)(B * B) * T.t[0][B - 1][c];
          if (c >= 2) v -= (float)(c * (c - 1)) * T.t[0][b][c - 2];
        } else {
          const int C = c - 1;
          v = -(float)(2 * C + 1) * z * T.t[0][0][C];
          if (C >= 1) v -= (float)(C * C) * T.t[0][0][C - 1];
        }
        T.t[a][b][c] = v * ir2;
      }
    }
  }
}

// Stored multi-indices of D (36): order 0; x y z; xx yy xy xz yz;
// xxx xyy xxy yyy xxz yyz xyz; xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz;
// x5 x4y x3y2 x2y3 xy4 y5 x4z x3yz x2y2z xy3z y4z.
#define CGPU_D36(X)                                                                          \
  X(0, 0, 0, 0) X(1, 1, 0, 0) X(2, 0, 1, 0) X(3, 0, 0, 1) X(4, 2, 0, 0) X(5, 0, 2, 0)       \
  X(6, 1, 1, 0) X(7, 1, 0, 1) X(8, 0, 1, 1) X(9, 3, 0, 0) X(10, 1, 2, 0) X(11, 2, 1, 0)     \
  X(12, 0, 3, 0) X(13, 2, 0, 1) X(14, 0, 2, 1) X(15, 1, 1, 1) X(16, 4, 0, 0)                \
  X(17, 1, 3, 0) X(18, 3, 1, 0) X(19, 0, 4, 0) X(20, 3, 0, 1) X(21, 0, 3, 1)                \
  X(22, 2, 2, 0) X(23, 2, 1, 1) X(24, 1, 2, 1) X(25, 5, 0, 0) X(26, 4, 1, 0)                \
  X(27, 3, 2, 0) X(28, 2, 3, 0) X(29, 1, 4, 0) X(30, 0, 5, 0) X(31, 4, 0, 1)                \
  X(32, 3, 1, 1) X(33, 2, 2, 1) X(34, 1, 3, 1) X(35, 0, 4, 1)

__constant__ int kD36[36][3] = {
#define CGPU_ROW(k, a, b, c) {a, b, c},
    CGPU_D36(CGPU_ROW)
#undef CGPU_ROW
};

// D_alpha += w . grad d^alpha(1/r)
__device__ __forceinline__ void accumulate_d(const RinvDer& T, float wx, float wy, float wz,
                                             float* D) {
#define CGPU_DACC(k, a, b, c) \
  D[k] += wx * T.t[(a) + 1][b][c] + wy * T.t[a][(b) + 1][c] + wz * T.t[a][b][(c) + 1];
  CGPU_D36(CGPU_DACC)
#undef CGPU_DACC
}

// Spline force kernel b and c = -(1/r) db/dr (pairwise Jacobian b I - c d d^T).
__device__ __forceinline__ void spline_bc(float r2, float twoh, float& b, float& c) {
  const float ir = rsq(r2), ir2 = ir * ir;
  const float ir3 = ir * ir2, ir5 = ir3 * ir2;
  if (r2 < twoh * twoh) {
    const float r = r2 * ir, dih = 2.0f / twoh, u = r * dih;
    const float dih3 = dih * dih * dih, dih5 = dih3 * dih * dih;
    if (u < 1.0f) {
      b = dih3 * (4.0f / 3.0f + u * u * (-6.0f / 5.0f + 0.5f * u));
      c = dih5 * (12.0f / 5.0f - 1.5f * u);
    } else {
      b = -(1.0f / 15.0f) * ir3 +
          dih3 * (8.0f / 3.0f + u * (-3.0f + u * (6.0f / 5.0f - u * (1.0f / 6.0f))));
      c = -0.2f * ir5 + dih5 * (3.0f / u - 12.0f / 5.0f + 0.5f * u);
    }
  } else {
    b = ir3;
    c = 3.0f * ir5;
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int y = __shfl_up(v, d, kWave);
    if (lane >= d) v += y;
  }
  return v;
}

// first index in [0, n) whose inclusive end exceeds q
__device__ __forceinline__ int find_run(const int* ends, int n, int q) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ends[mid] > q) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Transposed cell lists (int scratch):
//   count[nNodes], cursor[nNodes], start[nNodes + 1], item[nNodes + 1],
//   nItems[1], users[2 * entries] ({bucket, offsetID} per entry).
struct Transpose {
  int* count;
  int* cursor;
  int* start;
  int* item;
  int* nItems;
  int2* users;
};

struct AdjArgs {
  ChangaGravityLaunch pc, pp;
  const float* w;
  double* grad;
  double* node;  // kAdjStride doubles per node
  Transpose tr;
  int nNodes;
};

// ---------------------------------------------------------------------------
// (1) transpose
// ---------------------------------------------------------------------------
__global__ void adj_zero_int_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(64) void adj_count_kernel(AdjArgs a) {
  const int b = blockIdx.x;
  const ILCell* cells = static_cast<const ILCell*>(a.pc.list);
  for (int e = a.pc.marks[b] + threadIdx.x; e < a.pc.marks[b + 1]; e += kWave)
    atomicAdd(a.tr.count + cells[e].index, 1);
}

// one workgroup: exclusive scans of the user counts and of the work items
// (ceil(count / kUsersPerItem) per cell)
__global__ __launch_bounds__(1024) void adj_scan_kernel(AdjArgs a) {
  __shared__ int s_u[1024], s_i[1024];
  const int n = a.nNodes, t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  int su = 0, si = 0;
  for (int c = lo; c < hi; ++c) {
    su += a.tr.count[c];
    si += (a.tr.count[c] + kUsersPerItem - 1) / kUsersPerItem;
  }
  s_u[t] = su;
  s_i[t] = si;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int yu = t >= d ? s_u[t - d] : 0, yi = t >= d ? s_i[t - d] : 0;
    __syncthreads();
    s_u[t] += yu;
    s_i[t] += yi;
    __syncthreads();
  }
  int ru = s_u[t] - su, ri = s_i[t] - si;
  for (int c = lo; c < hi; ++c) {
    a.tr.start[c] = ru;
    a.tr.item[c] = ri;
    ru += a.tr.count[c];
    ri += (a.tr.count[c] + kUsersPerItem - 1) / kUsersPerItem;
  }
  if (t == 1023) {
    a.tr.start[n] = s_u[1023];
    a.tr.item[n] = s_i[1023];
    a.tr.nItems[0] = s_i[1023];
  }
}

__global__ __launch_bounds__(64) void adj_scatter_kernel(AdjArgs a) {
  const int b = blockIdx.x;
  const ILCell* cells = static_cast<const ILCell*>(a.pc.list);
  for (int e = a.pc.marks[b] + threadIdx.x; e < a.pc.marks[b + 1]; e += kWave) {
    const ILCell c = cells[e];
    const int k = atomicAdd(a.tr.cursor + c.index, 1);
    a.tr.users[a.tr.start[c.index] + k] = make_int2(b, c.offsetID);
  }
}

// ---------------------------------------------------------------------------
// (2) cell sources: one wavefront per work item (up to kUsersPerItem users of
// one cell); lanes walk the users' targets.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void adj_cells_kernel(AdjArgs a) {
  __shared__ int u_end[kUsersPerItem];
  __shared__ int u_start[kUsersPerItem];
  __shared__ int u_oid[kUsersPerItem];
  const int lane = threadIdx.x;
  const int n_items = a.tr.nItems[0];
  const float* wire = reinterpret_cast<const float*>(a.pc.moments);
  const double L = a.pc.fperiod;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    // the cell of this item: last c with item[c] <= it
    int lo = 0, hi = a.nNodes - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.tr.item[mid] <= it) lo = mid;
      else hi = mid - 1;
    }
    const int c = lo;
    const int u0 = a.tr.start[c] + (it - a.tr.item[c]) * kUsersPerItem;
    const int nu = min(kUsersPerItem, a.tr.start[c + 1] - u0);
    const double cx = a.pc.cm64[3 * (size_t)c], cy = a.pc.cm64[3 * (size_t)c + 1],
                 cz = a.pc.cm64[3 * (size_t)c + 2];
    const float cmass = wire[(size_t)c * 27 + 2], csoft = wire[(size_t)c * 27 + 1];
    // users -> target ranges, prefix of their sizes
    int base = 0;
    for (int k0 = 0; k0 < nu; k0 += kWave) {
      const int k = k0 + lane;
      int sz = 0;
      if (k < nu) {
        const int2 u = a.tr.users[u0 + k];
        sz = a.pc.sizes[u.x];
        u_start[k] = a.pc.starts[u.x];
        u_oid[k] = u.y;
      }
      const int incl = wave_incl_scan(sz, lane);
      if (k < nu) u_end[k] = base + incl;
      base += __shfl(incl, kWave - 1, kWave);
    }
    __syncthreads();
    const int total = base;
    float D[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) D[k] = 0.0f;
    float vs0 = 0.f, vs1 = 0.f, vs2 = 0.f;
    for (int q0 = 0; q0 < total; q0 += kWave) {
      const int q = q0 + lane;
      if (q < total) {
        const int k = find_run(u_end, nu, q);
        const int oid = u_oid[k];
        const int i = u_start[k] + q - (k ? u_end[k - 1] : 0);
        const float rx = (float)(a.pc.tpos64[3 * (size_t)i] - (cx + off_x(oid) * L));
        const float ry = (float)(a.pc.tpos64[3 * (size_t)i + 1] - (cy + off_y(oid) * L));
        const float rz = (float)(a.pc.tpos64[3 * (size_t)i + 2] - (cz + off_z(oid) * L));
        const float wx = a.w[3 * (size_t)i], wy = a.w[3 * (size_t)i + 1],
                    wz = a.w[3 * (size_t)i + 2];
        const float r2 = rx * rx + ry * ry + rz * rz;
        if (r2 != 0.0f) {
          if (oid & CHANGA_ILCELL_SOFTENED) {
            // monopole pseudo-particle at the cm: d = -r
            float bb, cc;
            spline_bc(r2, a.pc.targets[i].soft + csoft, bb, cc);
            const float dw = -(rx * wx + ry * wy + rz * wz);
            D[0] += bb * dw;
            vs0 -= cmass * (bb * wx + cc * rx * dw);
            vs1 -= cmass * (bb * wy + cc * ry * dw);
            vs2 -= cmass * (bb * wz + cc * rz * dw);
          } else {
            const float ir = rsq(r2);
            RinvDer T;
            rinv_derivs6(rx, ry, rz, ir, ir * ir, T);
            accumulate_d(T, wx, wy, wz, D);
          }
        }
      }
    }
    // reduce across the wavefront; lane k adds component k
    double* nd = a.node + (size_t)c * kAdjStride;
    float mine = 0.0f;
#pragma unroll
    for (int k = 0; k < 36; ++k) {
      const float s = wave_sum(D[k]);
      if (lane == k) mine = s;
    }
    {
      const float s0 = wave_sum(vs0), s1 = wave_sum(vs1), s2 = wave_sum(vs2);
      if (lane == 36) mine = s0;
      if (lane == 37) mine = s1;
      if (lane == 38) mine = s2;
    }
    if (lane < 39 && mine != 0.0f) atomicAdd(nd + lane, (double)mine);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// (3) target side of cells and both sides of particle pairs, per bucket
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(64) void adj_targets_kernel(AdjArgs a) {
  __shared__ float4 rec4[kAdjPcTile * (kPcRec / 4)];
  __shared__ float4 tp4[kWave];   // target x, y, z (bucket-centred), softening
  __shared__ float4 tw4[kWave];   // w
  __shared__ double tacc[3][kWave];
  __shared__ int r_idx[kWave], r_oid[kWave], r_end[kWave];

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.pc.numBuckets);
  const int pcs = a.pc.marks[b], pce = a.pc.marks[b + 1];
  const int pps = a.pp.marks[b], ppe = a.pp.marks[b + 1];
  const int bstart = a.pc.starts[b], n = a.pc.sizes[b];
  const float L = a.pc.fperiod;
  double cx0, cy0, cz0;
  bucket_centre<true>(a.pc.tpos64, bstart, n, cx0, cy0, cz0);
  const ILCell* cells = static_cast<const ILCell*>(a.pc.list);

  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int nt = min(kWave, n - t0);
    const Lanes l = lanes_for(lane, nt);
    if (lane < nt) {
      const int tp = bstart + t0 + lane;
      tp4[lane] = make_float4((float)(a.pc.tpos64[3 * (size_t)tp] - cx0),
                              (float)(a.pc.tpos64[3 * (size_t)tp + 1] - cy0),
                              (float)(a.pc.tpos64[3 * (size_t)tp + 2] - cz0),
                              a.pc.targets[tp].soft);
      tw4[lane] = make_float4(a.w[3 * (size_t)tp], a.w[3 * (size_t)tp + 1],
                              a.w[3 * (size_t)tp + 2], 0.0f);
    }
    tacc[0][lane] = 0.0;
    tacc[1][lane] = 0.0;
    tacc[2][lane] = 0.0;
    __syncthreads();
    const int j = l.live ? l.lane - l.slot * l.nt : 0;
    const float4 me = tp4[j], mw = tw4[j];

    // ---- cells: slots of targets over LDS-staged tiles
    double ax = 0.0, ay = 0.0, az = 0.0;
    WireCell wc;
    wc.oid = 0;
    load_cell<true>(a.pc, cells, pcs + lane, pce, wc);
    for (int base = pcs; base < pce; base += kAdjPcTile) {
      const int nc = min(kAdjPcTile, pce - base);
      if (lane < nc) stage_cell<true>(a.pc, wc, L, cx0, cy0, cz0, rec4 + lane * (kPcRec / 4));
      __syncthreads();
      load_cell<true>(a.pc, cells, base + kAdjPcTile + lane, pce, wc);
      if (l.live) {
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int c = l.slot; c < nc; c += l.nslot) {
          const float4* r4 = rec4 + c * (kPcRec / 4);
          const float4 h = r4[0];
          const float rx = me.x - h.x, ry = me.y - h.y, rz = me.z - h.z;
          const float r2 = rx * rx + ry * ry + rz * rz;
          if (r2 == 0.0f) continue;
          float tx, ty, tz;
          if (h.w < 0.0f) {  // softened: spline monopole at the cm (d = -r)
            float bb, cc;
            spline_bc(r2, me.w + r4[1].x, bb, cc);
            const float dw = -(rx * mw.x + ry * mw.y + rz * mw.z);
            tx = h.w * (bb * mw.x + cc * rx * dw);
            ty = h.w * (bb * mw.y + cc * ry * dw);
            tz = h.w * (bb * mw.z + cc * rz * dw);
          } else {
            pc_tidal(r4, rx, ry, rz, mw.x, mw.y, mw.z, tx, ty, tz);
          }
          sx += tx;
          sy += ty;
          sz += tz;
        }
        ax += sx;
        ay += sy;
        az += sz;
      }
      __syncthreads();
    }

    // ---- particles: one source per lane, targets broadcast
    auto source_pass = [&](int p, int oid, bool live) {
      float sx = 0.f, sy = 0.f, sz = 0.f, sm = 0.f, ss = 0.f;
      if (live) {
        const CompactPartData q = a.pp.sources[p];
        sx = (float)(a.pp.spos64[3 * (size_t)p] + (double)off_x(oid) * (double)L - cx0);
        sy = (float)(a.pp.spos64[3 * (size_t)p + 1] + (double)off_y(oid) * (double)L - cy0);
        sz = (float)(a.pp.spos64[3 * (size_t)p + 2] + (double)off_z(oid) * (double)L - cz0);
        sm = q.mass;
        ss = q.soft;
      }
      float gx = 0.f, gy = 0.f, gz = 0.f, gm = 0.f;
      for (int jj = 0; jj < nt; ++jj) {
        const float4 t = tp4[jj], wv = tw4[jj];
        float tx = 0.f, ty = 0.f, tz = 0.f;
        if (live) {
          const float dx = sx - t.x, dy = sy - t.y, dz = sz - t.z;
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (r2 != 0.0f) {
            float bb, cc;
            spline_bc(r2, t.w + ss, bb, cc);
            const float dw = dx * wv.x + dy * wv.y + dz * wv.z;
            const float jx = bb * wv.x - cc * dx * dw, jy = bb * wv.y - cc * dy * dw,
                        jz = bb * wv.z - cc * dz * dw;
            tx = -sm * jx;
            ty = -sm * jy;
            tz = -sm * jz;
            gx += sm * jx;
            gy += sm * jy;
            gz += sm * jz;
            gm += bb * dw;
          }
        }
        tx = wave_sum(tx);
        ty = wave_sum(ty);
        tz = wave_sum(tz);
        if (lane == 0) {
          tacc[0][jj] += tx;
          tacc[1][jj] += ty;
          tacc[2][jj] += tz;
        }
      }
      if (live) {
        double* g = a.grad + 4 * (size_t)p;
        atomicAdd(g, (double)gx);
        atomicAdd(g + 1, (double)gy);
        atomicAdd(g + 2, (double)gz);
        atomicAdd(g + 3, (double)gm);
      }
    };

    if (KIND == 0) {
      const ILCell* lst = static_cast<const ILCell*>(a.pp.list);
      for (int base = pps; base < ppe; base += kWave) {
        const bool live = base + lane < ppe;
        ILCell c = {0, 0};
        if (live) c = lst[base + lane];
        source_pass(c.index, c.offsetID, live);
      }
    } else {
      const ILPart* lst = static_cast<const ILPart*>(a.pp.list);
      for (int rb = pps; rb < ppe; rb += kWave) {
        const int nr = min(kWave, ppe - rb);
        int num = 0;
        if (lane < nr) {
          ILPart q = lst[rb + lane];
          r_idx[lane] = q.index;
          r_oid[lane] = q.off;
          num = q.num;
        }
        const int incl = wave_incl_scan(num, lane);
        r_end[lane] = incl;
        const int total = __shfl(incl, kWave - 1, kWave);
        __syncthreads();
        for (int sb = 0; sb < total; sb += kWave) {
          const int q = sb + lane;
          const bool live = q < total;
          int p = 0, oid = 0;
          if (live) {
            const int k = find_run(r_end, nr, q);
            p = r_idx[k] + (q - (k ? r_end[k - 1] : 0));
            oid = r_oid[k];
          }
          source_pass(p, oid, live);
        }
        __syncthreads();
      }
    }

    // ---- per target: slots of the cell part + the particle part
    __syncthreads();
    double(*red)[kWave] = reinterpret_cast<double(*)[kWave]>(rec4);
    red[0][lane] = l.live ? ax : 0.0;
    red[1][lane] = l.live ? ay : 0.0;
    red[2][lane] = l.live ? az : 0.0;
    __syncthreads();
    if (lane < nt) {
      double s0 = tacc[0][lane], s1 = tacc[1][lane], s2 = tacc[2][lane];
      for (int s = 0; s < l.nslot; ++s) {
        s0 += red[0][s * nt + lane];
        s1 += red[1][s * nt + lane];
        s2 += red[2][s * nt + lane];
      }
      double* g = a.grad + 4 * (size_t)(bstart + t0 + lane);
      atomicAdd(g, s0);
      atomicAdd(g + 1, s1);
      atomicAdd(g + 2, s2);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// (4, 5) tree passes (FP64). Monomials of degree <= 5 are indexed degree by
// degree, a descending, then b descending.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mono(int a, int b, int c) {
  const int n = a + b + c;
  return n * (n + 1) * (n + 2) / 6 + (n - a) * (n - a + 1) / 2 + (n - a - b);
}

__constant__ int kRedAbc[22][3] = {
    {0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {3, 0, 0}, {1, 2, 0},
    {2, 1, 0}, {0, 3, 0}, {2, 0, 1}, {0, 2, 1}, {1, 1, 1}, {4, 0, 0}, {1, 3, 0}, {3, 1, 0},
    {0, 4, 0}, {3, 0, 1}, {0, 3, 1}, {2, 2, 0}, {2, 1, 1}, {1, 2, 1}};
__constant__ double kInvFact[6] = {1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0};
__constant__ double kBinom[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0},
                                    {1, 3, 3, 1, 0}, {1, 4, 6, 4, 1}};

// trace-free completion in place over monomial slots of orders 2..nmax:
// X(a,b,c>=2) = -X(a+2,b,c-2) - X(a,b+2,c-2)
__device__ void complete_traceless(double* full, int nmax) {
  for (int n = 2; n <= nmax; ++n)
    for (int c = 2; c <= n; ++c)
      for (int a = n - c; a >= 0; --a) {
        const int b = n - c - a;
        full[mono(a, b, c)] = -full[mono(a + 2, b, c - 2)] - full[mono(a, b + 2, c - 2)];
      }
}

// Hhat of one node into P (35 coefficients)
__device__ void node_hhat(const double* nd, const double* g64, double* P) {
  double Df[56], Q[35];
  for (int q = 0; q < 56; ++q) Df[q] = 0.0;
  for (int k = 0; k < 36; ++k) Df[mono(kD36[k][0], kD36[k][1], kD36[k][2])] = nd[k];
  complete_traceless(Df, 5);
  for (int q = 0; q < 35; ++q) Q[q] = 0.0;
  const double u = g64[3];
  for (int k = 0; k < 22; ++k) {
    const int a = kRedAbc[k][0], b = kRedAbc[k][1], c = kRedAbc[k][2];
    double s = g64[12 + k];
    for (int i = 0; i < a + b + c; ++i) s *= u;
    Q[mono(a, b, c)] = s;
  }
  complete_traceless(Q, 4);
  for (int q = 0; q < 35; ++q) P[q] = 0.0;
  for (int n = 0; n <= 4; ++n) {
    if (n == 1) continue;
    for (int a = n; a >= 0; --a)
      for (int b = n - a; b >= 0; --b) {
        const int c = n - a - b;
        P[mono(a, b, c)] =
            ((n & 1) ? -1.0 : 1.0) * kInvFact[a] * kInvFact[b] * kInvFact[c] * Df[mono(a, b, c)];
      }
  }
  const double M = g64[5];
  for (int j = 0; j < 3; ++j) {
    double S = nd[kAdjVs + j];
    for (int a = 4; a >= 0; --a)
      for (int b = 4 - a; b >= 0; --b) {
        const int c = 4 - a - b;
        S += kInvFact[a] * kInvFact[b] * kInvFact[c] * Q[mono(a, b, c)] *
             Df[mono(a + (j == 0), b + (j == 1), c + (j == 2))];
      }
    const double U = Df[mono(j == 0, j == 1, j == 2)] + (M > 0.0 ? S / M : 0.0);
    P[mono(j == 0, j == 1, j == 2)] = -U;
  }
}

// P += Pp(s + d)
__device__ void add_shifted(const double* Pp, const double d[3], double* P) {
  double pw[3][5];
  for (int k = 0; k < 3; ++k) {
    pw[k][0] = 1.0;
    for (int i = 1; i < 5; ++i) pw[k][i] = pw[k][i - 1] * d[k];
  }
  for (int n = 0; n <= 4; ++n)
    for (int a = n; a >= 0; --a)
      for (int b = n - a; b >= 0; --b) {
        const int c = n - a - b;
        const double p = Pp[mono(a, b, c)];
        if (p == 0.0) continue;
        for (int i = 0; i <= a; ++i)
          for (int j = 0; j <= b; ++j)
            for (int k = 0; k <= c; ++k)
              P[mono(i, j, k)] += p * kBinom[a][i] * kBinom[b][j] * kBinom[c][k] *
                                  pw[0][a - i] * pw[1][b - j] * pw[2][c - k];
      }
}

__global__ void adjoint_root_kernel(double* node, const double* node64) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double P[35];
    node_hhat(node, node64, P);
    for (int q = 0; q < 35; ++q) node[kAdjPoly + q] = P[q];
  }
}

// children of the nodes [lo, hi): P_child = Hhat_child + P_parent(s + cm_child - cm_parent)
__global__ void adjoint_down_kernel(int lo, int hi, const int* child, double* node,
                                    const double* node64) {
  const int n = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= hi) return;
  const double* gp = node64 + (size_t)n * 34;
  const double* Pp = node + (size_t)n * kAdjStride + kAdjPoly;
  for (int k = 0; k < 2; ++k) {
    const int c = child[2 * n + k];
    if (c < 0) continue;
    const double* gc = node64 + (size_t)c * 34;
    double* nc = node + (size_t)c * kAdjStride;
    double P[35];
    node_hhat(nc, gc, P);
    const double d[3] = {gc[0] - gp[0], gc[1] - gp[1], gc[2] - gp[2]};
    add_shifted(Pp, d, P);
    for (int q = 0; q < 35; ++q) nc[kAdjPoly + q] = P[q];
  }
}

// bucket particles: dL/dm += P(s), dL/dx += m grad P(s)
__global__ void adjoint_particles_kernel(int n_nodes, const int* child, const int* first,
                                         const int* count, const double* node,
                                         const double* node64, const double* pos64,
                                         const double* mass64, double* grad) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes || child[2 * n] >= 0) return;
  const double* g = node64 + (size_t)n * 34;
  const double* P = node + (size_t)n * kAdjStride + kAdjPoly;
  for (int k = first[n]; k < first[n] + count[n]; ++k) {
    const double s[3] = {pos64[3 * (size_t)k] - g[0], pos64[3 * (size_t)k + 1] - g[1],
                         pos64[3 * (size_t)k + 2] - g[2]};
    double pw[3][5];
    for (int q = 0; q < 3; ++q) {
      pw[q][0] = 1.0;
      for (int i = 1; i < 5; ++i) pw[q][i] = pw[q][i - 1] * s[q];
    }
    double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int nn = 0; nn <= 4; ++nn)
      for (int a = nn; a >= 0; --a)
        for (int b = nn - a; b >= 0; --b) {
          const int c = nn - a - b;
          const double p = P[mono(a, b, c)];
          val += p * pw[0][a] * pw[1][b] * pw[2][c];
          if (a) gx += p * a * pw[0][a - 1] * pw[1][b] * pw[2][c];
          if (b) gy += p * b * pw[0][a] * pw[1][b - 1] * pw[2][c];
          if (c) gz += p * c * pw[0][a] * pw[1][b] * pw[2][c - 1];
        }
    const double m = mass64[k];
    double* gk = grad + 4 * (size_t)k;
    gk[0] += m * gx;
    gk[1] += m * gy;
    gk[2] += m * gz;
    gk[3] += val;
  }
}

__global__ void fill_zero_kernel(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

static int zero(double* p, size_t n, hipStream_t s) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

}  // namespace cgpu

using namespace cgpu;

extern "C" size_t changaGravityAdjointWorkInts(int nNodes, int numCellEntries) {
  if (nNodes <= 0 || numCellEntries < 0) return 0;
  return 4 * (size_t)nNodes + 2 + 1 + 2 * (size_t)numCellEntries + 8;
}

extern "C" int changaGravityAdjoint(const ChangaAdjointLaunch* a, void* stream) {
  if (!a || !a->w || !a->grad || !a->nodeWork || !a->listWork || !a->levelOffsets || !a->first ||
      !a->count || !a->child || !a->node64 || !a->pos64 || !a->mass64 || a->nNodes <= 0 ||
      a->nDepths <= 0 || a->nParticles <= 0 || a->numCellEntries < 0)
    return fail("changaGravityAdjoint: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  if (a->listWorkInts < changaGravityAdjointWorkInts(a->nNodes, a->numCellEntries))
    return fail("changaGravityAdjoint: listWork too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  const ChangaGravityLaunch& pc = a->cells;
  const ChangaGravityLaunch& pp = a->parts;
  if (!pc.marks || !pc.starts || !pc.sizes || !pc.targets || !pc.moments || !pc.list ||
      !pc.tpos64 || !pc.cm64 || !pp.marks || !pp.list || !pp.sources || !pp.spos64 ||
      pc.numBuckets != pp.numBuckets || pc.starts != pp.starts || pc.sizes != pp.sizes ||
      pc.fperiod != pp.fperiod || pc.listKind != 0 || (pp.listKind != 0 && pp.listKind != 1))
    return fail("changaGravityAdjoint: cell/particle launches (precision form, same buckets)",
                hipErrorInvalidValue, __FILE__, __LINE__);
  hipStream_t s = (hipStream_t)stream;
  const int nn = a->nNodes;
  if (int e = zero(a->grad, 4 * (size_t)a->nParticles, s)) return e;
  if (int e = zero(a->nodeWork, (size_t)kAdjStride * nn, s)) return e;
  AdjArgs args;
  args.pc = pc;
  args.pp = pp;
  args.w = a->w;
  args.grad = a->grad;
  args.node = a->nodeWork;
  args.nNodes = nn;
  int* iw = a->listWork;
  args.tr.count = iw;
  args.tr.cursor = iw + nn;
  args.tr.start = iw + 2 * (size_t)nn;
  args.tr.item = iw + 3 * (size_t)nn + 1;
  args.tr.nItems = iw + 4 * (size_t)nn + 2;
  args.tr.users = reinterpret_cast<int2*>(iw + 4 * (size_t)nn + 4);  // 8-byte aligned
  const int nb = pc.numBuckets;
  hipLaunchKernelGGL(adj_zero_int_kernel, dim3((2 * nn + 255) / 256), dim3(256), 0, s, iw,
                     2 * nn);
  CGPU_CHECK(hipGetLastError());
  if (nb > 0) {
    hipLaunchKernelGGL(adj_count_kernel, dim3(nb), dim3(kWave), 0, s, args);
    CGPU_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(adj_scan_kernel, dim3(1), dim3(1024), 0, s, args);
  CGPU_CHECK(hipGetLastError());
  if (nb > 0) {
    hipLaunchKernelGGL(adj_scatter_kernel, dim3(nb), dim3(kWave), 0, s, args);
    CGPU_CHECK(hipGetLastError());
    int dev = 0, ncu = 0;
    CGPU_CHECK(hipGetDevice(&dev));
    CGPU_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(adj_cells_kernel, dim3(ncu * 16), dim3(kWave), 0, s, args);
    CGPU_CHECK(hipGetLastError());
    if (pp.listKind == 1)
      hipLaunchKernelGGL(adj_targets_kernel<1>, dim3(nb), dim3(kWave), 0, s, args);
    else
      hipLaunchKernelGGL(adj_targets_kernel<0>, dim3(nb), dim3(kWave), 0, s, args);
    CGPU_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(adjoint_root_kernel, dim3(1), dim3(1), 0, s, a->nodeWork, a->node64);
  CGPU_CHECK(hipGetLastError());
  for (int d = 0; d < a->nDepths; ++d) {
    const int lo = a->levelOffsets[d], hi = a->levelOffsets[d + 1];
    if (hi <= lo) continue;
    hipLaunchKernelGGL(adjoint_down_kernel, dim3((hi - lo + 127) / 128), dim3(128), 0, s, lo, hi,
                       a->child, a->nodeWork, a->node64);
    CGPU_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(adjoint_particles_kernel, dim3((nn + 127) / 128), dim3(128), 0, s, nn,
                     a->child, a->first, a->count, a->nodeWork, a->node64, a->pos64, a->mass64,
                     a->grad);
  CGPU_CHECK(hipGetLastError());
  return 0;
}
