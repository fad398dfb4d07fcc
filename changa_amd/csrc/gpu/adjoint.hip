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
// Passes: (1) the cell list is transposed on the GPU: a stable radix sort of
// the entries by cell (rocPRIM) gives, for every cell, the (bucket, replica)
// pairs that use it in bucket order; (2) a persistent cell-major kernel
// accumulates D for chunks of each cell's users, one wave per chunk, lanes
// over the users' targets; (3) a bucket-major kernel adds the target-side
// terms of cells (lanes as in the forward kernel: slots of targets over
// LDS-staged cells) and both sides of particle pairs (one source per lane,
// targets broadcast); (4) Hhat of every node, then a top-down pass per depth
// (one wave per child, one lane per polynomial coefficient); (5) evaluation
// at the particles.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ewald_body.h"
#include "gravity_body.h"
#include "stf.h"

namespace cgpu {

constexpr int kAdjStride = 112;  // doubles per node: D[36], Vs[3], pad, Hhat[35] at 40,
                                 // P[35] at 75
constexpr int kAdjVs = 36;
constexpr int kAdjHhat = 40;
constexpr int kAdjPoly = 75;
constexpr int kUsersPerItem = 512;
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
// D_alpha = d^alpha f for the dipole potential f = w . grad(1/r) = -(w.r) g,
// g = 1/r^3. By Leibniz' rule
//   D_alpha = -[(w.r) G_alpha + sum_k alpha_k w_k G_(alpha-e_k)],
// with G_alpha = d^alpha g for |alpha| <= 5, c <= 1 (a set closed under the
// recursion). Differentiating r^2 d_x g + 3 x g = 0:
//   r^2 G(al+e_x) = -(2a+3) x G(al) - 2b y G(al+e_x-e_y) - 2c z G(al+e_x-e_z)
//                   - a(a+2) G(al-e_x) - b(b-1) G(al+e_x-2e_y) - c(c-1) G(al+e_x-2e_z),
// and the same with the axes permuted.
// ---------------------------------------------------------------------------
struct R3Der {
  float t[6][6][2];
};

__device__ __forceinline__ void r3_derivs5(float x, float y, float z, float ir, float ir2,
                                           R3Der& G) {
  G.t[0][0][0] = ir * ir2;
#pragma unroll
  for (int n = 0; n < 5; ++n) {
#pragma unroll
    for (int a = n + 1; a >= 0; --a) {
#pragma unroll
      for (int b = n + 1 - a; b >= 0; --b) {
        const int c = n + 1 - a - b;
        if (c > 1) continue;
        float v;
        if (a > 0) {
          const int A = a - 1;
          v = -(float)(2 * A + 3) * x * G.t[A][b][c];
          if (b >= 1) v -= (float)(2 * b) * y * G.t[a][b - 1][c];
          if (c >= 1) v -= (float)(2 * c) * z * G.t[a][b][c - 1];
          if (A >= 1) v -= (float)(A * (A + 2)) * G.t[A - 1][b][c];
          if (b >= 2) v -= (float)(b * (b - 1)) * G.t[a][b - 2][c];
        } else if (b > 0) {
          const int B = b - 1;
          v = -(float)(2 * B + 3) * y * G.t[0][B][c];
          if (c >= 1) v -= (float)(2 * c) * z * G.t[0][b][c - 1];
          if (B >= 1) v -= (float)(B * (B + 2)) * G.t[0][B - 1][c];
        } else {
          const int C = c - 1;
          v = -(float)(2 * C + 3) * z * G.t[0][0][C];
          if (C >= 1) v -= (float)(C * (C + 2)) * G.t[0][0][C - 1];
        }
        G.t[a][b][c] = v * ir2;
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

template <int A, int B, int C>
__device__ __forceinline__ float dipole_term(const R3Der& G, float wr, float wx, float wy,
                                             float wz) {
  float v = wr * G.t[A][B][C];
  if constexpr (A > 0) v += (float)A * wx * G.t[A - 1][B][C];
  if constexpr (B > 0) v += (float)B * wy * G.t[A][B - 1][C];
  if constexpr (C > 0) v += (float)C * wz * G.t[A][B][C - 1];
  return v;
}

// D_alpha += d^alpha (w . grad(1/r)) at r
__device__ __forceinline__ void accumulate_d(float rx, float ry, float rz, float ir, float ir2,
                                             float wx, float wy, float wz, float* D) {
  R3Der G;
  r3_derivs5(rx, ry, rz, ir, ir2, G);
  const float wr = wx * rx + wy * ry + wz * rz;
#define CGPU_DACC(k, a, b, c) D[k] -= dipole_term<a, b, c>(G, wr, wx, wy, wz);
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

// Transposed cell lists (scratch):
//   key[E], value[E] (the entries as sorted: cell, then {bucket, replica and
//   softened bits of offsetID} packed in 32 bits), start[nNodes + 1] (users
//   of cell c at [start[c], start[c+1])), chunks[nNodes + 1], item[nNodes + 1]
//   (work items per cell, exclusive prefix; item[nNodes] is their number),
//   item_cell[items] (the cell of each item).
struct Transpose {
  const unsigned* key;
  const unsigned* value;  // bucket << kBucketShift | offsetID >> 22
  int* start;
  int* chunks;
  int* item;
  int* item_cell;  // cell of each work item
};

// Transposed particle lists (ILPart runs), scratch: value[R] (the runs as
// sorted by their first source particle: bucket << 32 | replica bits (31..22)
// | run length), key[nruns] (distinct first source particles), start[nruns +
// 1] (users of key k at value[start[k], start[k + 1])), nruns (device),
// item[nruns + 1] (work items of up to kRunUsersPerItem users).
struct PartTranspose {
  const unsigned long long* value;
  const int* key;
  const int* start;
  const int* nruns;
  const int* item;      // first work item of key k (exclusive scan of its chunks)
  const int* item_run;  // key index of each work item
};

struct AdjArgs {
  ChangaGravityLaunch pc, pp;
  const float* w;
  const int* tindex;  // particle of each target, or NULL (identity)
  double* grad;
  double* node;  // kAdjStride doubles per node
  Transpose tr;
  PartTranspose ptr;
  int nNodes;
  int entries;
};

// ---------------------------------------------------------------------------
// (1) transpose
// ---------------------------------------------------------------------------
// value: bucket << 10 | offsetID bits 31..22 (replica and softened flag;
// bits 21..0 carry no information for a cell entry)
constexpr int kBucketShift = 10;
constexpr int kMaxAdjBuckets = 1 << 22;

__global__ __launch_bounds__(64) void adj_keys_kernel(ChangaGravityLaunch pc, unsigned* key,
                                                      unsigned* value) {
  const int b = blockIdx.x;
  const ILCell* cells = static_cast<const ILCell*>(pc.list);
  for (int e = pc.marks[b] + threadIdx.x; e < pc.marks[b + 1]; e += kWave) {
    const ILCell c = cells[e];
    key[e] = (unsigned)c.index;
    value[e] = ((unsigned)b << kBucketShift) | ((unsigned)c.offsetID >> 22);
  }
}

// start[c] = first sorted position whose cell is >= c
__global__ void adj_bounds_kernel(const unsigned* key, int entries, int n_nodes, int* start) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (entries == 0) {
    if (k <= n_nodes) start[k] = 0;
    return;
  }
  if (k >= entries) return;
  const int hi = (int)key[k];
  const int lo = k ? (int)key[k - 1] + 1 : 0;
  for (int c = lo; c <= hi; ++c) start[c] = k;
  if (k == entries - 1)
    for (int c = hi + 1; c <= n_nodes; ++c) start[c] = entries;
}

__global__ void adj_chunks_kernel(const int* start, int n_nodes, int* chunks) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n_nodes) chunks[c] = (start[c + 1] - start[c] + kUsersPerItem - 1) / kUsersPerItem;
  if (c == n_nodes) chunks[c] = 0;
}

// owner of each work item: the owner's index at its first item (items of
// owner c: item[c] .. item[c + 1] - 1), then a running maximum fills the rest
__global__ void adj_item_heads_kernel(const int* item, const int* n_owners_dev, int n_owners,
                                      int* owner) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = n_owners_dev ? *n_owners_dev : n_owners;
  if (c < n && item[c + 1] > item[c]) owner[item[c]] = c;
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
  const int n_items = a.tr.item[a.nNodes];
  const float* wire = reinterpret_cast<const float*>(a.pc.moments);
  const double L = a.pc.fperiod;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int c = a.tr.item_cell[it];
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
        const unsigned u = a.tr.value[u0 + k];
        const int bk = (int)(u >> kBucketShift);
        sz = a.pc.sizes[bk];
        u_start[k] = a.pc.starts[bk];
        u_oid[k] = (int)((u & ((1u << kBucketShift) - 1u)) << 22);
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
            accumulate_d(rx, ry, rz, ir, ir * ir, wx, wy, wz, D);
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
// (2b) particle sources, from the particle lists transposed by source run:
// one wavefront per work item (up to kRunUsersPerItem users of one run
// start, a source bucket); the users' targets are staged in LDS 64 at a
// time; lanes form groups of P >= run length lanes, one source particle per
// lane, each group taking every G-th staged target. Per source particle and
// item: the sum over targets of m_s (b w - c d (d.w)) and b (d.w),
// d = x_s - x_t, added into grad (FP64 atomics; one per particle and item).
// ---------------------------------------------------------------------------
constexpr unsigned kRunLenMask = 0x3fffffu;
// A source bucket with at most nMinParticleNode particles is opened by every
// walk, so its run has every bucket as a user: users are split into items.
constexpr int kRunUsersPerItem = 256;

__global__ void adj_pchunks_kernel(const int* count, const int* nruns, int n, int* chunks) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) chunks[i] = i < *nruns ? (count[i] + kRunUsersPerItem - 1) / kRunUsersPerItem : 0;
}

__global__ __launch_bounds__(64) void adj_pkeys_kernel(ChangaGravityLaunch pp, unsigned* key,
                                                       unsigned long long* value) {
  const int b = blockIdx.x;
  const ILPart* lst = static_cast<const ILPart*>(pp.list);
  for (int r = pp.marks[b] + (int)threadIdx.x; r < pp.marks[b + 1]; r += kWave) {
    const ILPart q = lst[r];
    key[r] = (unsigned)q.index;
    value[r] = ((unsigned long long)(unsigned)b << 32) | ((unsigned)q.off & ~kRunLenMask) |
               ((unsigned)q.num & kRunLenMask);
  }
}

__global__ __launch_bounds__(64) void adj_pp_sources_kernel(AdjArgs a) {
  __shared__ float4 st_pos[kWave];  // staged target: x, y, z relative to the run's origin, soft
  __shared__ float4 st_w[kWave];    // w, run length of its user (as int bits)
  __shared__ int u_end[kWave];
  __shared__ int u_first[kWave];
  __shared__ unsigned long long u_val[kWave];
  const int lane = threadIdx.x;
  const int nruns = *a.ptr.nruns;
  const int n_items = a.ptr.item[nruns];
  const double L = a.pp.fperiod;
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int lo = a.ptr.item_run[it];
    const int k = a.ptr.key[lo];
    const int v0 = a.ptr.start[lo] + (it - a.ptr.item[lo]) * kRunUsersPerItem;
    const int v1 = min(v0 + kRunUsersPerItem, a.ptr.start[lo + 1]);
    // longest run among the users (runs sharing a start normally have one length)
    int maxlen = 0;
    for (int u = v0 + lane; u < v1; u += kWave)
      maxlen = max(maxlen, (int)(a.ptr.value[u] & kRunLenMask));
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, m, kWave));
    int P = 1;
    while (P < maxlen && P < kWave) P <<= 1;
    const int G = kWave / P, grp = lane / P, sl = lane % P;
    const double ox = a.pp.spos64[3 * (size_t)k], oy = a.pp.spos64[3 * (size_t)k + 1],
                 oz = a.pp.spos64[3 * (size_t)k + 2];
    for (int s0 = 0; s0 < maxlen; s0 += P) {
      const int si = s0 + sl;  // source index within the run
      const bool src_live = si < maxlen;
      const int p = k + (src_live ? si : 0);
      const CompactPartData q = a.pp.sources[p];
      const float sx = (float)(a.pp.spos64[3 * (size_t)p] - ox);
      const float sy = (float)(a.pp.spos64[3 * (size_t)p + 1] - oy);
      const float sz = (float)(a.pp.spos64[3 * (size_t)p + 2] - oz);
      const float sm = q.mass, ss = q.soft;
      float gx = 0.f, gy = 0.f, gz = 0.f, gm = 0.f;
      for (int ub = v0; ub < v1; ub += kWave) {
        const int nu = min(kWave, v1 - ub);
        int sz_u = 0;
        if (lane < nu) {
          const unsigned long long val = a.ptr.value[ub + lane];
          const int bk = (int)(val >> 32);
          sz_u = a.pc.sizes[bk];
          u_first[lane] = a.pc.starts[bk];
          u_val[lane] = val;
        }
        const int incl = wave_incl_scan(sz_u, lane);
        u_end[lane] = incl;
        const int total = __shfl(incl, kWave - 1, kWave);
        __syncthreads();
        for (int q0 = 0; q0 < total; q0 += kWave) {
          const int qq = q0 + lane;
          if (qq < total) {
            const int uk = find_run(u_end, nu, qq);
            const int t = u_first[uk] + qq - (uk ? u_end[uk - 1] : 0);
            const unsigned long long val = u_val[uk];
            const int oid = (int)(unsigned)(val & 0xffffffffull);
            // d = x_s + off L - x_t  =>  target relative to the run's origin: x_t - off L - o
            st_pos[lane] = make_float4(
                (float)(a.pc.tpos64[3 * (size_t)t] - off_x(oid) * L - ox),
                (float)(a.pc.tpos64[3 * (size_t)t + 1] - off_y(oid) * L - oy),
                (float)(a.pc.tpos64[3 * (size_t)t + 2] - off_z(oid) * L - oz),
                a.pc.targets[t].soft);
            st_w[lane] = make_float4(a.w[3 * (size_t)t], a.w[3 * (size_t)t + 1],
                                     a.w[3 * (size_t)t + 2],
                                     __int_as_float((int)(val & kRunLenMask)));
          }
          __syncthreads();
          const int ns = min(kWave, total - q0);
          if (src_live) {
            for (int j = grp; j < ns; j += G) {
              const float4 tq = st_pos[j], wv = st_w[j];
              if (si >= __float_as_int(wv.w)) continue;
              const float dx = sx - tq.x, dy = sy - tq.y, dz = sz - tq.z;
              const float r2 = dx * dx + dy * dy + dz * dz;
              if (r2 == 0.0f) continue;
              float bb, cc;
              spline_bc(r2, tq.w + ss, bb, cc);
              const float dw = dx * wv.x + dy * wv.y + dz * wv.z;
              gx += sm * (bb * wv.x - cc * dx * dw);
              gy += sm * (bb * wv.y - cc * dy * dw);
              gz += sm * (bb * wv.z - cc * dz * dw);
              gm += bb * dw;
            }
          }
          __syncthreads();
        }
      }
      // sum the G groups (lanes sl, sl + P, ...), then one update per source
#pragma unroll
      for (int m = kWave / 2; m >= 1; m >>= 1) {
        if (m < P) break;
        gx += __shfl_xor(gx, m, kWave);
        gy += __shfl_xor(gy, m, kWave);
        gz += __shfl_xor(gz, m, kWave);
        gm += __shfl_xor(gm, m, kWave);
      }
      if (grp == 0 && src_live) {
        double* g = a.grad + 4 * (size_t)p;
        atomicAdd(g, (double)gx);
        atomicAdd(g + 1, (double)gy);
        atomicAdd(g + 2, (double)gz);
        atomicAdd(g + 3, (double)gm);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// (3) target side of cells and particles, and source side of particles, per
// bucket. Target sides run in the forward kernels' layout (slots of targets
// over LDS-staged sources, sums in registers); the particle source side puts
// one source on each lane with the targets broadcast.
// ---------------------------------------------------------------------------
template <int KIND>
struct TargetsLds {
  static constexpr size_t pc = (size_t)kAdjPcTile * kPcRec * sizeof(float);
  static constexpr size_t pp = sizeof(PpLds<kPpTile>);
  static constexpr size_t bytes = pc > pp ? pc : pp;
  static constexpr size_t f4 = (bytes + 15) / 16;
};

template <int KIND>
__global__ __launch_bounds__(64) void adj_targets_kernel(AdjArgs a) {
  __shared__ float4 lds[TargetsLds<KIND>::f4];
  __shared__ float4 tp4[kWave];  // target x, y, z (bucket-centred), softening
  __shared__ float4 tw4[kWave];  // w
  float4* rec4 = lds;
  PpLds<kPpTile>& ps = *reinterpret_cast<PpLds<kPpTile>*>(lds);

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
    __syncthreads();
    const int j = l.live ? l.lane - l.slot * l.nt : 0;
    const float4 me = tp4[j], mw = tw4[j];
    double ax = 0.0, ay = 0.0, az = 0.0;

    // ---- cells, target side
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

    // ---- particles, target side: -m_s (b w - c d (d.w)), d = x_s - x_i
    pp_tiles<true, KIND, kPpTile>(a.pp, pps, ppe, cx0, cy0, cz0, lane, ps, [&](int ns) {
      if (l.live) {
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int k = l.slot; k < ns; k += l.nslot) {
          const float4 v = ps.src4[k];
          const float dx = v.x - me.x, dy = v.y - me.y, dz = v.z - me.z;
          const float r2 = dx * dx + dy * dy + dz * dz;
          if (r2 == 0.0f) continue;
          float bb, cc;
          spline_bc(r2, me.w + ps.srcs[k], bb, cc);
          const float dw = dx * mw.x + dy * mw.y + dz * mw.z;
          sx -= v.w * (bb * mw.x - cc * dx * dw);
          sy -= v.w * (bb * mw.y - cc * dy * dw);
          sz -= v.w * (bb * mw.z - cc * dz * dw);
        }
        ax += sx;
        ay += sy;
        az += sz;
      }
    });

    // ---- particles, source side: one source per lane
    auto source_pass = [&](int p, int oid, bool live) {
      if (!live) return;
      const CompactPartData q = a.pp.sources[p];
      const float sx = (float)(a.pp.spos64[3 * (size_t)p] + (double)off_x(oid) * (double)L - cx0);
      const float sy =
          (float)(a.pp.spos64[3 * (size_t)p + 1] + (double)off_y(oid) * (double)L - cy0);
      const float sz =
          (float)(a.pp.spos64[3 * (size_t)p + 2] + (double)off_z(oid) * (double)L - cz0);
      const float sm = q.mass, ss = q.soft;
      float gx = 0.f, gy = 0.f, gz = 0.f, gm = 0.f;
      for (int jj = 0; jj < nt; ++jj) {
        const float4 t = tp4[jj], wv = tw4[jj];
        const float dx = sx - t.x, dy = sy - t.y, dz = sz - t.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0.0f) continue;
        float bb, cc;
        spline_bc(r2, t.w + ss, bb, cc);
        const float dw = dx * wv.x + dy * wv.y + dz * wv.z;
        gx += sm * (bb * wv.x - cc * dx * dw);
        gy += sm * (bb * wv.y - cc * dy * dw);
        gz += sm * (bb * wv.z - cc * dz * dw);
        gm += bb * dw;
      }
      double* g = a.grad + 4 * (size_t)p;
      atomicAdd(g, (double)gx);
      atomicAdd(g + 1, (double)gy);
      atomicAdd(g + 2, (double)gz);
      atomicAdd(g + 3, (double)gm);
    };
    // (run lists: the source side is adj_pp_sources_kernel's, over the
    // transposed lists)
    if (KIND == 0) {
      const ILCell* lst = static_cast<const ILCell*>(a.pp.list);
      for (int base = pps; base < ppe; base += kWave) {
        const bool live = base + lane < ppe;
        ILCell c = {0, 0};
        if (live) c = lst[base + lane];
        source_pass(c.index, c.offsetID, live);
      }
    }

    // ---- per target: sum of the slots
    __syncthreads();
    double(*red)[kWave] = reinterpret_cast<double(*)[kWave]>(lds);
    red[0][lane] = l.live ? ax : 0.0;
    red[1][lane] = l.live ? ay : 0.0;
    red[2][lane] = l.live ? az : 0.0;
    __syncthreads();
    if (lane < nt) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (int sl = 0; sl < l.nslot; ++sl) {
        s0 += red[0][sl * nt + lane];
        s1 += red[1][sl * nt + lane];
        s2 += red[2][sl * nt + lane];
      }
      const int tp = bstart + t0 + lane;
      double* g = a.grad + 4 * (size_t)(a.tindex ? a.tindex[tp] : tp);
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
__host__ __device__ constexpr int mono(int a, int b, int c) {
  return (a + b + c) * (a + b + c + 1) * (a + b + c + 2) / 6 +
         (b + c) * (b + c + 1) / 2 + c;
}
__host__ __device__ constexpr double inv_fact(int n) {
  return n <= 1 ? 1.0 : n == 2 ? 0.5 : n == 3 ? 1.0 / 6.0 : n == 4 ? 1.0 / 24.0 : 1.0 / 120.0;
}

// trace-free completion over monomial slots of orders 2..NMAX:
// X(a,b,c>=2) = -X(a+2,b,c-2) - X(a,b+2,c-2)
template <int NMIN, int NMAX>
__device__ __forceinline__ void complete_traceless(double* f) {
#pragma unroll
  for (int n = NMIN; n <= NMAX; ++n)
#pragma unroll
    for (int c = 2; c <= n; ++c)
#pragma unroll
      for (int a = n - c; a >= 0; --a) {
        const int b = n - c - a;
        f[mono(a, b, c)] = -f[mono(a + 2, b, c - 2)] - f[mono(a, b + 2, c - 2)];
      }
}

// Hhat of every node (one thread per node); the root's polynomial is its Hhat
__global__ void adj_hhat_kernel(int n_nodes, double* node, const double* node64) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double* nd = node + (size_t)n * kAdjStride;
  const double* g = node64 + (size_t)n * 34;
  double Df[56];
#pragma unroll
  for (int q = 0; q < 56; ++q) Df[q] = 0.0;
#define CGPU_LOAD(k, a, b, c) Df[mono(a, b, c)] = nd[k];
  CGPU_D36(CGPU_LOAD)
#undef CGPU_LOAD
  complete_traceless<2, 5>(Df);
  // order-4 moments about cm (unscaled), completed
  double Q[35];
  const double u = g[3], u4 = (u * u) * (u * u);
#pragma unroll
  for (int q = 0; q < 35; ++q) Q[q] = 0.0;
  Q[mono(4, 0, 0)] = u4 * g[12 + 13];
  Q[mono(1, 3, 0)] = u4 * g[12 + 14];
  Q[mono(3, 1, 0)] = u4 * g[12 + 15];
  Q[mono(0, 4, 0)] = u4 * g[12 + 16];
  Q[mono(3, 0, 1)] = u4 * g[12 + 17];
  Q[mono(0, 3, 1)] = u4 * g[12 + 18];
  Q[mono(2, 2, 0)] = u4 * g[12 + 19];
  Q[mono(2, 1, 1)] = u4 * g[12 + 20];
  Q[mono(1, 2, 1)] = u4 * g[12 + 21];
  complete_traceless<4, 4>(Q);
  double H[35];
#pragma unroll
  for (int m = 0; m <= 4; ++m)
#pragma unroll
    for (int a = m; a >= 0; --a)
#pragma unroll
      for (int b = m - a; b >= 0; --b) {
        const int c = m - a - b;
        H[mono(a, b, c)] = (m == 1) ? 0.0
                                    : ((m & 1) ? -1.0 : 1.0) * inv_fact(a) * inv_fact(b) *
                                          inv_fact(c) * Df[mono(a, b, c)];
      }
  const double M = g[5];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double S = nd[kAdjVs + j];
#pragma unroll
    for (int a = 4; a >= 0; --a)
#pragma unroll
      for (int b = 4 - a; b >= 0; --b) {
        const int c = 4 - a - b;
        S += inv_fact(a) * inv_fact(b) * inv_fact(c) * Q[mono(a, b, c)] *
             Df[mono(a + (j == 0), b + (j == 1), c + (j == 2))];
      }
    H[mono(j == 0, j == 1, j == 2)] = -(Df[mono(j == 0, j == 1, j == 2)] + (M > 0.0 ? S / M : 0.0));
  }
#pragma unroll
  for (int q = 0; q < 35; ++q) nd[kAdjHhat + q] = H[q];
  if (n == 0)
#pragma unroll
    for (int q = 0; q < 35; ++q) nd[kAdjPoly + q] = H[q];
}

// P(s) -> P(s + d) in place, for a polynomial of degree <= 4 in monomial
// slots: one Taylor shift per axis (repeated synthetic division along each
// line of the other two exponents), 35 FMAs per axis.
template <int AXIS>
__device__ __forceinline__ void taylor_shift_axis(double* P, double d) {
#pragma unroll
  for (int u = 0; u <= 4; ++u)
#pragma unroll
    for (int v = 0; v <= 4 - u; ++v) {
      const int n = 4 - u - v;  // degree along AXIS on this line
#pragma unroll
      for (int k = 0; k < n; ++k)
#pragma unroll
        for (int e = n - 1; e >= k; --e) {
          const int lo = AXIS == 0 ? mono(e, u, v) : AXIS == 1 ? mono(u, e, v) : mono(u, v, e);
          const int hi = AXIS == 0 ? mono(e + 1, u, v)
                                   : AXIS == 1 ? mono(u, e + 1, v) : mono(u, v, e + 1);
          P[lo] = fma(d, P[hi], P[lo]);
        }
    }
}

// children of the nodes [lo, hi): one thread per (parent, child slot),
// P_child(s) = Hhat_child(s) + P_parent(s + cm_child - cm_parent)
__global__ __launch_bounds__(128) void adj_shift_kernel(int lo, int hi, const int* child,
                                                        double* node, const double* node64) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int parent = lo + (i >> 1);
  if (parent >= hi) return;
  const int c = child[2 * parent + (i & 1)];
  if (c < 0) return;
  const double* gp = node64 + (size_t)parent * 34;
  const double* gc = node64 + (size_t)c * 34;
  const double* Pp = node + (size_t)parent * kAdjStride + kAdjPoly;
  double P[35];
#pragma unroll
  for (int q = 0; q < 35; ++q) P[q] = Pp[q];
  taylor_shift_axis<0>(P, gc[0] - gp[0]);
  taylor_shift_axis<1>(P, gc[1] - gp[1]);
  taylor_shift_axis<2>(P, gc[2] - gp[2]);
  const double* Hc = node + (size_t)c * kAdjStride + kAdjHhat;
  double* Pc = node + (size_t)c * kAdjStride + kAdjPoly;
#pragma unroll
  for (int q = 0; q < 35; ++q) Pc[q] = P[q] + Hc[q];
}

// bucket particles: dL/dm += P(s), dL/dx += m grad P(s)
__global__ void adjoint_particles_kernel(int n_nodes, const int* child, const int* first,
                                         const int* count, const double* node,
                                         const double* node64, const double* pos64,
                                         const double* mass64, double* grad) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes || child[2 * n] >= 0) return;
  const double* g = node64 + (size_t)n * 34;
  double P[35];
#pragma unroll
  for (int q = 0; q < 35; ++q) P[q] = node[(size_t)n * kAdjStride + kAdjPoly + q];
  for (int k = first[n]; k < first[n] + count[n]; ++k) {
    const double s[3] = {pos64[3 * (size_t)k] - g[0], pos64[3 * (size_t)k + 1] - g[1],
                         pos64[3 * (size_t)k + 2] - g[2]};
    double pw[3][5];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      pw[t][0] = 1.0;
#pragma unroll
      for (int e = 1; e < 5; ++e) pw[t][e] = pw[t][e - 1] * s[t];
    }
    double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int m = 0; m <= 4; ++m)
#pragma unroll
      for (int a = m; a >= 0; --a)
#pragma unroll
        for (int b = m - a; b >= 0; --b) {
          const int c = m - a - b;
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

// ---------------------------------------------------------------------------
// (6) Ewald term of L (periodic boxes). L_E = sum_t w_t . a_E(x_t - cm_root),
// a_E the forward kernel's acceleration (ewald.hip: replicas with g_0..g_5,
// then the h-table). Per target:
//   - target side: the gradient of w . a_E, by dual numbers through
//     ewald_term with dual radial kernels (D_w g_m = -g_(m+1) (r.w) for the
//     closed form, 2 B_m (r.w) for the two-term series g_m = A_m + B_m r^2),
//     and sum_h k (k.w)(hC cos + hS sin) for the k-terms;
//   - source side, summed over targets: G_n = dL/dQ_n of the root moments,
//     (1/n!) [g_n sum_j alpha_j w_j r^(alpha-e_j) - (w.r) g_(n+1) r^alpha] per
//     replica; A_h = sum (k.w) sin, B_h = -sum (k.w) cos per h-entry (the
//     h-table is linear in M, Q_2, Q_3, Q_4); V = sum of the target-side terms.
// The root kernel turns these into the root's polynomial as for a cell, with
// H(s) = G_0 + sum_n STF(G_n) . s^n (the Ewald kernels are not harmonic, so G
// is projected) and U = (sum_k m_k grad H(s_k) + V) / M.
// ---------------------------------------------------------------------------
template <typename S>
struct Dual {
  S v, d;
};
template <typename S>
__device__ __forceinline__ Dual<S> operator+(Dual<S> a, Dual<S> b) { return {a.v + b.v, a.d + b.d}; }
template <typename S>
__device__ __forceinline__ Dual<S> operator-(Dual<S> a, Dual<S> b) { return {a.v - b.v, a.d - b.d}; }
template <typename S>
__device__ __forceinline__ Dual<S> operator*(Dual<S> a, Dual<S> b) {
  return {a.v * b.v, a.v * b.d + a.d * b.v};
}
template <typename S>
__device__ __forceinline__ Dual<S> operator*(S s, Dual<S> a) { return {s * a.v, s * a.d}; }
template <typename S>
__device__ __forceinline__ Dual<S> operator*(Dual<S> a, S s) { return {s * a.v, s * a.d}; }

__device__ __forceinline__ float erf_s(float x) { return erff(x); }
__device__ __forceinline__ double erf_s(double x) { return erf(x); }
__device__ __forceinline__ float erfc_s(float x) { return erfcf(x); }
__device__ __forceinline__ double erfc_s(double x) { return erfc(x); }
__device__ __forceinline__ float exp_s(float x) { return expf(x); }
__device__ __forceinline__ double exp_s(double x) { return exp(x); }
__device__ __forceinline__ float sqrt_s(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_s(double x) { return sqrt(x); }

// g_0..g_6 as in the forward kernel's radial32 / radial64, and gam_0..gam_5
// with grad g_m = gam_m r: -g_(m+1) for the closed form, 2 B_m for the series
// g_m = A_m + B_m r^2
template <typename S>
__device__ __forceinline__ bool radial_dual(S r2, bool hole, S alpha, S alpha2, S ka, S inner2,
                                            S* g, S* gam) {
  if (r2 < inner2) {
    S sc = ka;
    const S x = alpha2 * r2;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      g[k] = sc * (x / S(2 * k + 3) - S(1) / S(2 * k + 1));
      if (k < 6) gam[k] = S(2) * sc * alpha2 / S(2 * k + 3);
      sc *= S(2) * alpha2;
    }
    return true;
  }
  const S r = sqrt_s(r2), ir2 = S(1) / r2;
  const S e = ka * exp_s(-alpha2 * r2) * ir2;
  g[0] = (hole ? -erf_s(alpha * r) : erfc_s(alpha * r)) / r;
  S sc = S(1);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    g[k + 1] = S(2 * k + 1) * g[k] * ir2 + sc * e;
    sc *= S(2) * alpha2;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) gam[k] = -g[k + 1];
  return false;
}

// one replica: target side (grad of w . a) into t, root source tensors into G.
// The directional derivative of a along w (dual numbers) is the gradient of
// w . a where a is a gradient field (the closed form); the two-term series is
// not one, and there the antisymmetric rest
// sum_n (1/(n-1)!) (gam_n + g_(n+1)) [(V_n . w) r - (r . w) V_n] is added.
template <typename S>
__device__ __forceinline__ void ewald_adj_term(const Mom22<S>& Q, S rx, S ry, S rz, S wx, S wy,
                                               S wz, bool hole, const EwaldParams& p, S* t,
                                               S* G) {
  const S r2 = rx * rx + ry * ry + rz * rz;
  const S rw = rx * wx + ry * wy + rz * wz;
  S g[7], gam[6];
  const bool series =
      radial_dual<S>(r2, hole, (S)p.alpha, (S)p.alpha2, (S)p.ka, (S)p.inner2, g, gam);
  Dual<S> gd[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) gd[k] = {g[k], gam[k] * rw};
  Dual<S> pot = {S(0), S(0)}, ax = pot, ay = pot, az = pot;
  ewald_term<S, Dual<S>>(Q, Dual<S>{rx, wx}, Dual<S>{ry, wy}, Dual<S>{rz, wz}, gd, pot, ax, ay, az);
  t[0] += ax.d;
  t[1] += ay.d;
  t[2] += az.d;
  if (series) {
    const EwaldVP<S> v = ewald_vp<S, S>(Q, rx, ry, rz);
    const S c2 = gam[2] + g[3], c3 = (gam[3] + g[4]) * S(0.5), c4 = (gam[4] + g[5]) / S(6);
    const S vx[3] = {v.v2x, v.v3x, v.v4x}, vy[3] = {v.v2y, v.v3y, v.v4y},
            vz[3] = {v.v2z, v.v3z, v.v4z}, cf[3] = {c2, c3, c4};
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const S vw = vx[n] * wx + vy[n] * wy + vz[n] * wz;
      t[0] += cf[n] * (vw * rx - rw * vx[n]);
      t[1] += cf[n] * (vw * ry - rw * vy[n]);
      t[2] += cf[n] * (vw * rz - rw * vz[n]);
    }
  }
  S pw[3][5];
  pw[0][0] = pw[1][0] = pw[2][0] = S(1);
#pragma unroll
  for (int e = 1; e < 5; ++e) {
    pw[0][e] = pw[0][e - 1] * rx;
    pw[1][e] = pw[1][e - 1] * ry;
    pw[2][e] = pw[2][e - 1] * rz;
  }
#pragma unroll
  for (int n = 0; n <= 4; ++n) {
    if (n == 1) continue;
#pragma unroll
    for (int a = n; a >= 0; --a)
#pragma unroll
      for (int b = n - a; b >= 0; --b) {
        const int c = n - a - b;
        S sw = S(0);
        if (a) sw += S(a) * wx * pw[0][a - 1] * pw[1][b] * pw[2][c];
        if (b) sw += S(b) * wy * pw[0][a] * pw[1][b - 1] * pw[2][c];
        if (c) sw += S(c) * wz * pw[0][a] * pw[1][b] * pw[2][c - 1];
        G[mono(a, b, c)] += S(inv_fact(n)) * (g[n] * sw - rw * g[n + 1] * pw[0][a] * pw[1][b] *
                                                             pw[2][c]);
      }
  }
}

constexpr int kEwAdjBlock = 256;
// monomial (a, b, c) of each reduced moment, in FMOMR order
__device__ constexpr int kRedAbc[22][3] = {
    {0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {3, 0, 0}, {1, 2, 0},
    {2, 1, 0}, {0, 3, 0}, {2, 0, 1}, {0, 2, 1}, {1, 1, 1}, {4, 0, 0}, {1, 3, 0}, {3, 1, 0},
    {0, 4, 0}, {3, 0, 1}, {0, 3, 1}, {2, 2, 0}, {2, 1, 1}, {1, 2, 1}};
constexpr int kEwAdjFixed = 38;  // G[35] (orders 0, 2, 3, 4 used), V[3]; then A[nH], B[nH]

struct EwAdjArgs {
  const double* pos64;
  const float* w;
  const int* tindex;
  double* grad;
  const EwtData* ewt;
  double* part;  // per block: kEwAdjFixed + 2 nH doubles
  int n;
  EwaldParams p;
};

template <typename T>
__device__ __forceinline__ T wave_sum_t(T v) {
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

__global__ __launch_bounds__(kEwAdjBlock) void adj_ewald_kernel(EwAdjArgs a) {
  extern __shared__ double dyn[];  // red[4][nvals], then the table (2 float4 per entry)
  __shared__ float4 cand[kEwAdjBlock / kWave][kEwMaxCand];
  __shared__ int ncand[kEwAdjBlock / kWave];
  const EwaldParams& p = a.p;
  const int nvals = kEwAdjFixed + 2 * p.nH;
  double* red = dyn;
  float4* tab = reinterpret_cast<float4*>(dyn + 4 * (size_t)nvals);
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  for (int i = tid; i < p.nH; i += kEwAdjBlock) {
    const EwtData e = a.ewt[i];
    tab[2 * i] = make_float4(e.hx, e.hy, e.hz, e.hCfac);
    tab[2 * i + 1] = make_float4(0.0f, 0.0f, 0.0f, e.hSfac);
  }
  const int gi = blockIdx.x * kEwAdjBlock + tid;
  const bool live = gi < a.n;
  const int id = live ? gi : a.n - 1;
  const double dx = a.pos64[3 * (size_t)id] - p.cmx, dy = a.pos64[3 * (size_t)id + 1] - p.cmy,
               dz = a.pos64[3 * (size_t)id + 2] - p.cmz;
  const float wx = live ? a.w[3 * (size_t)id] : 0.0f, wy = live ? a.w[3 * (size_t)id + 1] : 0.0f,
              wz = live ? a.w[3 * (size_t)id + 2] : 0.0f;
  // lattice vectors some particle of the wavefront needs (as the forward kernel)
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
    const int c = c0 + lane;
    bool need = false;
    int ix = 0, iy = 0, iz = 0;
    bool hole = false;
    if (c < NC) {
      ix = c / (W * W) - p.nEwReps;
      iy = (c / W) % W - p.nEwReps;
      iz = c % W - p.nEwReps;
      hole = abs(ix) <= p.nReps && abs(iy) <= p.nReps && abs(iz) <= p.nReps;
      const bool centre = ix == 0 && iy == 0 && iz == 0;
      const int iv[3] = {ix, iy, iz};
      float d2 = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float l = lo[k] + iv[k] * Lf, h = hi[k] + iv[k] * Lf;
        const float e = l > 0.f ? l : (h < 0.f ? -h : 0.f);
        d2 += e * e;
      }
      need = !centre && (hole || d2 <= cut2f * 1.0001f + 1e-6f * Lf * Lf);
    }
    const unsigned long long mask = __ballot(need);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (need) cand[wv][ncand[wv] + before] = make_float4(ix * Lf, iy * Lf, iz * Lf, hole ? 1.0f : 0.0f);
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
  // replicas other than the central one in FP32, then the central one in FP64
  double t64[3] = {0.0, 0.0, 0.0}, G64[35];
  {
    float t32[3] = {0.f, 0.f, 0.f}, G32[35];
#pragma unroll
    for (int q = 0; q < 35; ++q) G32[q] = 0.0f;
    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
    for (int k = 0; live && k < nc; ++k) {
      const float4 cv = cand[wv][k];
      const bool hole = cv.w != 0.0f;
      const float x = fx + cv.x, y = fy + cv.y, z = fz + cv.z;
      if (x * x + y * y + z * z > cut2f && !hole) continue;
      ewald_adj_term<float>(Qf, x, y, z, wx, wy, wz, hole, p, t32, G32);
    }
#pragma unroll
    for (int q = 0; q < 35; ++q) G64[q] = (double)G32[q];
#pragma unroll
    for (int k = 0; k < 3; ++k) t64[k] = (double)t32[k];
  }
  if (live)
    ewald_adj_term<double>(Qd, dx, dy, dz, (double)wx, (double)wy, (double)wz, p.nReps >= 0, p,
                           t64, G64);
  // k space (the forward's phase form): t += k (k.w)(hC cos + hS sin)
  const float iL = (float)(1.0 / p.L), kscale = (float)(p.L / (2.0 * M_PI));
  const float ux = (float)dx * iL, uy = (float)dy * iL, uz = (float)dz * iL;
  __syncthreads();
  for (int i = 0; i < p.nH; ++i) {
    const float4 e0 = tab[2 * i];
    const float hs = tab[2 * i + 1].w;
    float ph = rintf(e0.x * kscale) * ux + rintf(e0.y * kscale) * uy + rintf(e0.z * kscale) * uz;
    ph -= rintf(ph);
    float sn, cs;
    sincospif(2.0f * ph, &sn, &cs);
    const float kw = e0.x * wx + e0.y * wy + e0.z * wz;
    const float f = kw * (e0.w * cs + hs * sn);
    t64[0] += (double)(e0.x * f);
    t64[1] += (double)(e0.y * f);
    t64[2] += (double)(e0.z * f);
    const float A = wave_sum_t(live ? kw * sn : 0.0f);
    const float B = wave_sum_t(live ? -kw * cs : 0.0f);
    if (lane == 0) {
      red[wv * nvals + kEwAdjFixed + i] = (double)A;
      red[wv * nvals + kEwAdjFixed + p.nH + i] = (double)B;
    }
  }
  if (live) {  // one target per particle: a plain update
    double* g = a.grad + 4 * (size_t)(a.tindex ? a.tindex[id] : id);
    g[0] += t64[0];
    g[1] += t64[1];
    g[2] += t64[2];
  }
#pragma unroll
  for (int q = 0; q < 35; ++q) {
    const double s = wave_sum_t(G64[q]);
    if (lane == 0) red[wv * nvals + q] = s;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double s = wave_sum_t(live ? t64[k] : 0.0);
    if (lane == 0) red[wv * nvals + 35 + k] = s;
  }
  __syncthreads();
  for (int v = tid; v < nvals; v += kEwAdjBlock) {
    double s = 0.0;
    for (int k = 0; k < kEwAdjBlock / kWave; ++k) s += red[k * nvals + v];
    a.part[(size_t)blockIdx.x * nvals + v] = s;
  }
}

// totals over blocks, one workgroup per value: fixed strided partial sums,
// then a fixed-order tree in LDS (deterministic)
constexpr int kEwSumBlock = 256;
__global__ __launch_bounds__(kEwSumBlock) void adj_ewald_sum_kernel(const double* part,
                                                                    int nblocks, int nvals,
                                                                    double* total) {
  __shared__ double sh[kEwSumBlock];
  const int v = blockIdx.x, tid = threadIdx.x;
  double s = 0.0;
  for (int b = tid; b < nblocks; b += kEwSumBlock) s += part[(size_t)b * nvals + v];
  sh[tid] = s;
  __syncthreads();
#pragma unroll
  for (int m = kEwSumBlock / 2; m >= 1; m >>= 1) {
    if (tid < m) sh[tid] += sh[tid + m];
    __syncthreads();
  }
  if (tid == 0) total[v] = sh[0];
}

// Hhat of the Ewald term about the root's centre of mass, added to node 0
__global__ void adj_ewald_root_kernel(const double* total, const EwtData* ewt, EwaldParams p,
                                      double* node) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double G[35];
  for (int q = 0; q < 35; ++q) G[q] = total[q];
  const double* A = total + kEwAdjFixed;
  const double* B = A + p.nH;
  const double V3 = 1.0 / (p.L * p.L * p.L);
  // h-table: hC = -w_h (M - Q2.kk/2 + Q4.k^4/24), hS = w_h Q3.k^3/6
  for (int i = 0; i < p.nH; ++i) {
    const double k[3] = {ewt[i].hx, ewt[i].hy, ewt[i].hz};
    const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    const double wh = 4.0 * M_PI * V3 * exp(-k2 / (4.0 * p.alpha2)) / k2;
    G[0] -= A[i] * wh;
    for (int n = 2; n <= 4; ++n)
      for (int a = n; a >= 0; --a)
        for (int b = n - a; b >= 0; --b) {
          const int c = n - a - b;
          double kp = 1.0;
          for (int e = 0; e < a; ++e) kp *= k[0];
          for (int e = 0; e < b; ++e) kp *= k[1];
          for (int e = 0; e < c; ++e) kp *= k[2];
          const double f = n == 2 ? 0.5 * A[i] : n == 3 ? B[i] / 6.0 : -A[i] / 24.0;
          G[mono(a, b, c)] += f * wh * kp;
        }
  }
  // H: h_0 = G_0; h_al = mult(al) STF(G)_al (orders 2..4)
  double T[35];
  for (int q = 0; q < 35; ++q) T[q] = G[q];
  stf_project(T);
  double H[35];
  for (int q = 0; q < 35; ++q) H[q] = 0.0;
  H[0] = G[0];
  const double fct[5] = {1.0, 1.0, 2.0, 6.0, 24.0};
  for (int n = 2; n <= 4; ++n)
    for (int a = n; a >= 0; --a)
      for (int b = n - a; b >= 0; --b) {
        const int c = n - a - b;
        H[mono(a, b, c)] = fct[n] / (fct[a] * fct[b] * fct[c]) * T[mono(a, b, c)];
      }
  // root moments about cm (unit scale), completed
  double C[35];
  for (int q = 0; q < 35; ++q) C[q] = 0.0;
  for (int k = 0; k < 22; ++k) C[mono(kRedAbc[k][0], kRedAbc[k][1], kRedAbc[k][2])] = p.q[k];
  complete_traceless<2, 4>(C);
  const double M = p.q[0];
  double Hj[3];
  for (int j = 0; j < 3; ++j) {
    double S = 0.0;
    for (int n = 3; n <= 4; ++n)
      for (int a = n; a >= 0; --a)
        for (int b = n - a; b >= 0; --b) {
          const int c = n - a - b;
          const int e[3] = {a, b, c};
          if (e[j] == 0) continue;
          S += H[mono(a, b, c)] * e[j] *
               C[mono(a - (j == 0), b - (j == 1), c - (j == 2))];
        }
    Hj[j] = -(S + total[35 + j]) / M;
  }
  H[mono(1, 0, 0)] = Hj[0];
  H[mono(0, 1, 0)] = Hj[1];
  H[mono(0, 0, 1)] = Hj[2];
  for (int q = 0; q < 35; ++q) {
    node[kAdjHhat + q] += H[q];
    node[kAdjPoly + q] += H[q];
  }
}

__global__ void fill_zero_int_kernel(int* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
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

// owner[0, cap) of the work items whose per-owner first items are item[0 ..
// n_owners] (n_owners read on the device when n_owners_dev is set)
static int item_owners(const int* item, const int* n_owners_dev, int n_owners, int* heads,
                       int* owner, size_t cap, void* temp, size_t temp_bytes, hipStream_t s) {
  hipLaunchKernelGGL(fill_zero_int_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s,
                     heads, cap);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(adj_item_heads_kernel, dim3((unsigned)((n_owners + 255) / 256)), dim3(256), 0,
                     s, item, n_owners_dev, n_owners, heads);
  CGPU_CHECK(hipGetLastError());
  size_t tb = temp_bytes;
  CGPU_CHECK(rocprim::inclusive_scan(temp, tb, heads, owner, cap, rocprim::maximum<int>(), s));
  return 0;
}

// Scratch layout of listWork (offsets in ints, 256-byte aligned segments).
struct WorkLayout {
  size_t key, key_out, val, val_out, start, chunks, item, item_cell;
  size_t pkey, pkey_out, pval, pval_out, puniq, pcount, pstart, pnruns, pchunks, pitem, pitem_run;
  size_t item_cap, pitem_cap, item_heads, pitem_heads;
  size_t temp, total;
  size_t temp_bytes;
  unsigned end_bit, p_end_bit;
};

static size_t align64(size_t x) { return (x + 63) & ~(size_t)63; }

static unsigned bits_for(long long n) {
  unsigned b = 1;
  while ((1ll << b) < n && b < 32) ++b;
  return b;
}

// part_entries: runs of the particle list when it is an ILPart list, else 0
static hipError_t work_layout(int n_nodes, int entries, int part_entries, int n_particles,
                              WorkLayout& L) {
  L.end_bit = bits_for(n_nodes);
  L.p_end_bit = bits_for(n_particles);
  size_t sort_bytes = 0, scan_bytes = 0, psort_bytes = 0, rle_bytes = 0, pscan_bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, sort_bytes, (const unsigned*)nullptr,
                                           (unsigned*)nullptr, (const unsigned*)nullptr,
                                           (unsigned*)nullptr, (size_t)entries, 0, L.end_bit);
  if (e != hipSuccess) return e;
  e = rocprim::exclusive_scan(nullptr, scan_bytes, (const int*)nullptr, (int*)nullptr, 0,
                              (size_t)n_nodes + 1, rocprim::plus<int>());
  if (e != hipSuccess) return e;
  if (part_entries > 0) {
    e = rocprim::radix_sort_pairs(nullptr, psort_bytes, (const unsigned*)nullptr,
                                  (unsigned*)nullptr,
                                  (const unsigned long long*)nullptr,
                                  (unsigned long long*)nullptr, (size_t)part_entries, 0,
                                  L.p_end_bit);
    if (e != hipSuccess) return e;
    e = rocprim::run_length_encode(nullptr, rle_bytes, (const unsigned*)nullptr,
                                   (size_t)part_entries, (unsigned*)nullptr, (int*)nullptr,
                                   (int*)nullptr);
    if (e != hipSuccess) return e;
    e = rocprim::exclusive_scan(nullptr, pscan_bytes, (const int*)nullptr, (int*)nullptr, 0,
                                (size_t)part_entries + 1, rocprim::plus<int>());
    if (e != hipSuccess) return e;
  }
  size_t own_bytes = 0, pown_bytes = 0;
  const size_t item_cap = (size_t)n_nodes + (size_t)entries / kUsersPerItem + 1;
  const size_t pitem_cap = (size_t)part_entries + (size_t)part_entries / kRunUsersPerItem + 1;
  e = rocprim::inclusive_scan(nullptr, own_bytes, (const int*)nullptr, (int*)nullptr, item_cap,
                              rocprim::maximum<int>());
  if (e != hipSuccess) return e;
  if (part_entries > 0) {
    e = rocprim::inclusive_scan(nullptr, pown_bytes, (const int*)nullptr, (int*)nullptr,
                                pitem_cap, rocprim::maximum<int>());
    if (e != hipSuccess) return e;
  }
  size_t tb = sort_bytes;
  for (size_t x : {scan_bytes, psort_bytes, rle_bytes, pscan_bytes, own_bytes, pown_bytes})
    tb = x > tb ? x : tb;
  L.temp_bytes = tb;
  const size_t E = (size_t)entries, R = (size_t)part_entries;
  size_t o = 0;
  auto take = [&](size_t ints) {
    const size_t at = o;
    o = align64(o + ints);
    return at;
  };
  L.key = take(E);
  L.key_out = take(E);
  L.val = take(E);
  L.val_out = take(E);
  L.start = take((size_t)n_nodes + 1);
  L.chunks = take((size_t)n_nodes + 1);
  L.item = take((size_t)n_nodes + 1);
  L.item_cap = (size_t)n_nodes + E / kUsersPerItem + 1;
  L.item_cell = take(L.item_cap);
  L.item_heads = take(L.item_cap);
  L.pkey = take(R);
  L.pkey_out = take(R);
  L.pval = take(2 * R);
  L.pval_out = take(2 * R);
  L.puniq = take(R);
  L.pcount = take(R + 1);
  L.pstart = take(R + 1);
  L.pnruns = take(1);
  L.pchunks = take(R + 1);
  L.pitem = take(R + 1);
  L.pitem_cap = R + R / kRunUsersPerItem + 1;
  L.pitem_run = take(L.pitem_cap);
  L.pitem_heads = take(L.pitem_cap);
  L.temp = take((L.temp_bytes + 3) / 4);
  L.total = o;
  return hipSuccess;
}

}  // namespace cgpu

using namespace cgpu;

extern "C" size_t changaGravityAdjointWorkInts(int nNodes, int numCellEntries,
                                               int numPartEntries, int nParticles) {
  if (nNodes <= 0 || numCellEntries < 0 || numPartEntries < 0 || nParticles <= 0) return 0;
  WorkLayout L;
  if (work_layout(nNodes, numCellEntries, numPartEntries, nParticles, L) != hipSuccess) return 0;
  return L.total;
}

static size_t ewald_adj_vals(int nH) { return (size_t)kEwAdjFixed + 2 * (size_t)nH; }
static int ewald_adj_blocks(int n) { return (n + kEwAdjBlock - 1) / kEwAdjBlock; }

// target side into grad, root source side into node 0's Hhat and polynomial
static int launch_ewald_adjoint(const ChangaAdjointLaunch* a, hipStream_t s) {
  const EwaldParams p = make_params(a->ewaldRO, a->rootCm64);
  const int n = a->nTargets;
  if (n <= 0 || !a->ewt || !a->ewaldWork || !a->cells.tpos64)
    return fail("changaGravityAdjoint: Ewald term needs nTargets, ewt, ewaldWork, tpos64",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (p.nEwReps > 3 || p.nH < 0)
    return fail("changaGravityAdjoint: Ewald nEwReps > 3", hipErrorInvalidValue, __FILE__,
                __LINE__);
  const size_t nvals = ewald_adj_vals(p.nH);
  const int nb = ewald_adj_blocks(n);
  if (a->ewaldWorkDoubles < ((size_t)nb + 1) * nvals)
    return fail("changaGravityAdjoint: ewaldWork too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  // reduction rows of the four wavefronts, then the h-table
  const size_t shm = 4 * nvals * sizeof(double) + 2 * sizeof(float4) * (size_t)p.nH;
  if (shm + sizeof(float4) * (kEwAdjBlock / kWave) * kEwMaxCand + 256 > 65536)
    return fail("changaGravityAdjoint: Ewald h-table too large", hipErrorInvalidValue, __FILE__,
                __LINE__);
  EwAdjArgs ea;
  ea.pos64 = a->cells.tpos64;
  ea.w = a->w;
  ea.tindex = a->targetIndex;
  ea.grad = a->grad;
  ea.ewt = a->ewt;
  ea.part = a->ewaldWork;
  ea.n = n;
  ea.p = p;
  double* total = a->ewaldWork + (size_t)nb * nvals;
  hipLaunchKernelGGL(adj_ewald_kernel, dim3(nb), dim3(kEwAdjBlock), shm, s, ea);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(adj_ewald_sum_kernel, dim3((unsigned)nvals), dim3(kEwSumBlock), 0, s,
                     a->ewaldWork, nb, (int)nvals, total);
  CGPU_CHECK(hipGetLastError());
  hipLaunchKernelGGL(adj_ewald_root_kernel, dim3(1), dim3(kWave), 0, s, total, a->ewt, p,
                     a->nodeWork);
  CGPU_CHECK(hipGetLastError());
  return 0;
}

extern "C" size_t changaGravityAdjointEwaldWorkDoubles(int nTargets, int nH) {
  if (nTargets <= 0 || nH < 0) return 0;
  return ((size_t)ewald_adj_blocks(nTargets) + 1) * ewald_adj_vals(nH);
}

extern "C" int changaGravityAdjoint(const ChangaAdjointLaunch* a, void* stream) {
  if (a && a->cells.numBuckets == 0 && a->grad && a->nParticles > 0)
    return zero(a->grad, 4 * (size_t)a->nParticles, (hipStream_t)stream);  // no targets: L = 0
  if (!a || !a->w || !a->grad || !a->nodeWork || !a->listWork || !a->levelOffsets || !a->first ||
      !a->count || !a->child || !a->node64 || !a->pos64 || !a->mass64 || a->nNodes <= 0 ||
      a->nDepths <= 0 || a->nParticles <= 0 || a->numCellEntries < 0 || a->numPartEntries < 0)
    return fail("changaGravityAdjoint: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  const ChangaGravityLaunch& pc = a->cells;
  const ChangaGravityLaunch& pp = a->parts;
  if (!pc.marks || !pc.starts || !pc.sizes || !pc.targets || !pc.moments || !pc.list ||
      !pc.tpos64 || !pc.cm64 || !pp.marks || !pp.list || !pp.sources || !pp.spos64 ||
      pc.numBuckets != pp.numBuckets || pc.starts != pp.starts || pc.sizes != pp.sizes ||
      pc.fperiod != pp.fperiod || pc.listKind != 0 || (pp.listKind != 0 && pp.listKind != 1))
    return fail("changaGravityAdjoint: cell/particle launches (precision form, same buckets)",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (pc.numBuckets >= kMaxAdjBuckets)
    return fail("changaGravityAdjoint: more than 2^22 buckets", hipErrorInvalidValue, __FILE__,
                __LINE__);
  for (const ChangaGravityLaunch* l : {&pc, &pp})
    if ((l->fperiodY != 0.0f && l->fperiodY != l->fperiod) ||
        (l->fperiodZ != 0.0f && l->fperiodZ != l->fperiod))
      return fail("changaGravityAdjoint: the adjoint takes one period on all three axes",
                  hipErrorInvalidValue, __FILE__, __LINE__);
  // the ILPart particle list is transposed by source run (its source side
  // needs no atomics per pair); per-particle ILCell lists keep the per-bucket
  // source pass
  const int R = pp.listKind == 1 ? a->numPartEntries : 0;
  WorkLayout W;
  CGPU_CHECK(work_layout(a->nNodes, a->numCellEntries, R, a->nParticles, W));
  if (a->listWorkInts < W.total)
    return fail("changaGravityAdjoint: listWork too small", hipErrorInvalidValue, __FILE__,
                __LINE__);
  hipStream_t s = (hipStream_t)stream;
  const int nn = a->nNodes, nb = pc.numBuckets, E = a->numCellEntries;
  if (int e = zero(a->grad, 4 * (size_t)a->nParticles, s)) return e;
  if (int e = zero(a->nodeWork, (size_t)kAdjStride * nn, s)) return e;
  int* iw = a->listWork;
  unsigned* key = reinterpret_cast<unsigned*>(iw + W.key);
  unsigned* key_out = reinterpret_cast<unsigned*>(iw + W.key_out);
  unsigned* val = reinterpret_cast<unsigned*>(iw + W.val);
  unsigned* val_out = reinterpret_cast<unsigned*>(iw + W.val_out);
  AdjArgs args;
  args.pc = pc;
  args.pp = pp;
  args.w = a->w;
  args.tindex = a->targetIndex;
  args.grad = a->grad;
  args.node = a->nodeWork;
  args.nNodes = nn;
  args.entries = E;
  args.tr.key = key_out;
  args.tr.value = val_out;
  args.tr.start = iw + W.start;
  args.tr.chunks = iw + W.chunks;
  args.tr.item = iw + W.item;
  args.tr.item_cell = iw + W.item_cell;
  args.ptr.value = reinterpret_cast<const unsigned long long*>(iw + W.pval_out);
  args.ptr.key = iw + W.puniq;
  args.ptr.start = iw + W.pstart;
  args.ptr.nruns = iw + W.pnruns;
  args.ptr.item = iw + W.pitem;
  args.ptr.item_run = iw + W.pitem_run;
  // (1) transpose: entries sorted by cell, stable (bucket order within a cell)
  if (nb > 0) {
    hipLaunchKernelGGL(adj_keys_kernel, dim3(nb), dim3(kWave), 0, s, pc, key, val);
    CGPU_CHECK(hipGetLastError());
  }
  if (E > 0) {
    size_t tb = W.temp_bytes;
    CGPU_CHECK(rocprim::radix_sort_pairs(iw + W.temp, tb, key, key_out, val, val_out, (size_t)E,
                                         0, W.end_bit, s));
  }
  {
    const int n = E > nn + 1 ? E : nn + 1;
    hipLaunchKernelGGL(adj_bounds_kernel, dim3((n + 255) / 256), dim3(256), 0, s, key_out, E, nn,
                       args.tr.start);
    CGPU_CHECK(hipGetLastError());
    hipLaunchKernelGGL(adj_chunks_kernel, dim3((nn + 1 + 255) / 256), dim3(256), 0, s,
                       args.tr.start, nn, args.tr.chunks);
    CGPU_CHECK(hipGetLastError());
    size_t tb = W.temp_bytes;
    CGPU_CHECK(rocprim::exclusive_scan(iw + W.temp, tb, args.tr.chunks, args.tr.item, 0,
                                       (size_t)nn + 1, rocprim::plus<int>(), s));
    if (int e = item_owners(args.tr.item, nullptr, nn, iw + W.item_heads, args.tr.item_cell,
                            W.item_cap, iw + W.temp, W.temp_bytes, s))
      return e;
  }
  // (1b) runs sorted by first source particle, stable; distinct starts and
  // their user ranges
  if (R > 0) {
    unsigned* pkey = reinterpret_cast<unsigned*>(iw + W.pkey);
    unsigned* pkey_out = reinterpret_cast<unsigned*>(iw + W.pkey_out);
    unsigned long long* pval = reinterpret_cast<unsigned long long*>(iw + W.pval);
    unsigned long long* pval_out = reinterpret_cast<unsigned long long*>(iw + W.pval_out);
    int* pcount = iw + W.pcount;
    hipLaunchKernelGGL(adj_pkeys_kernel, dim3(nb), dim3(kWave), 0, s, pp, pkey, pval);
    CGPU_CHECK(hipGetLastError());
    size_t tb = W.temp_bytes;
    CGPU_CHECK(rocprim::radix_sort_pairs(iw + W.temp, tb, pkey, pkey_out, pval, pval_out,
                                         (size_t)R, 0, W.p_end_bit, s));
    hipLaunchKernelGGL(fill_zero_int_kernel, dim3((unsigned)((R + 1 + 255) / 256)), dim3(256), 0,
                       s, pcount, (size_t)R + 1);
    CGPU_CHECK(hipGetLastError());
    tb = W.temp_bytes;
    CGPU_CHECK(rocprim::run_length_encode(iw + W.temp, tb, pkey_out, (size_t)R,
                                          reinterpret_cast<unsigned*>(iw + W.puniq), pcount,
                                          iw + W.pnruns, s));
    tb = W.temp_bytes;
    CGPU_CHECK(rocprim::exclusive_scan(iw + W.temp, tb, pcount, iw + W.pstart, 0, (size_t)R + 1,
                                       rocprim::plus<int>(), s));
    hipLaunchKernelGGL(adj_pchunks_kernel, dim3((unsigned)((R + 1 + 255) / 256)), dim3(256), 0, s,
                       pcount, iw + W.pnruns, R + 1, iw + W.pchunks);
    CGPU_CHECK(hipGetLastError());
    tb = W.temp_bytes;
    CGPU_CHECK(rocprim::exclusive_scan(iw + W.temp, tb, iw + W.pchunks, iw + W.pitem, 0,
                                       (size_t)R + 1, rocprim::plus<int>(), s));
    if (int e = item_owners(iw + W.pitem, iw + W.pnruns, R, iw + W.pitem_heads, iw + W.pitem_run,
                            W.pitem_cap, iw + W.temp, W.temp_bytes, s))
      return e;
  }
  // (2) cell sources, (3) target side and particle pairs
  if (nb > 0) {
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
    if (R > 0) {
      hipLaunchKernelGGL(adj_pp_sources_kernel, dim3(ncu * 16), dim3(kWave), 0, s, args);
      CGPU_CHECK(hipGetLastError());
    }
  }
  // (4) Hhat per node (plus the Ewald term's at the root), then top-down
  hipLaunchKernelGGL(adj_hhat_kernel, dim3((nn + 127) / 128), dim3(128), 0, s, nn, a->nodeWork,
                     a->node64);
  CGPU_CHECK(hipGetLastError());
  if (a->ewaldRO)
    if (int e = launch_ewald_adjoint(a, s)) return e;
  for (int d = 0; d < a->nDepths; ++d) {
    const int lo = a->levelOffsets[d], hi = a->levelOffsets[d + 1];
    if (hi <= lo) continue;
    hipLaunchKernelGGL(adj_shift_kernel, dim3((2 * (hi - lo) + 127) / 128), dim3(128), 0, s, lo,
                       hi, a->child, a->nodeWork, a->node64);
    CGPU_CHECK(hipGetLastError());
  }
  // (5) particles
  hipLaunchKernelGGL(adjoint_particles_kernel, dim3((nn + 127) / 128), dim3(128), 0, s, nn,
                     a->child, a->first, a->count, a->nodeWork, a->node64, a->pos64, a->mass64,
                     a->grad);
  CGPU_CHECK(hipGetLastError());
  return 0;
}
