// Particle-cell (hexadecapole) and particle-particle gravity kernels.
//
// CPU semantics followed: nodeBucketForce (gravity.h:398-475) with
// momEvalFmomrcm (moments.c:1469-1525) for cells, partBucketForce
// (gravity.h:266-303) with the spline of gravity.h:147-182 for particles and
// for cells whose softening overlaps the bucket (gravity.h:414-421). The GPU
// work decomposition the reference uses for the same lists is
// nodeGravityComputation / particleGravityComputation (HostCUDA.cu:1008-1205,
// 1565-1751).
//
// Decomposition. One 64-lane wavefront (one workgroup) owns one target
// bucket: lanes are split into `nslot` slots of `nt` targets (nt = bucket
// size, nslot = 64 / nt), so a 16-particle bucket keeps 4 source slots busy
// and an 11-particle bucket 5. Sources are staged through LDS in tiles: the
// cells' 27-float records are fetched with coalesced flat loads, pre-scaled
// (moments times n c_n radius^n, replica offset applied) once per tile, and
// read back as broadcast ds_read_b128. Each lane accumulates its slot's
// contributions in FP32 within a tile and folds them into FP64 after each
// tile; the slots are reduced in a fixed order through LDS, so the result is
// deterministic. Accelerations and potentials are added into the
// VariablePartData array (+=), dtGrav takes the max.
#include "common.h"

namespace cgpu {

constexpr int kPcTile = 64;   // cells per staged tile (one lane stages one cell)
constexpr int kPcRec = 36;    // floats per staged cell record (9 x float4; 36 keeps the
                              // records read together 4 LDS banks apart)
constexpr int kPpTile = 256;  // source particles per staged tile

// Staged cell record. Moments of order n are multiplied by n c_n radius^n
// (c_n = (2n-1)!!/n!), the replica offset is applied to the centre, and the
// combinations the trace-free contraction needs are formed once per cell.
// A cell whose softening overlaps the bucket is stored with negative mass
// and its softening in slot 4.
enum {
  R_CX = 0, R_CY, R_CZ, R_M,
  R_Q2 = 4,    // xx xy xz yy yz, then -(xx+yy)
  R_Q3 = 10,   // xxx xyy xxy yyy xxz yyz xyz, then -(xxx+xyy), -(xxy+yyy)
  R_Q4 = 19,   // xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz,
               // then -(xxxx+xxyy), -(yyyy+xxyy), -(xxxy+xyyy)
  R_SOFT = 4
};

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

struct Acc32 {
  float ax, ay, az, pot, idt;
};

// Hernquist-Katz spline (gravity.h:147-182) in FP32; returns potential kernel
// a and force kernel b for separation^2 r2 and softening sum twoh.
__device__ __forceinline__ void spline32(float r2, float twoh, float& a, float& b) {
  float ir = rsq(r2);
  if (r2 < twoh * twoh) {
    float r = r2 * ir;
    float dih = 2.0f / twoh;
    float u = r * dih, u2 = u * u;
    if (u < 1.0f) {
      a = dih * (7.0f / 5.0f + u2 * (-2.0f / 3.0f + u2 * (3.0f / 10.0f - u * (1.0f / 10.0f))));
      b = dih * dih * dih * (4.0f / 3.0f + u2 * (-6.0f / 5.0f + u * 0.5f));
    } else {
      a = -(1.0f / 15.0f) * ir +
          dih * (8.0f / 5.0f + u2 * (-4.0f / 3.0f + u * (1.0f + u * (-3.0f / 10.0f + u * (1.0f / 30.0f)))));
      b = -(1.0f / 15.0f) * ir * ir * ir +
          dih * dih * dih * (8.0f / 3.0f + u * (-3.0f + u * (6.0f / 5.0f - u * (1.0f / 6.0f))));
    }
  } else {
    a = ir;
    b = ir * ir * ir;
  }
}

// One source particle (or softened-cell monopole) at separation r = src - tgt.
__device__ __forceinline__ void pp_interact(float rx, float ry, float rz, float ms, float soft_s,
                                            float mt, float soft_t, Acc32& A) {
  float r2 = rx * rx + ry * ry + rz * rz;
  if (r2 != 0.0f) {
    float a, b;
    spline32(r2, soft_s + soft_t, a, b);
    float bm = b * ms;
    A.ax += rx * bm;
    A.ay += ry * bm;
    A.az += rz * bm;
    A.pot -= ms * a;
    A.idt = fmaxf(A.idt, (mt + ms) * b);
  }
}

// Hexadecapole cell at r = tgt - cm from a staged record (see the enum).
// With n = r/|r|, V_n = Q_n . n^(n-1) follows from the rank-(n-1) harmonic
// polynomials of n (the gradients of the harmonic polynomials Q_n . n^n, z
// eliminated by tracelessness), P_n = V_n . n, and
//   phi -= dir (m + dir^2 (P2/2 + dir (P3/3 + dir P4/4)))
//   a   += dir^4 (V2 + dir (V3 + dir V4)) - dir^2 S n,
//   S = m + dir^2 (5/2 P2 + dir (7/3 P3 + dir 9/4 P4)).
__device__ __forceinline__ void pc_interact(float rx, float ry, float rz, const float4* rec,
                                            float mt, Acc32& A) {
  float r2 = rx * rx + ry * ry + rz * rz;
  if (r2 == 0.0f) return;
  float dir = rsq(r2);
  const float4 c0 = rec[0], c1 = rec[1], c2 = rec[2], c3 = rec[3];
  const float4 c4 = rec[4], c5 = rec[5], c6 = rec[6], c7 = rec[7];
  const float m = c0.w;
  const float qxx = c1.x, qxy = c1.y, qxz = c1.z, qyy = c1.w, qyz = c2.x, qzz = c2.y;
  const float qxxx = c2.z, qxyy = c2.w, qxxy = c3.x, qyyy = c3.y, qxxz = c3.z, qyyz = c3.w,
              qxyz = c4.x, q3a = c4.y, q3b = c4.z;
  const float qxxxx = c4.w, qxyyy = c5.x, qxxxy = c5.y, qyyyy = c5.z, qxxxz = c5.w, qyyyz = c6.x,
              qxxyy = c6.y, qxxyz = c6.z, qxyyz = c6.w, q4a = c7.x, q4b = c7.y, q4c = c7.z;
  const float x = rx * dir, y = ry * dir, z = rz * dir;
  const float x2 = x * x, y2 = y * y, z2 = z * z;
  // rank-2 harmonic polynomials
  const float fxx = x2 - z2, fyy = y2 - z2;
  const float x_2 = x + x, y_2 = y + y;
  const float fxy = x_2 * y, fxz = x_2 * z, fyz = y_2 * z;
  // rank-3 harmonic polynomials
  const float z3 = 3.0f * z2;
  const float fxxx = x * (x2 - z3), fyyy = y * (y2 - z3);
  const float fxyy = (3.0f * x) * fyy, fxxy = (3.0f * y) * fxx;
  const float fxxz = z * (3.0f * x2 - z2), fyyz = z * (3.0f * y2 - z2);
  const float fxyz = (3.0f * z) * fxy;
  // V2 = Q2 . n
  const float v2x = qxx * x + qxy * y + qxz * z;
  const float v2y = qxy * x + qyy * y + qyz * z;
  const float v2z = qxz * x + qyz * y + qzz * z;
  // V3 = Q3 . n n
  const float v3x = qxxx * fxx + qxyy * fyy + qxxy * fxy + qxxz * fxz + qxyz * fyz;
  const float v3y = qxyy * fxy + qxxy * fxx + qyyy * fyy + qyyz * fyz + qxyz * fxz;
  const float v3z = qxxz * fxx + qyyz * fyy + qxyz * fxy + q3a * fxz + q3b * fyz;
  // V4 = Q4 . n n n
  const float v4x = qxxxx * fxxx + qxxyy * fxyy + qxxxy * fxxy + qxyyy * fyyy + qxxxz * fxxz +
                    qxyyz * fyyz + qxxyz * fxyz;
  const float v4y = qyyyy * fyyy + qxxyy * fxxy + qxxxy * fxxx + qxyyy * fxyy + qxyyz * fxyz +
                    qxxyz * fxxz + qyyyz * fyyz;
  const float v4z = qxxxz * fxxx + qxyyz * fxyy + qxxyz * fxxy + qyyyz * fyyy + q4a * fxxz +
                    q4b * fyyz + q4c * fxyz;
  const float p2 = v2x * x + v2y * y + v2z * z;
  const float p3 = v3x * x + v3y * y + v3z * z;
  const float p4 = v4x * x + v4y * y + v4z * z;
  const float d2 = dir * dir;
  const float pin = m + d2 * (0.5f * p2 + dir * ((1.0f / 3.0f) * p3 + (0.25f * dir) * p4));
  const float S = m + d2 * (2.5f * p2 + dir * ((7.0f / 3.0f) * p3 + (2.25f * dir) * p4));
  const float e = d2 * d2, f = S * d2;
  A.pot -= dir * pin;
  A.ax += e * (v2x + dir * (v3x + dir * v4x)) - f * x;
  A.ay += e * (v2y + dir * (v3y + dir * v4y)) - f * y;
  A.az += e * (v2z + dir * (v3z + dir * v4z)) - f * z;
  A.idt = fmaxf(A.idt, (mt + m) * (d2 * dir));
}

// Slot reduction in a fixed order and the += into the result array.
__device__ __forceinline__ void reduce_store(double (*red)[kWave], float* redi, int lane, int nt,
                                             int nslot, int slot, double ax, double ay, double az,
                                             double pot, float idt, VariablePartData* out) {
  bool live = slot < nslot;
  red[0][lane] = live ? ax : 0.0;
  red[1][lane] = live ? ay : 0.0;
  red[2][lane] = live ? az : 0.0;
  red[3][lane] = live ? pot : 0.0;
  redi[lane] = live ? idt : 0.0f;
  __syncthreads();
  if (lane < nt) {
    double sx = 0, sy = 0, sz = 0, sp = 0;
    float si = 0.0f;
    for (int s = 0; s < nslot; ++s) {
      int l = s * nt + lane;
      sx += red[0][l];
      sy += red[1][l];
      sz += red[2][l];
      sp += red[3][l];
      si = fmaxf(si, redi[l]);
    }
    VariablePartData v = *out;
    v.a.x += (float)sx;
    v.a.y += (float)sy;
    v.a.z += (float)sz;
    v.potential += (float)sp;
    v.dtGrav = fmaxf(v.dtGrav, si);
    *out = v;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Particle-cell kernel. Tiles of 64 cells: lane c loads cell c's 27-float
// wire record and its ILCell, builds the staged record in registers and
// writes it to LDS; the next tile's wire record is already in flight while
// the current tile is evaluated. A tile without softened cells (ballot) runs
// the hexadecapole loop with no per-cell test.
// ---------------------------------------------------------------------------
struct WireCell {
  float v[27];
  int oid;
  int idx;
};

template <bool PRECISE>
__device__ __forceinline__ void load_cell(const ChangaGravityLaunch& a, const ILCell* cells, int e,
                                          int end, WireCell& w) {
  if (e < end) {
    ILCell c = cells[e];
    w.idx = c.index;
    w.oid = c.offsetID;
    const float* src = reinterpret_cast<const float*>(a.moments) + (size_t)c.index * 27;
#pragma unroll
    for (int k = 0; k < 27; ++k) w.v[k] = src[k];
  }
}

// wire record -> staged record (in registers), then to LDS
template <bool PRECISE>
__device__ __forceinline__ void stage_cell(const ChangaGravityLaunch& a, const WireCell& w, float L,
                                           double cx0, double cy0, double cz0, float4* dst) {
  // wire: 0 radius, 1 soft, 2 mass, 3-5 cm, 6-10 Q2 (xx xy xz yy yz),
  // 11-17 Q3 (xxx xyy xxy yyy xxz yyz xyz), 18-26 Q4
  float r[kPcRec];
  const int oid = w.oid;
  if (PRECISE) {
    const size_t ci = (size_t)w.idx * 3;
    r[R_CX] = (float)(a.cm64[ci] + (double)off_x(oid) * (double)L - cx0);
    r[R_CY] = (float)(a.cm64[ci + 1] + (double)off_y(oid) * (double)L - cy0);
    r[R_CZ] = (float)(a.cm64[ci + 2] + (double)off_z(oid) * (double)L - cz0);
  } else {
    r[R_CX] = (float)off_x(oid) * L + w.v[3];
    r[R_CY] = (float)off_y(oid) * L + w.v[4];
    r[R_CZ] = (float)off_z(oid) * L + w.v[5];
  }
  const float u = w.v[0], u2 = u * u;
  const float s2 = 3.0f * u2, s3 = 7.5f * u2 * u, s4 = 17.5f * u2 * u2;
#pragma unroll
  for (int k = 0; k < 5; ++k) r[R_Q2 + k] = s2 * w.v[6 + k];
  r[R_Q2 + 5] = -(r[R_Q2 + 0] + r[R_Q2 + 3]);
#pragma unroll
  for (int k = 0; k < 7; ++k) r[R_Q3 + k] = s3 * w.v[11 + k];
  r[R_Q3 + 7] = -(r[R_Q3 + 0] + r[R_Q3 + 1]);
  r[R_Q3 + 8] = -(r[R_Q3 + 2] + r[R_Q3 + 3]);
#pragma unroll
  for (int k = 0; k < 9; ++k) r[R_Q4 + k] = s4 * w.v[18 + k];
  r[R_Q4 + 9] = -(r[R_Q4 + 0] + r[R_Q4 + 6]);
  r[R_Q4 + 10] = -(r[R_Q4 + 3] + r[R_Q4 + 6]);
  r[R_Q4 + 11] = -(r[R_Q4 + 2] + r[R_Q4 + 1]);
#pragma unroll
  for (int k = R_Q4 + 12; k < kPcRec; ++k) r[k] = 0.0f;
  r[R_M] = w.v[2];
  if (oid & CHANGA_ILCELL_SOFTENED) {
    r[R_M] = -w.v[2];
    r[R_SOFT] = w.v[1];
  }
#pragma unroll
  for (int k = 0; k < kPcRec / 4; ++k)
    dst[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

template <bool PRECISE>
__global__ __launch_bounds__(64) void pc_kernel(ChangaGravityLaunch a) {
  // staged tile; reused for the slot reduction at the end of each bucket
  __shared__ float4 rec4[kPcTile * (kPcRec / 4)];
  static_assert(sizeof(rec4) >= 4 * kWave * sizeof(double) + kWave * sizeof(float), "reduction");
  double(*red)[kWave] = reinterpret_cast<double(*)[kWave]>(rec4);
  float* redi = reinterpret_cast<float*>(reinterpret_cast<double*>(rec4) + 4 * kWave);

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  const float L = a.fperiod;
  const ILCell* cells = static_cast<const ILCell*>(a.list);

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

    WireCell w;
    load_cell<PRECISE>(a, cells, start + lane, end, w);
    for (int base = start; base < end; base += kPcTile) {
      const int nc = min(kPcTile, end - base);
      const bool mine = lane < nc;
      const bool soft_tile = __ballot(mine && (w.oid & CHANGA_ILCELL_SOFTENED)) != 0ull;
      if (mine) stage_cell<PRECISE>(a, w, L, cx0, cy0, cz0, rec4 + lane * (kPcRec / 4));
      __syncthreads();
      // the next tile's records travel while this one is evaluated (issued
      // after the barrier, whose fence would otherwise wait for them)
      load_cell<PRECISE>(a, cells, base + kPcTile + lane, end, w);
      if (live) {
        Acc32 A = {0.f, 0.f, 0.f, 0.f, idt};
        if (!soft_tile) {
          for (int c = slot; c < nc; c += nslot) {
            const float4* r4 = rec4 + c * (kPcRec / 4);
            const float4 h = r4[0];
            pc_interact(tx - h.x, ty - h.y, tz - h.z, r4, tm, A);
          }
        } else {
          for (int c = slot; c < nc; c += nslot) {
            const float4* r4 = rec4 + c * (kPcRec / 4);
            const float4 h = r4[0];
            if (h.w < 0.0f)  // softened cell: spline monopole at the cm
              pp_interact(h.x - tx, h.y - ty, h.z - tz, -h.w, r4[1].x, tm, tsoft, A);
            else
              pc_interact(tx - h.x, ty - h.y, tz - h.z, r4, tm, A);
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
        ILCell c = cells[base + lane];
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
// Particle-particle kernel. KIND 0: one ILCell per source particle (the list
// ChaNGa's GenericList<ILPart>::serialize produces, Compute.cpp:1118-1218);
// KIND 1: ILPart runs expanded here through a wavefront prefix sum.
// ---------------------------------------------------------------------------
template <bool PRECISE, int KIND>
__global__ __launch_bounds__(64) void pp_kernel(ChangaGravityLaunch a) {
  __shared__ float4 src4[kPpTile];
  __shared__ float srcs[kPpTile];
  __shared__ int r_idx[kWave], r_oid[kWave], r_end[kWave];
  __shared__ double red[4][kWave];
  __shared__ float redi[kWave];

  const int lane = threadIdx.x;
  const int b = xcd_bucket(blockIdx.x, a.numBuckets);
  const int start = a.marks[b], end = a.marks[b + 1];
  const int bstart = a.starts[b], n = a.sizes[b];
  const float L = a.fperiod;

  double cx0 = 0, cy0 = 0, cz0 = 0;
  if (PRECISE && n > 0) {
    cx0 = a.tpos64[3 * (size_t)bstart];
    cy0 = a.tpos64[3 * (size_t)bstart + 1];
    cz0 = a.tpos64[3 * (size_t)bstart + 2];
  }

  // stage source s (particle index p, replica oid) into tile slot i
  auto stage = [&](int i, int p, int oid) {
    CompactPartData s = a.sources[p];
    float4 v;
    if (PRECISE) {
      v.x = (float)(a.spos64[3 * (size_t)p] + (double)off_x(oid) * (double)L - cx0);
      v.y = (float)(a.spos64[3 * (size_t)p + 1] + (double)off_y(oid) * (double)L - cy0);
      v.z = (float)(a.spos64[3 * (size_t)p + 2] + (double)off_z(oid) * (double)L - cz0);
    } else {
      v.x = (float)off_x(oid) * L + s.position.x;
      v.y = (float)off_y(oid) * L + s.position.y;
      v.z = (float)off_z(oid) * L + s.position.z;
    }
    v.w = s.mass;
    src4[i] = v;
    srcs[i] = s.soft;
  };

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

    auto consume = [&](int ns) {
      if (live) {
        Acc32 A = {0.f, 0.f, 0.f, 0.f, idt};
        for (int s = slot; s < ns; s += nslot) {
          float4 v = src4[s];
          pp_interact(v.x - tx, v.y - ty, v.z - tz, v.w, srcs[s], tm, tsoft, A);
        }
        dax += A.ax;
        day += A.ay;
        daz += A.az;
        dpot += A.pot;
        idt = A.idt;
      }
    };

    if (KIND == 0) {
      const ILCell* lst = static_cast<const ILCell*>(a.list);
      for (int base = start; base < end; base += kPpTile) {
        const int ns = min(kPpTile, end - base);
        for (int i = lane; i < ns; i += kWave) {
          ILCell c = lst[base + i];
          stage(i, c.index, c.offsetID);
        }
        __syncthreads();
        consume(ns);
        __syncthreads();
      }
    } else {
      const ILPart* lst = static_cast<const ILPart*>(a.list);
      for (int rb = start; rb < end; rb += kWave) {
        const int nr = min(kWave, end - rb);
        int num = 0;
        if (lane < nr) {
          ILPart p = lst[rb + lane];
          r_idx[lane] = p.index;
          r_oid[lane] = p.off;
          num = p.num;
        }
        // inclusive prefix sum of run lengths across the wavefront
        int incl = num;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          int y = __shfl_up(incl, d, kWave);
          if (lane >= d) incl += y;
        }
        r_end[lane] = incl;
        const int total = __shfl(incl, kWave - 1, kWave);
        __syncthreads();
        for (int sb = 0; sb < total; sb += kPpTile) {
          const int ns = min(kPpTile, total - sb);
          for (int i = lane; i < ns; i += kWave) {
            int s = sb + i;
            // first run whose inclusive end exceeds s
            int lo = 0, hi = nr - 1;
            while (lo < hi) {
              int mid = (lo + hi) >> 1;
              if (r_end[mid] > s) hi = mid;
              else lo = mid + 1;
            }
            int begin = lo ? r_end[lo - 1] : 0;
            stage(i, r_idx[lo] + (s - begin), r_oid[lo]);
          }
          __syncthreads();
          consume(ns);
          __syncthreads();
        }
      }
    }
    reduce_store(red, redi, lane, nt, nslot, slot, dax, day, daz, dpot, idt,
                 a.vars + bstart + t0 + (lane < nt ? lane : 0));
  }
}

__global__ void zero_vars_kernel(VariablePartData* v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = VariablePartData{{0.f, 0.f, 0.f}, 0.f, 0.f};
}

static int check_launch(const ChangaGravityLaunch* a) {
  if (!a || a->numBuckets < 0 || !a->marks || !a->starts || !a->sizes || !a->vars ||
      !a->targets)
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

extern "C" int changaZeroVars(VariablePartData* vars, int n, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(zero_vars_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     vars, n);
  CGPU_CHECK(hipGetLastError());
  return 0;
}
