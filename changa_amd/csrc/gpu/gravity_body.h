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
#pragma once
#include <utility>

#include "common.h"

namespace cgpu {

constexpr int kPcTile = 64;   // cells per staged tile of the plain kernel
constexpr int kPcRec = 36;    // floats per staged cell record of the adjoint's target pass
                              // (9 x float4, stored whole: 36 keeps records read together
                              // 4 LDS banks apart)
constexpr int kFwdRec = 32;   // floats per staged cell record of the forward loops
                              // (8 float4 component planes)
constexpr int kPpTile = 256;  // source particles per staged tile

// Staged cell record of the adjoint's target pass (36 floats). Moments of
// order n are multiplied by n c_n radius^n (c_n = (2n-1)!!/n!), the replica
// offset is applied to the centre, and the combinations the trace-free
// contraction needs are formed once per cell. A cell whose softening
// overlaps the bucket is stored with negative mass and its softening in
// slot 4.
enum {
  R_CX = 0, R_CY, R_CZ, R_M,
  R_Q2 = 4,    // xx xy xz yy yz, then -(xx+yy)
  R_Q3 = 10,   // xxx xyy xxy yyy xxz yyz xyz, then -(xxx+xyy), -(xxy+yyy)
  R_Q4 = 19,   // xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz,
               // then -(xxxx+xxyy), -(yyyy+xxyy), -(xxxy+xyyy)
  R_SOFT = 4
};

// Staged cell record of the forward loops (32 floats, 8 float4): the same
// scaled moments and trace combinations, with the centre (replica offset
// applied, relative to the bucket centre in the precision form) in the first
// three floats, the mass in F_M and, in F_SOFT, the cell's softening when
// its softening overlaps the bucket (a spline monopole) or -1 for a
// hexadecapole cell.
enum {
  F_CX = 0, F_CY, F_CZ,
  F_Q4A = 3, F_Q4B, F_Q4C,  // -(xxxx+xxyy), -(yyyy+xxyy), -(xxxy+xyyy)
  F_M = 6, F_SOFT = 7,
  F_Q2 = 8,    // xx xy xz yy yz, then -(xx+yy)
  F_Q3 = 14,   // xxx xyy xxy yyy xxz yyz xyz, then -(xxx+xyy), -(xxy+yyy)
  F_Q4 = 23    // xxxx xyyy xxxy yyyy xxxz yyyz xxyy xxyz xyyz
};

// Cell record in device memory (changaCellRecords): 128 bytes, one cache
// line per cell, so that a tile of cells is fetched as whole lines (eight
// lanes per record) instead of one record per lane. Layout: the centre of
// mass as 3 doubles (floats 0-5), mass (6), softening (7), then floats
// F_Q2.. of the forward record. The three Q4 combinations and the centre
// relative to the bucket are formed when a tile is staged.
constexpr int kCellRecFloats = 32;

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

struct Acc32 {
  float ax, ay, az, pot, idt;
};

// max of two non-negative floats (dtGrav terms: masses and 1/r^3 are
// non-negative) on their bit patterns, which order like the values: one
// integer max, with no NaN canonicalisation of the loop-carried operand.
__device__ __forceinline__ float max_nonneg(float a, float b) {
  return __builtin_bit_cast(float, max(__builtin_bit_cast(unsigned, a),
                                       __builtin_bit_cast(unsigned, b)));
}

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
    A.idt = max_nonneg(A.idt, __builtin_fmaf(mt, b, bm));
  }
}

// Hexadecapole cell at r = tgt - cm from a staged forward record (F_ enum).
// With the harmonic polynomials of r (not of the unit vector: the powers of
// 1/|r| are applied at the end), V_n = Q_n . r^(n-1) (the gradients of the
// harmonic polynomials Q_n . r^n, z eliminated by tracelessness),
// P_n = V_n . r, d = 1/|r|, and
//   phi -= d (m + d^4 (P2/2 + d^2 (P3/3 + d^2 P4/4)))
//   a   += d^5 (V2 + d^2 (V3 + d^2 V4)) - d^3 S r,
//   S = m + d^4 (5/2 P2 + d^2 (7/3 P3 + d^2 9/4 P4)).
// The record's float4 k is rec[k * STRIDE] (STRIDE 1: records stored whole;
// STRIDE = tile width: component planes), read by load_rec.
// The staged record of one cell, read once per cell and evaluated for every
// target a lane holds.
struct PcRec {
  float4 c0, c1, c2, c3, c4, c5, c6, c7;
};

template <int STRIDE = 1>
__device__ __forceinline__ PcRec load_rec(const float4* rec) {
  PcRec r;
  r.c0 = rec[0];
  r.c1 = rec[STRIDE];
  r.c2 = rec[2 * STRIDE];
  r.c3 = rec[3 * STRIDE];
  r.c4 = rec[4 * STRIDE];
  r.c5 = rec[5 * STRIDE];
  r.c6 = rec[6 * STRIDE];
  r.c7 = rec[7 * STRIDE];
  return r;
}

__device__ __forceinline__ void pc_eval(float x, float y, float z, const PcRec& R, float mt,
                                        Acc32& A) {
  const float r2 = x * x + y * y + z * z;
  if (r2 == 0.0f) return;
  const float dir = rsq(r2);
  const float4 c0 = R.c0, c1 = R.c1, c2 = R.c2, c3 = R.c3, c4 = R.c4, c5 = R.c5, c6 = R.c6,
               c7 = R.c7;
  // forward record: c0 (cx cy cz q4a), c1 (q4b q4c m soft), then Q2, Q3, Q4
  const float m = c1.z;
  const float qxx = c2.x, qxy = c2.y, qxz = c2.z, qyy = c2.w, qyz = c3.x, qzz = c3.y;
  const float qxxx = c3.z, qxyy = c3.w, qxxy = c4.x, qyyy = c4.y, qxxz = c4.z, qyyz = c4.w,
              qxyz = c5.x, q3a = c5.y, q3b = c5.z;
  const float qxxxx = c5.w, qxyyy = c6.x, qxxxy = c6.y, qyyyy = c6.z, qxxxz = c6.w, qyyyz = c7.x,
              qxxyy = c7.y, qxxyz = c7.z, qxyyz = c7.w, q4a = c0.w, q4b = c1.x, q4c = c1.y;
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
  // V2 = Q2 . r
  const float v2x = qxx * x + qxy * y + qxz * z;
  const float v2y = qxy * x + qyy * y + qyz * z;
  const float v2z = qxz * x + qyz * y + qzz * z;
  // V3 = Q3 . r r
  const float v3x = qxxx * fxx + qxyy * fyy + qxxy * fxy + qxxz * fxz + qxyz * fyz;
  const float v3y = qxyy * fxy + qxxy * fxx + qyyy * fyy + qyyz * fyz + qxyz * fxz;
  const float v3z = qxxz * fxx + qyyz * fyy + qxyz * fxy + q3a * fxz + q3b * fyz;
  // V4 = Q4 . r r r
  const float v4x = qxxxx * fxxx + qxxyy * fxyy + qxxxy * fxxy + qxyyy * fyyy + qxxxz * fxxz +
                    qxyyz * fyyz + qxxyz * fxyz;
  const float v4y = qyyyy * fyyy + qxxyy * fxxy + qxxxy * fxxx + qxyyy * fxyy + qxyyz * fxyz +
                    qxxyz * fxxz + qyyyz * fyyz;
  const float v4z = qxxxz * fxxx + qxyyz * fxyy + qxxyz * fxxy + qyyyz * fyyy + q4a * fxxz +
                    q4b * fyyz + q4c * fxyz;
  const float p2 = v2x * x + v2y * y + v2z * z;
  const float p3 = v3x * x + v3y * y + v3z * z;
  const float p4 = v4x * x + v4y * y + v4z * z;
  const float d2 = dir * dir, d4 = d2 * d2;
  const float pin = m + d4 * (0.5f * p2 + d2 * ((1.0f / 3.0f) * p3 + (0.25f * d2) * p4));
  const float S = m + d4 * (2.5f * p2 + d2 * ((7.0f / 3.0f) * p3 + (2.25f * d2) * p4));
  const float d3 = d2 * dir, d5 = d3 * d2, d3S = d3 * S;
  A.pot = __builtin_fmaf(-dir, pin, A.pot);
  A.ax = __builtin_fmaf(d5, v2x + d2 * (v3x + d2 * v4x), A.ax);
  A.ay = __builtin_fmaf(d5, v2y + d2 * (v3y + d2 * v4y), A.ay);
  A.az = __builtin_fmaf(d5, v2z + d2 * (v3z + d2 * v4z), A.az);
  A.ax = __builtin_fmaf(-d3S, x, A.ax);
  A.ay = __builtin_fmaf(-d3S, y, A.ay);
  A.az = __builtin_fmaf(-d3S, z, A.az);
  A.idt = max_nonneg(A.idt, (mt + m) * d3);
}

template <int STRIDE = 1>
__device__ __forceinline__ void pc_interact(float x, float y, float z, const float4* rec,
                                            float mt, Acc32& A) {
  pc_eval(x, y, z, load_rec<STRIDE>(rec), mt, A);
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

// List entries are read once per evaluation: streaming (non-temporal) loads,
// so that they do not displace the cell records and particles that
// neighbouring buckets reuse from L2.
__device__ __forceinline__ ILCell load_stream(const ILCell* p) {
  const long long v = __builtin_nontemporal_load(reinterpret_cast<const long long*>(p));
  ILCell c;
  c.index = (int)(unsigned)(v & 0xffffffffll);
  c.offsetID = (int)(unsigned)((unsigned long long)v >> 32);
  return c;
}

__device__ __forceinline__ ILPart load_stream(const ILPart* p) {
  ILPart r;
  r.index = __builtin_nontemporal_load(&p->index);
  r.off = __builtin_nontemporal_load(&p->off);
  r.num = __builtin_nontemporal_load(&p->num);
  return r;
}

// STREAM false: plain loads, for lists that are re-read or were just written
// (a streaming load of a dirty line sends it back to memory early).
template <bool STREAM, typename T>
__device__ __forceinline__ T load_entry(const T* p) {
  if constexpr (STREAM) return load_stream(p);
  else return *p;
}

// Wire record of list entry e, for every lane (no branch: a load under a
// divergent branch makes the compiler wait for it at the branch's end, which
// would serialize the prefetch): the caller clamps e into the list.
template <bool STREAM>
__device__ __forceinline__ void load_cell_all(const ChangaGravityLaunch& a, const ILCell* cells,
                                              int e, WireCell& w) {
  const ILCell c = load_entry<STREAM>(cells + e);
  w.idx = c.index;
  w.oid = c.offsetID;
  const float* src = reinterpret_cast<const float*>(a.moments) + (size_t)c.index * kWireStride;
#pragma unroll
  for (int k = 0; k < 27; ++k) w.v[k] = src[wire_float(k)];
}

template <bool PRECISE, bool STREAM = true>
__device__ __forceinline__ void load_cell(const ChangaGravityLaunch& a, const ILCell* cells, int e,
                                          int end, WireCell& w) {
  if (e < end) {
    const ILCell c = load_entry<STREAM>(cells + e);
    w.idx = c.index;
    w.oid = c.offsetID;
    const float* src = reinterpret_cast<const float*>(a.moments) + (size_t)c.index * kWireStride;
#pragma unroll
    for (int k = 0; k < 27; ++k) w.v[k] = src[wire_float(k)];
  }
}

// wire record -> staged record (in registers), then to LDS (float4 k at
// dst[k * STRIDE])
template <bool PRECISE, int STRIDE = 1>
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
    dst[k * STRIDE] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

// Replica code of the central image (x, y, z offsets 0) in offsetID bits
// 22-30, and whether any of a tile's cells (mine: this lane's is one) comes
// from another image: a wavefront-uniform answer, so that tiles of central
// cells skip the offset arithmetic (adding a zero offset changes no bit).
constexpr int kCentralCode = 3 | (3 << 3) | (3 << 6);
__device__ __forceinline__ bool tile_shifted(bool mine, int oid) {
  return __ballot(mine && ((oid >> 22) & 0x1ff) != kCentralCode) != 0ull;
}

// Replica periods per axis (fperiodY / fperiodZ 0: the cube's fperiod).
struct Period {
  float x, y, z;
};
__device__ __forceinline__ Period period_of(const ChangaGravityLaunch& a) {
  return {a.fperiod, a.fperiodY > 0.0f ? a.fperiodY : a.fperiod,
          a.fperiodZ > 0.0f ? a.fperiodZ : a.fperiod};
}

// Replica-shifted centre of a cell relative to the bucket centre (precision
// form: FP64 subtraction, then FP32) or absolute (plain form). shifted
// (wavefront-uniform) false: every cell of the tile is in the central image.
template <bool PRECISE>
__device__ __forceinline__ void cell_centre(double mx, double my, double mz, float fx, float fy,
                                            float fz, int oid, const Period& L, bool shifted,
                                            double cx0, double cy0, double cz0, float& x, float& y,
                                            float& z) {
  // every operation rounded on its own (no contraction into FMAs), so that
  // both staging paths and the records give the same bits
#pragma clang fp contract(off)
  if (PRECISE) {
    if (shifted) {
      x = (float)(mx + (double)off_x(oid) * (double)L.x - cx0);
      y = (float)(my + (double)off_y(oid) * (double)L.y - cy0);
      z = (float)(mz + (double)off_z(oid) * (double)L.z - cz0);
    } else {
      x = (float)(mx - cx0);
      y = (float)(my - cy0);
      z = (float)(mz - cz0);
    }
  } else if (shifted) {
    x = (float)off_x(oid) * L.x + fx;
    y = (float)off_y(oid) * L.y + fy;
    z = (float)off_z(oid) * L.z + fz;
  } else {
    x = fx;
    y = fy;
    z = fz;
  }
}

// wire record -> forward record (in registers), then to LDS (float4 k at
// dst[k * STRIDE])
template <bool PRECISE, int STRIDE>
__device__ __forceinline__ void stage_cell_fwd(const ChangaGravityLaunch& a, const WireCell& w,
                                               const Period& L, bool shifted, double cx0,
                                               double cy0, double cz0, float4* dst) {
#pragma clang fp contract(off)
  float r[kFwdRec];
  const int oid = w.oid;
  double mx = 0.0, my = 0.0, mz = 0.0;
  if (PRECISE) {
    const size_t ci = (size_t)w.idx * 3;
    mx = a.cm64[ci];
    my = a.cm64[ci + 1];
    mz = a.cm64[ci + 2];
  }
  cell_centre<PRECISE>(mx, my, mz, w.v[3], w.v[4], w.v[5], oid, L, shifted, cx0, cy0, cz0,
                       r[F_CX], r[F_CY], r[F_CZ]);
  const float u = w.v[0], u2 = u * u;
  const float s2 = 3.0f * u2, s3 = 7.5f * u2 * u, s4 = 17.5f * u2 * u2;
#pragma unroll
  for (int k = 0; k < 5; ++k) r[F_Q2 + k] = s2 * w.v[6 + k];
  r[F_Q2 + 5] = -(r[F_Q2 + 0] + r[F_Q2 + 3]);
#pragma unroll
  for (int k = 0; k < 7; ++k) r[F_Q3 + k] = s3 * w.v[11 + k];
  r[F_Q3 + 7] = -(r[F_Q3 + 0] + r[F_Q3 + 1]);
  r[F_Q3 + 8] = -(r[F_Q3 + 2] + r[F_Q3 + 3]);
#pragma unroll
  for (int k = 0; k < 9; ++k) r[F_Q4 + k] = s4 * w.v[18 + k];
  r[F_Q4A] = -(r[F_Q4 + 0] + r[F_Q4 + 6]);
  r[F_Q4B] = -(r[F_Q4 + 3] + r[F_Q4 + 6]);
  r[F_Q4C] = -(r[F_Q4 + 2] + r[F_Q4 + 1]);
  r[F_M] = w.v[2];
  r[F_SOFT] = (oid & CHANGA_ILCELL_SOFTENED) ? w.v[1] : -1.0f;
#pragma unroll
  for (int k = 0; k < kFwdRec / 4; ++k)
    dst[k * STRIDE] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

// Cell records (changaCellRecords): one 128-byte line per cell. A tile of
// TILE cells is fetched by TILE / 8 loads per lane: load j gives lane l the
// float4 l % 8 of the tile's cell 8 j + l / 8, so eight consecutive lanes
// read one record's 128 contiguous bytes and every load covers eight whole
// cache lines. In the component planes (stride PS = TILE + 1 float4) the
// eight float4 of one cell then land in eight different bank groups.
// (native vectors: HIP's float4 is a struct whose copies become memcpy calls,
// which keep the tile in scratch memory instead of registers)
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int TILE>
struct RecTile {
  f32x4 v[TILE / 8];
};

// cidx: this lane's cell index (its entry of the tile), valid for lanes < nc;
// slots past nc fetch the tile's last cell again (no branch around a load)
template <int J, int TILE>
__device__ __forceinline__ void load_rec_tile_j(const float4* recs, int cidx, int nc, int k,
                                                int c0, RecTile<TILE>& f) {
  if constexpr (J < TILE / 8) {
    const int idx = __shfl(cidx, min(8 * J + c0, nc - 1), kWave);
    // a 32-bit byte offset from the (wavefront-uniform) base: the records
    // of at most 2^25 cells span less than 4 GiB
    const unsigned off = (unsigned)idx * (unsigned)CHANGA_CELL_RECORD_BYTES + (unsigned)k * 16u;
    f.v[J] = *reinterpret_cast<const f32x4*>(reinterpret_cast<const char*>(recs) + off);
    load_rec_tile_j<J + 1, TILE>(recs, cidx, nc, k, c0, f);
  }
}

// (the tile's slots are indexed at compile time, so that the tile stays in
// registers)
template <int TILE>
__device__ __forceinline__ void load_rec_tile(const float4* recs, int cidx, int nc, int lane,
                                              RecTile<TILE>& f) {
  load_rec_tile_j<0, TILE>(recs, cidx, nc, lane & 7, lane >> 3, f);
}

// Write a fetched tile into the component planes (float4 k of cell c at
// rec4[k * PS + c]), then form each cell's centre (replica offset, bucket
// centre), its softened flag and the three Q4 combinations: lane c finishes
// cell c. Ends with a barrier.
template <bool PRECISE, int TILE, int PS>
__device__ __forceinline__ void stage_rec_tile(const RecTile<TILE>& f, int nc, int lane, int oid,
                                               const Period& L, bool shifted, double cx0,
                                               double cy0, double cz0, float4* rec4) {
#pragma clang fp contract(off)
  f32x4* dst = reinterpret_cast<f32x4*>(rec4) + (lane & 7) * PS + (lane >> 3);
  [&]<int... J>(std::integer_sequence<int, J...>) {
    ((dst[8 * J] = f.v[J]), ...);
  }(std::make_integer_sequence<int, TILE / 8>());
  __syncthreads();
  if (lane < nc) {
    float4* r = rec4 + lane;
    const float4 p0 = r[0], p1 = r[PS], p5 = r[5 * PS], p6 = r[6 * PS], p7 = r[7 * PS];
    const double mx = __hiloint2double(__float_as_int(p0.y), __float_as_int(p0.x));
    const double my = __hiloint2double(__float_as_int(p0.w), __float_as_int(p0.z));
    const double mz = __hiloint2double(__float_as_int(p1.y), __float_as_int(p1.x));
    float x, y, z;
    cell_centre<PRECISE>(mx, my, mz, (float)mx, (float)my, (float)mz, oid, L, shifted, cx0, cy0,
                         cz0, x, y, z);
    // F_Q4 + 0 = p5.w (xxxx), +1 = p6.x (xyyy), +2 = p6.y (xxxy), +3 = p6.z (yyyy),
    // +6 = p7.y (xxyy)
    const float q4a = -(p5.w + p7.y), q4b = -(p6.z + p7.y), q4c = -(p6.y + p6.x);
    r[0] = make_float4(x, y, z, q4a);
    r[PS] = make_float4(q4b, q4c, p1.z, (oid & CHANGA_ILCELL_SOFTENED) ? p1.w : -1.0f);
  }
  __syncthreads();
}


// ---------------------------------------------------------------------------
// Per-chunk loops shared by the per-bucket kernels and the persistent
// batched kernel. A chunk is up to 64 targets of one bucket; `Lanes` is the
// slot layout of the wavefront over them.
// ---------------------------------------------------------------------------
// Lane layout of a chunk of nt <= 64 targets with K targets per lane: the
// lanes form nslot = 64 / ng slots of ng = ceil(nt / K) lanes each; lane g of
// a slot holds the targets g, g + ng, ..., g + (K - 1) ng that are < nt, and
// every slot takes every nslot-th source. K > 1 packs more slots into the
// wavefront when nt does not divide 64 well (nt = 17: 3 slots, 51 lanes busy
// with K = 1; 7 slots of 9 lanes, 119 of 128 target places, with K = 2).
struct Lanes {
  int lane, nt, nslot, slot;
  bool live;
  int ng, g;
};

template <int K = 1>
__device__ __forceinline__ Lanes lanes_for(int lane, int nt) {
  Lanes l;
  l.lane = lane;
  l.nt = nt;
  l.ng = (nt + K - 1) / K;
  l.nslot = kWave / l.ng;
  l.slot = lane / l.ng;
  l.g = lane - l.slot * l.ng;
  l.live = l.slot < l.nslot;
  return l;
}

// Targets per lane (1, 2 or 3) that keep the most target places busy,
// nslot(K) nt / (64 K), the smaller K on ties.
__device__ __forceinline__ int targets_per_lane(int nt) {
  int best = 1;
  int bnum = (kWave / nt) * nt, bden = 1;  // utilisation bnum / (64 bden)
  for (int k = 2; k <= 3; ++k) {
    const int ng = (nt + k - 1) / k;
    const int num = (kWave / ng) * nt;
    if (num * bden > bnum * k) {
      best = k;
      bnum = num;
      bden = k;
    }
  }
  return best;
}

template <int K>
struct KTag {
  static constexpr int value = K;
};

// Calls body(KTag<K>()) with K = targets_per_lane(nt) (nt uniform across the
// wavefront, so one of the three instantiations runs).
template <typename Body>
__device__ __forceinline__ void with_targets_per_lane(int nt, Body&& body) {
  switch (targets_per_lane(nt)) {
    case 1:
      body(KTag<1>());
      break;
    case 2:
      body(KTag<2>());
      break;
    default:
      body(KTag<3>());
      break;
  }
}

struct Target {
  float x, y, z, m, soft;
};

struct Acc64 {
  double ax, ay, az, pot;
  float idt;
};

// FP64 reference point of a bucket in the precision form: its first particle.
template <bool PRECISE>
__device__ __forceinline__ void bucket_centre(const double* tpos64, int bstart, int n, double& cx0,
                                              double& cy0, double& cz0) {
  cx0 = cy0 = cz0 = 0.0;
  if (PRECISE && n > 0) {
    cx0 = tpos64[3 * (size_t)bstart];
    cy0 = tpos64[3 * (size_t)bstart + 1];
    cz0 = tpos64[3 * (size_t)bstart + 2];
  }
}

template <bool PRECISE>
__device__ __forceinline__ Target load_target(const ChangaGravityLaunch& a, int tp, double cx0,
                                              double cy0, double cz0) {
  Target t;
  CompactPartData p = a.targets[tp];
  t.m = p.mass;
  t.soft = p.soft;
  if (PRECISE) {
    t.x = (float)(a.tpos64[3 * (size_t)tp] - cx0);
    t.y = (float)(a.tpos64[3 * (size_t)tp + 1] - cy0);
    t.z = (float)(a.tpos64[3 * (size_t)tp + 2] - cz0);
  } else {
    t.x = p.position.x;
    t.y = p.position.y;
    t.z = p.position.z;
  }
  return t;
}

// The K targets of a lane: target g + j ng of the chunk (a place past nt
// holds target g again; the reduction ignores it).
template <bool PRECISE, int K>
__device__ __forceinline__ void load_targets(const ChangaGravityLaunch& a, int tbase,
                                             const Lanes& l, double cx0, double cy0, double cz0,
                                             Target (&t)[K]) {
  const int g = l.live ? l.g : 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int tj = g + j * l.ng;
    t[j] = load_target<PRECISE>(a, tbase + (tj < l.nt ? tj : g), cx0, cy0, cz0);
  }
}

// Doubles of LDS the slot reduction of K targets per lane needs (the floats
// of dtGrav included).
template <int K>
constexpr size_t reduce_bytes() {
  return (size_t)K * kWave * (4 * sizeof(double) + sizeof(float));
}

// Slot reduction in a fixed order and the += into the result array (+= for
// a and phi, max for dtGrav): every lane stores the sums of its places, the
// slots are added pairwise (slot s += slot s + h, h = ceil(n / 2), while
// n > 1 slots are left: log2(nslot) rounds, all lanes at once), then lane
// i < nt adds target i's total, now in slot 0. lds: reduce_bytes<K>().
template <int K>
__device__ __forceinline__ void reduce_store(void* lds, const Lanes& l, const Acc64 (&acc)[K],
                                             VariablePartData* out) {
  constexpr int W = K * kWave;
  double* red = static_cast<double*>(lds);
  float* redi = reinterpret_cast<float*>(red + 4 * W);
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int i = j * kWave + l.lane;
    const bool v = l.live && l.g + j * l.ng < l.nt;
    red[i] = v ? acc[j].ax : 0.0;
    red[W + i] = v ? acc[j].ay : 0.0;
    red[2 * W + i] = v ? acc[j].az : 0.0;
    red[3 * W + i] = v ? acc[j].pot : 0.0;
    redi[i] = v ? acc[j].idt : 0.0f;
  }
  __syncthreads();
  for (int n = l.nslot; n > 1;) {
    const int h = (n + 1) >> 1;
    if (l.slot < n - h) {
      const int d = h * l.ng;  // the partner lane: same g, slot + h
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int i = j * kWave + l.lane;
        red[i] += red[i + d];
        red[W + i] += red[W + i + d];
        red[2 * W + i] += red[2 * W + i + d];
        red[3 * W + i] += red[3 * W + i + d];
        redi[i] = fmaxf(redi[i], redi[i + d]);
      }
    }
    n = h;
    __syncthreads();
  }
  if (l.lane < l.nt) {
    const int j = l.lane / l.ng, i = j * kWave + (l.lane - j * l.ng);
    VariablePartData v = out[l.lane];
    v.a.x += (float)red[i];
    v.a.y += (float)red[W + i];
    v.a.z += (float)red[2 * W + i];
    v.potential += (float)red[3 * W + i];
    v.dtGrav = fmaxf(v.dtGrav, redi[i]);
    out[l.lane] = v;
  }
  __syncthreads();
}

// Cells [start, end) of a.list onto the chunk's targets. Tiles of TILE <= 64
// cells are staged into rec4 ((TILE + 1) * kFwdRec floats of LDS) as component
// planes: float4 k of cell c at rec4[k * (TILE + 1) + c], so lanes reading or
// writing different cells touch consecutive 16-byte words (no bank
// conflicts however many distinct cells a read instruction touches: one per
// lane when a bucket has a single target). The next tile's data is in flight
// while this one is evaluated. REC false: lane c fetches cell c's wire
// record and stages it; REC true: the tile's 128-byte cell records
// (a.cellRecords) are fetched as whole lines, eight lanes per record
// (load_rec_tile, stage_rec_tile). A tile without softened cells (ballot)
// runs the hexadecapole loop with no per-cell test.
template <bool PRECISE, int TILE, bool STREAM = true, int K = 1, bool REC = false>
__device__ __forceinline__ void pc_cells(const ChangaGravityLaunch& a, int start, int end,
                                         double cx0, double cy0, double cz0,
                                         const Target (&t)[K], const Lanes& l, float4* rec4,
                                         Acc64 (&acc)[K]) {
  static_assert(TILE <= kWave && TILE % 8 == 0, "one staged cell per lane");
  constexpr int PS = TILE + 1;  // plane stride (float4)
  const Period L = period_of(a);
  const ILCell* cells = static_cast<const ILCell*>(a.list);
  // cells per tile: the most slot rounds that fit TILE, so that the last
  // round of a tile keeps every slot busy (with 9 slots, 63 cells in 7 full
  // rounds rather than 64 in 8, the last of them one cell); a tile of fewer
  // than two rounds (sparse buckets: many slots) keeps TILE cells, its
  // per-tile work then weighing more than the partial round
  const int tc = 2 * l.nslot <= TILE ? l.nslot * (TILE / l.nslot) : TILE;
  // list entry of this lane in the tile at base (clamped into the list)
  auto entry_at = [&](int base) { return min(base + min(l.lane, tc - 1), end - 1); };
  // (end - 1 >= start: the loop below runs only for a non-empty list)
  // REC false: the next tile's wire records. REC true: the next tile's
  // record lines, its offsetIDs, and the entries of the tile after it (read
  // a tile ahead, so that the index shuffle of a fetch never waits on them).
  WireCell w;
  RecTile<TILE> rt;
  ILCell enext;
  int oid_next = 0;
  const float4* recs = static_cast<const float4*>(a.cellRecords);
  // every fetch is unconditional (past the list's end it fetches the last
  // cell again): a load under any branch makes the compiler wait for it where
  // the branch joins, which would put the fetch's latency before the
  // evaluation it is meant to overlap
  auto fetch = [&](int base) {
    if constexpr (REC) {
      load_rec_tile<TILE>(recs, enext.index, max(1, min(tc, end - base)), l.lane, rt);
      oid_next = enext.offsetID;
      enext = load_entry<STREAM>(cells + entry_at(base + tc));
    } else {
      load_cell_all<STREAM>(a, cells, entry_at(base), w);
    }
  };
  if (start >= end) return;
  if constexpr (REC) enext = load_entry<STREAM>(cells + entry_at(start));
  fetch(start);
  for (int base = start; base < end; base += tc) {
    const int nc = min(tc, end - base);
    const bool mine = l.lane < nc;
    const int oid = REC ? oid_next : w.oid;
    const bool soft_tile = __ballot(mine && (oid & CHANGA_ILCELL_SOFTENED)) != 0ull;
    const bool shifted = tile_shifted(mine, oid);
    if constexpr (REC) {
      stage_rec_tile<PRECISE, TILE, PS>(rt, nc, l.lane, oid, L, shifted, cx0, cy0, cz0, rec4);
    } else {
      if (mine) stage_cell_fwd<PRECISE, PS>(a, w, L, shifted, cx0, cy0, cz0, rec4 + l.lane);
      __syncthreads();
    }
    // issued after the barrier, whose fence would otherwise wait for them
    fetch(base + tc);
    if (l.live) {
      Acc32 A[K];
#pragma unroll
      for (int j = 0; j < K; ++j) A[j] = {0.f, 0.f, 0.f, 0.f, acc[j].idt};
      // cell slot + i nslot: full rounds over every slot, then one partial
      const int full = nc / l.nslot;
      const float4* r4 = rec4 + l.slot;
      auto hex = [&](const float4* r) {
        const PcRec R = load_rec<PS>(r);
#pragma unroll
        for (int j = 0; j < K; ++j) {
          pc_eval(t[j].x - R.c0.x, t[j].y - R.c0.y, t[j].z - R.c0.z, R, t[j].m, A[j]);
          // one target's evaluation at a time: interleaving them would need
          // K times the temporaries
          if (j + 1 < K) __builtin_amdgcn_sched_barrier(0);
        }
      };
      if (!soft_tile) {
        for (int i = 0; i < full; ++i, r4 += l.nslot) hex(r4);
        if (l.slot < nc - full * l.nslot) hex(r4);
      } else {
        auto cell = [&](const float4* r) {
          const float4 p1 = r[PS];
          if (p1.w >= 0.0f) {  // softened cell: spline monopole at the cm
            const float4 h = r[0];
#pragma unroll
            for (int j = 0; j < K; ++j)
              pp_interact(h.x - t[j].x, h.y - t[j].y, h.z - t[j].z, p1.z, p1.w, t[j].m,
                          t[j].soft, A[j]);
          } else {
            hex(r);
          }
        };
        for (int i = 0; i < full; ++i, r4 += l.nslot) cell(r4);
        if (l.slot < nc - full * l.nslot) cell(r4);
      }
#pragma unroll
      for (int j = 0; j < K; ++j) {
        acc[j].ax += A[j].ax;
        acc[j].ay += A[j].ay;
        acc[j].az += A[j].az;
        acc[j].pot += A[j].pot;
        acc[j].idt = A[j].idt;
      }
    }
    __syncthreads();
  }
}

// LDS of the particle-particle loop for TILE staged sources.
template <int TILE>
struct PpLds {
  float4 src4[TILE];
  float srcs[TILE];
  int2 map[TILE];  // ILPart runs expanded: (source particle, offsetID) per tile slot
};

// Tiles of the sources [start, end) of a.list staged into LDS (s.src4: x, y,
// z relative to the bucket centre with the replica offset applied, mass;
// s.srcs: softening); consume(ns) runs after each tile is staged. KIND 0: one
// ILCell per source particle (the list ChaNGa's GenericList<ILPart>::serialize
// produces, Compute.cpp:1118-1218); KIND 1: ILPart runs expanded here: a
// wavefront prefix sum places each run, and each lane writes its run's slots
// of the tile into LDS.
template <bool PRECISE, int KIND, int TILE, bool STREAM = true, typename Consume>
__device__ __forceinline__ void pp_tiles(const ChangaGravityLaunch& a, int start, int end,
                                         double cx0, double cy0, double cz0, int lane,
                                         PpLds<TILE>& s, Consume&& consume) {
  const Period L = period_of(a);

  // stage source particle p (replica oid) into tile slot i
  auto stage = [&](int i, int p, int oid) {
    CompactPartData q = a.sources[p];
    float4 v;
    if (PRECISE) {
      v.x = (float)(a.spos64[3 * (size_t)p] + (double)off_x(oid) * (double)L.x - cx0);
      v.y = (float)(a.spos64[3 * (size_t)p + 1] + (double)off_y(oid) * (double)L.y - cy0);
      v.z = (float)(a.spos64[3 * (size_t)p + 2] + (double)off_z(oid) * (double)L.z - cz0);
    } else {
      v.x = (float)off_x(oid) * L.x + q.position.x;
      v.y = (float)off_y(oid) * L.y + q.position.y;
      v.z = (float)off_z(oid) * L.z + q.position.z;
    }
    v.w = q.mass;
    s.src4[i] = v;
    s.srcs[i] = q.soft;
  };

  if (KIND == 0) {
    const ILCell* lst = static_cast<const ILCell*>(a.list);
    for (int base = start; base < end; base += TILE) {
      const int ns = min(TILE, end - base);
      for (int i = lane; i < ns; i += kWave) {
        const ILCell c = load_entry<STREAM>(lst + base + i);
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
      int num = 0, first = 0, oid = 0;
      if (lane < nr) {
        const ILPart p = load_entry<STREAM>(lst + rb + lane);
        first = p.index;
        oid = p.off;
        num = p.num;
      }
      // inclusive prefix sum of run lengths across the wavefront: this lane's
      // run covers [incl - num, incl) of the expanded sequence
      int incl = num;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        int y = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += y;
      }
      const int total = __shfl(incl, kWave - 1, kWave);
      const int begin = incl - num;
      for (int sb = 0; sb < total; sb += TILE) {
        const int ns = min(TILE, total - sb);
        // each lane expands the part of its run that falls in this tile
        const int j1 = min(incl, sb + ns);
        for (int j = max(begin, sb); j < j1; ++j)
          s.map[j - sb] = make_int2(first + (j - begin), oid);
        __syncthreads();
        for (int i = lane; i < ns; i += kWave) {
          const int2 m = s.map[i];
          stage(i, m.x, m.y);
        }
        __syncthreads();
        consume(ns);
        __syncthreads();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward particle-particle loop (partBucketForce) with far and near sources
// kept apart. A source is far from the chunk when its distance from the
// chunk's centre exceeds the chunk's radius plus both softenings (with a
// relative margin of 2e-4, far above FP32 rounding): no pair of it with a
// target of the chunk can need the spline or coincide, so its pairs run a
// Newtonian loop with no per-pair test. Near sources run the general loop.
//
// LDS of a staged tile: rec[kPpFwdCap] float4 (x, y, z, m), far sources from
// the front, near sources from the back (near source k at rec[cap - 1 - k],
// its softening at soft[k]). Far sources are consumed slot-major: slot s
// takes the q consecutive records [s q, s q + q) with q uniform,
// the records past the far count holding null sources (m = 0 at 1e18: b
// underflows to 0, so they add exactly nothing), so the loop runs a fixed
// trip count with immediate-offset LDS reads (pairs, then one odd record).
// ---------------------------------------------------------------------------
template <int TILE>
struct PpFwdLds {
  static constexpr int cap = TILE + kWave;  // far padding is < nslot <= kWave
  float4 rec[cap];
  float soft[TILE];
  int2 map[TILE];  // ILPart runs expanded: (source particle, offsetID) per tile slot
};

// Chunk sphere for the far test: centre (the chunk's first target; the
// origin of the staged frame in the precision form), radius + the largest
// target softening, times the margin.
struct ChunkSphere {
  float cx, cy, cz, r;
};

template <bool PRECISE, int K>
__device__ __forceinline__ ChunkSphere chunk_sphere(const Target (&t)[K], const Lanes& l) {
  ChunkSphere c;
  if (PRECISE) {
    c.cx = c.cy = c.cz = 0.0f;  // the bucket centre, first target of the bucket
  } else {
    c.cx = __shfl(t[0].x, 0, kWave);  // lane 0 holds the chunk's first target
    c.cy = __shfl(t[0].y, 0, kWave);
    c.cz = __shfl(t[0].z, 0, kWave);
  }
  float d2 = 0.0f, sm = 0.0f;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const float dx = t[j].x - c.cx, dy = t[j].y - c.cy, dz = t[j].z - c.cz;
    // a place past nt holds a real target of the chunk again
    if (l.live) {
      d2 = fmaxf(d2, dx * dx + dy * dy + dz * dz);
      sm = fmaxf(sm, t[j].soft);
    }
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    d2 = fmaxf(d2, __shfl_xor(d2, d, kWave));
    sm = fmaxf(sm, __shfl_xor(sm, d, kWave));
  }
  c.r = (sqrtf(d2) + sm) * 1.0002f;
  return c;
}

// Newtonian pair at r = src - tgt (the far loop: no softening, r != 0).
__device__ __forceinline__ void pp_far(const float4 v, const Target& t, Acc32& A) {
  const float rx = v.x - t.x, ry = v.y - t.y, rz = v.z - t.z;
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float ir = rsq(r2);
  const float b = ir * ir * ir;
  const float bm = b * v.w;
  A.ax += rx * bm;
  A.ay += ry * bm;
  A.az += rz * bm;
  A.pot -= v.w * ir;
  A.idt = max_nonneg(A.idt, __builtin_fmaf(t.m, b, bm));
}

// Calls tile(ns, src) for each tile of the sources [start, end) of a.list,
// where src(i) gives (source particle, offsetID) of tile slot i < ns. KIND
// 0: one ILCell per source particle (the list ChaNGa's
// GenericList<ILPart>::serialize produces, Compute.cpp:1118-1218); KIND 1:
// ILPart runs expanded into map: a wavefront prefix sum places each run,
// and each lane writes its run's slots of the tile.
template <int KIND, int TILE, bool STREAM, typename Tile>
__device__ __forceinline__ void pp_source_tiles(const ChangaGravityLaunch& a, int start, int end,
                                                int lane, int2* map, Tile&& tile) {
  if (KIND == 0) {
    const ILCell* lst = static_cast<const ILCell*>(a.list);
    for (int base = start; base < end; base += TILE) {
      tile(min(TILE, end - base), [&](int i) {
        const ILCell c = load_entry<STREAM>(lst + base + i);
        return make_int2(c.index, c.offsetID);
      });
    }
  } else {
    const ILPart* lst = static_cast<const ILPart*>(a.list);
    for (int rb = start; rb < end; rb += kWave) {
      const int nr = min(kWave, end - rb);
      int num = 0, first = 0, oid = 0;
      if (lane < nr) {
        const ILPart p = load_entry<STREAM>(lst + rb + lane);
        first = p.index;
        oid = p.off;
        num = p.num;
      }
      int incl = num;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        int y = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += y;
      }
      const int total = __shfl(incl, kWave - 1, kWave);
      const int begin = incl - num;
      for (int sb = 0; sb < total; sb += TILE) {
        const int ns = min(TILE, total - sb);
        const int j1 = min(incl, sb + ns);
        for (int j = max(begin, sb); j < j1; ++j) map[j - sb] = make_int2(first + (j - begin), oid);
        __syncthreads();
        tile(ns, [&](int i) { return map[i]; });
      }
    }
  }
}

// Sources [start, end) of a.list onto the chunk's targets (partBucketForce).
template <bool PRECISE, int KIND, int TILE, bool STREAM = true, int K = 1>
__device__ __forceinline__ void pp_sources(const ChangaGravityLaunch& a, int start, int end,
                                           double cx0, double cy0, double cz0,
                                           const Target (&t)[K], const Lanes& l,
                                           PpFwdLds<TILE>& s, Acc64 (&acc)[K]) {
  constexpr int cap = PpFwdLds<TILE>::cap;
  const Period L = period_of(a);
  const int lane = l.lane;
  const ChunkSphere cs = chunk_sphere<PRECISE, K>(t, l);
  const unsigned long long below = (1ull << lane) - 1ull;
  pp_source_tiles<KIND, TILE, STREAM>(a, start, end, lane, s.map, [&](int ns, auto&& src) {
    // stage: classify each source, far ones to the front, near ones to the back
    int nfar = 0, nnear = 0;
    for (int i0 = 0; i0 < ns; i0 += kWave) {
      const int i = i0 + lane;
      const bool valid = i < ns;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      float ss = 0.0f;
      bool far = false;
      int2 e = make_int2(0, kCentralCode << 22);
      if (valid) e = src(i);
      // sources of the central image only: no offset arithmetic (adding a
      // zero offset changes no bit)
      const bool shifted = tile_shifted(valid, e.y);
      if (valid) {
        const int p = e.x, oid = e.y;
        const CompactPartData q = a.sources[p];
        if (PRECISE) {
          if (shifted) {
            v.x = (float)(a.spos64[3 * (size_t)p] + (double)off_x(oid) * (double)L.x - cx0);
            v.y = (float)(a.spos64[3 * (size_t)p + 1] + (double)off_y(oid) * (double)L.y - cy0);
            v.z = (float)(a.spos64[3 * (size_t)p + 2] + (double)off_z(oid) * (double)L.z - cz0);
          } else {
            v.x = (float)(a.spos64[3 * (size_t)p] - cx0);
            v.y = (float)(a.spos64[3 * (size_t)p + 1] - cy0);
            v.z = (float)(a.spos64[3 * (size_t)p + 2] - cz0);
          }
        } else if (shifted) {
          v.x = (float)off_x(oid) * L.x + q.position.x;
          v.y = (float)off_y(oid) * L.y + q.position.y;
          v.z = (float)off_z(oid) * L.z + q.position.z;
        } else {
          v.x = q.position.x;
          v.y = q.position.y;
          v.z = q.position.z;
        }
        v.w = q.mass;
        ss = q.soft;
        const float dx = v.x - cs.cx, dy = v.y - cs.cy, dz = v.z - cs.cz;
        const float thr = __builtin_fmaf(ss, 1.0002f, cs.r);
        far = dx * dx + dy * dy + dz * dz > fmaxf(thr * thr, 1e-30f);
      }
      const unsigned long long fm = __ballot(far), nm = __ballot(valid && !far);
      if (far) {
        s.rec[nfar + __popcll(fm & below)] = v;
      } else if (valid) {
        const int k = nnear + __popcll(nm & below);
        s.rec[cap - 1 - k] = v;
        s.soft[k] = ss;
      }
      nfar += __popcll(fm);
      nnear += __popcll(nm);
    }
    // slot-major far layout: q records per slot, padded with null sources
    const int q = (nfar + l.nslot - 1) / l.nslot;
    for (int k = nfar + lane; k < l.nslot * q; k += kWave)
      s.rec[k] = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
    __syncthreads();
    if (l.live) {
      Acc32 A[K];
#pragma unroll
      for (int j = 0; j < K; ++j) A[j] = {0.f, 0.f, 0.f, 0.f, acc[j].idt};
      const float4* r = s.rec + l.slot * q;
      for (int i = 0; i + 1 < q; i += 2, r += 2) {
        const float4 r0 = r[0], r1 = r[1];
#pragma unroll
        for (int j = 0; j < K; ++j) pp_far(r0, t[j], A[j]);
        if (K > 1) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < K; ++j) pp_far(r1, t[j], A[j]);
      }
      if (q & 1) {
        const float4 r0 = r[0];
#pragma unroll
        for (int j = 0; j < K; ++j) pp_far(r0, t[j], A[j]);
      }
      for (int k = l.slot; k < nnear; k += l.nslot) {
        const float4 v = s.rec[cap - 1 - k];
        const float sk = s.soft[k];
#pragma unroll
        for (int j = 0; j < K; ++j)
          pp_interact(v.x - t[j].x, v.y - t[j].y, v.z - t[j].z, v.w, sk, t[j].m, t[j].soft, A[j]);
      }
#pragma unroll
      for (int j = 0; j < K; ++j) {
        acc[j].ax += A[j].ax;
        acc[j].ay += A[j].ay;
        acc[j].az += A[j].az;
        acc[j].pot += A[j].pot;
        acc[j].idt = A[j].idt;
      }
    }
    __syncthreads();
  });
}

}  // namespace cgpu
