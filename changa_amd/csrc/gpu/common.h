// Device-side helpers shared by the gravity, Ewald and moments kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "changa_gpu.h"

namespace cgpu {

constexpr int kWave = 64;
constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

// Float offsets of a wire cell record (CudaMultipoleMoments): radius, soft,
// mass and cm at 0-5, the 21 moments (xx .. xyyz) from kWireMom on;
// kWireStride floats per node. 27 and 6 in ChaNGa's default layout, 40 and
// 19 with the local-tree-walk fields between cm and the moments.
constexpr int kWireStride = (int)(sizeof(CudaMultipoleMoments) / sizeof(float));
constexpr int kWireMom = (int)(offsetof(CudaMultipoleMoments, xx) / sizeof(float));
// float k of the 27 (0-5 header, 6-26 moments) read from a wire record
__host__ __device__ constexpr int wire_float(int k) { return k < 6 ? k : k - 6 + kWireMom; }

// Record an error (and abort unless the caller switched that off). Returns
// the non-zero code so launchers can `return fail(...)`.
int fail(const char* what, hipError_t e, const char* file, int line);

#define CGPU_CHECK(expr)                                              \
  do {                                                                \
    hipError_t _e = (expr);                                           \
    if (_e != hipSuccess) return cgpu::fail(#expr, _e, __FILE__, __LINE__); \
  } while (0)

// Replica offset of offsetID bits 22-30 (TreePiece.cpp:3631-3645;
// ParallelGravity.h:2064-2073) as integer multiples of the period.
__device__ __forceinline__ int off_x(int oid) { return ((oid >> 22) & 7) - 3; }
__device__ __forceinline__ int off_y(int oid) { return ((oid >> 25) & 7) - 3; }
__device__ __forceinline__ int off_z(int oid) { return ((oid >> 28) & 7) - 3; }

// Workgroup -> bucket such that every XCD walks one contiguous range of
// buckets: neighbouring buckets share most of their source cells, so each
// XCD's L2 keeps reusing the same moments.
__device__ __forceinline__ int xcd_bucket(int wg, int n) {
  int x = wg % kNumXcd, k = wg / kNumXcd;
  int q = n / kNumXcd, r = n % kNumXcd;
  return x * q + (x < r ? x : r) + k;
}

}  // namespace cgpu
