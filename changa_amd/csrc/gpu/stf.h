// Symmetric tensors of orders 0..4 as "complete" components C(a,b,c)
// (a x's, b y's, c z's; 35 slots indexed degree by degree, a descending,
// then b descending) and their projection onto the symmetric trace-free
// subspace. Shared by the multipole build and the adjoint.
#pragma once
#include "common.h"

namespace cgpu {

// slot of the complete component (a,b,c)
__host__ __device__ constexpr int cidx(int a, int b, int c) {
  return (a + b + c == 0 ? 0 : a + b + c == 1 ? 1 : a + b + c == 2 ? 4 : a + b + c == 3 ? 10 : 20) +
         (b + c) * (b + c + 1) / 2 + c;
}
__host__ __device__ constexpr int binom(int n, int k) {
  return k < 0 || k > n ? 0 : (k == 0 || k == n) ? 1 : binom(n - 1, k - 1) + binom(n - 1, k);
}
__host__ __device__ constexpr int pairs(int a) { return a * (a - 1) / 2; }

// Project raw symmetric components of orders 2..4 onto the STF subspace.
__device__ __forceinline__ void stf_project(double* T) {
  // order 2: T - delta tr/3
  double tr = T[cidx(2, 0, 0)] + T[cidx(0, 2, 0)] + T[cidx(0, 0, 2)];
  T[cidx(2, 0, 0)] -= tr / 3.0;
  T[cidx(0, 2, 0)] -= tr / 3.0;
  T[cidx(0, 0, 2)] -= tr / 3.0;
  // order 3: T - (1/5) sum_pairs delta V, V_i = T_ikk
  double V[3] = {T[cidx(3, 0, 0)] + T[cidx(1, 2, 0)] + T[cidx(1, 0, 2)],
                 T[cidx(2, 1, 0)] + T[cidx(0, 3, 0)] + T[cidx(0, 1, 2)],
                 T[cidx(2, 0, 1)] + T[cidx(0, 2, 1)] + T[cidx(0, 0, 3)]};
#pragma unroll
  for (int a = 0; a <= 3; ++a)
#pragma unroll
    for (int b = 0; b <= 3 - a; ++b) {
      const int c = 3 - a - b;
      double s = 0.0;
      if (a >= 2) s += pairs(a) * V[(a - 2) == 1 ? 0 : (b == 1 ? 1 : 2)];
      if (b >= 2) s += pairs(b) * V[a == 1 ? 0 : ((b - 2) == 1 ? 1 : 2)];
      if (c >= 2) s += pairs(c) * V[a == 1 ? 0 : (b == 1 ? 1 : 2)];
      T[cidx(a, b, c)] -= s / 5.0;
    }
  // order 4: T - (1/7) sum_pairs delta W + (1/35) w (delta delta), W_ij = T_ijkk
  double W[6];  // xx xy xz yy yz zz
  W[0] = T[cidx(4, 0, 0)] + T[cidx(2, 2, 0)] + T[cidx(2, 0, 2)];
  W[1] = T[cidx(3, 1, 0)] + T[cidx(1, 3, 0)] + T[cidx(1, 1, 2)];
  W[2] = T[cidx(3, 0, 1)] + T[cidx(1, 2, 1)] + T[cidx(1, 0, 3)];
  W[3] = T[cidx(2, 2, 0)] + T[cidx(0, 4, 0)] + T[cidx(0, 2, 2)];
  W[4] = T[cidx(2, 1, 1)] + T[cidx(0, 3, 1)] + T[cidx(0, 1, 3)];
  W[5] = T[cidx(2, 0, 2)] + T[cidx(0, 2, 2)] + T[cidx(0, 0, 4)];
  const double w = W[0] + W[3] + W[5];
#pragma unroll
  for (int a = 0; a <= 4; ++a)
#pragma unroll
    for (int b = 0; b <= 4 - a; ++b) {
      const int c = 4 - a - b;
      // W of the multiset left after removing an equal-index pair
      auto Wof = [&](int p, int q, int r) -> double {
        return p == 2 ? W[0] : q == 2 ? W[3] : r == 2 ? W[5] : (p == 1 && q == 1) ? W[1]
                                                              : (p == 1) ? W[2] : W[4];
      };
      double s = 0.0;
      if (a >= 2) s += pairs(a) * Wof(a - 2, b, c);
      if (b >= 2) s += pairs(b) * Wof(a, b - 2, c);
      if (c >= 2) s += pairs(c) * Wof(a, b, c - 2);
      int dd = (a == 4 || b == 4 || c == 4) ? 3 : ((a == 2) + (b == 2) + (c == 2) == 2 ? 1 : 0);
      T[cidx(a, b, c)] += -s / 7.0 + w * dd / 35.0;
    }
}

}  // namespace cgpu
