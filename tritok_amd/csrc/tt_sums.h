// tt_sums.h -- device helper for the column-slice partial sums of the dense
// products: the sum of p[0], p[stride], ..., p[(n-1)*stride] added in that
// order (so the result is that of the plain loop, bit for bit), with eight
// loads in flight at a time instead of one.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace tt {

template <typename T>
__device__ __forceinline__ double slice_sum(const T *__restrict__ p, size_t stride, int n) {
  double s = 0.0;
  int k = 0;
  for (; k + 8 <= n; k += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(k + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (double)v[u];
  }
  for (; k < n; ++k) s += (double)p[(size_t)k * stride];
  return s;
}

}  // namespace tt
