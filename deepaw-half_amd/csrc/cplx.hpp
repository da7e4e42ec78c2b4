// Complex arithmetic on HIP's float2 / double2 vector types.
#pragma once

#include <hip/hip_runtime.h>

namespace half {

template <typename V>
struct scalar_of;
template <>
struct scalar_of<double2> {
  using type = double;
};
template <>
struct scalar_of<float2> {
  using type = float;
};

template <typename V>
__host__ __device__ inline V mk(typename scalar_of<V>::type x, typename scalar_of<V>::type y) {
  V r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename V>
__host__ __device__ inline V cadd(V a, V b) {
  return mk<V>(a.x + b.x, a.y + b.y);
}
template <typename V>
__host__ __device__ inline V csub(V a, V b) {
  return mk<V>(a.x - b.x, a.y - b.y);
}
template <typename V>
__host__ __device__ inline V cmul(V a, V b) {
  return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename V>
__host__ __device__ inline V cmulc(V a, V b) {
  return mk<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename V>
__host__ __device__ inline V cscale(V a, typename scalar_of<V>::type s) {
  return mk<V>(a.x * s, a.y * s);
}
// multiply by i * s (s = +-1 for +-i)
template <typename V>
__host__ __device__ inline V cmul_i(V a, typename scalar_of<V>::type s) {
  return mk<V>(-a.y * s, a.x * s);
}

}  // namespace half
