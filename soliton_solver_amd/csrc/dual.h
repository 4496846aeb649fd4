// Forward-mode dual numbers with K tangent directions, used to differentiate a
// theory's energy density with respect to its local inputs (field values and
// first derivatives) exactly, without a hand-written adjoint per theory.
#pragma once

#include <hip/hip_runtime.h>

namespace ss {

// Keep the plain overloads visible next to the Dual ones below.
using ::cos;
using ::sin;
using ::sqrt;

template <int K>
struct Dual {
  double v;
  double d[K];

  __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  __device__ __forceinline__ Dual(double x) : v(x) {  // NOLINT: implicit from constants
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  __device__ __forceinline__ static Dual seed(double x, int k) {
    Dual r(x);
#pragma unroll
    for (int j = 0; j < K; ++j) r.d[j] = (j == k) ? 1.0 : 0.0;
    return r;
  }

  __device__ __forceinline__ Dual& operator+=(const Dual& o) {
    v += o.v;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] += o.d[k];
    return *this;
  }
  __device__ __forceinline__ Dual& operator-=(const Dual& o) {
    v -= o.v;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] -= o.d[k];
    return *this;
  }
  __device__ __forceinline__ Dual& operator*=(const Dual& o) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = d[k] * o.v + v * o.d[k];
    v *= o.v;
    return *this;
  }
  __device__ __forceinline__ Dual& operator/=(const Dual& o) {
    const double inv = 1.0 / o.v;
    const double q = v * inv;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = (d[k] - q * o.d[k]) * inv;
    v = q;
    return *this;
  }
};

template <int K>
__device__ __forceinline__ Dual<K> operator+(Dual<K> a, const Dual<K>& b) { return a += b; }
template <int K>
__device__ __forceinline__ Dual<K> operator-(Dual<K> a, const Dual<K>& b) { return a -= b; }
template <int K>
__device__ __forceinline__ Dual<K> operator*(Dual<K> a, const Dual<K>& b) { return a *= b; }
template <int K>
__device__ __forceinline__ Dual<K> operator/(Dual<K> a, const Dual<K>& b) { return a /= b; }
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ bool operator!=(const Dual<K>& a, const Dual<K>& b) { return a.v != b.v; }
template <int K>
__device__ __forceinline__ bool operator==(const Dual<K>& a, const Dual<K>& b) { return a.v == b.v; }
template <int K>
__device__ __forceinline__ bool operator>(const Dual<K>& a, const Dual<K>& b) { return a.v > b.v; }
template <int K>
__device__ __forceinline__ bool operator<(const Dual<K>& a, const Dual<K>& b) { return a.v < b.v; }

template <int K>
__device__ __forceinline__ Dual<K> sqrt(const Dual<K>& a) {
  Dual<K> r;
  r.v = ::sqrt(a.v);
  const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> sin(const Dual<K>& a) {
  Dual<K> r;
  r.v = ::sin(a.v);
  const double c = ::cos(a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * c;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> cos(const Dual<K>& a) {
  Dual<K> r;
  r.v = ::cos(a.v);
  const double s = -::sin(a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * s;
  return r;
}

}  // namespace ss
