// Dual numbers (value, tangent) for the tangent-linear form of the TVD step.
//
// Instantiating the forward kernels with T = Dual gives the exact derivative
// of the code path the parity tests pin to the oracle: every branch
// (upwind/limited, max/min choice, flow sign) is taken on the value part, and
// the tangent follows the chosen operand.  The hand-written adjoint is tested
// against this tangent-linear model (dot-product test), so the two must use
// the same one-sided conventions at kinks: d|x| = sign-bit(x) ? -dx : dx.
#pragma once
#include <hip/hip_runtime.h>

namespace tvd {

// the overloads below must not hide the scalar functions inside the namespace
using ::copysign;
using ::fabs;
using ::signbit;
using ::__shfl;
using ::__shfl_up;

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
  __host__ __device__ explicit Dual(int x) : v((double)x), d(0.0) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(Dual a, Dual b) { return a.v == b.v; }
__device__ __forceinline__ Dual fabs(Dual a) { return __builtin_signbit(a.v) ? -a : a; }
__device__ __forceinline__ Dual copysign(Dual a, Dual b) { return Dual(__builtin_copysign(a.v, b.v), 0.0); }
__device__ __forceinline__ bool signbit(Dual a) { return __builtin_signbit(a.v); }
__device__ __forceinline__ Dual __shfl_up(Dual a, unsigned delta) {
  return Dual(::__shfl_up(a.v, delta), ::__shfl_up(a.d, delta));
}
__device__ __forceinline__ Dual __shfl(Dual a, int lane) { return Dual(::__shfl(a.v, lane), ::__shfl(a.d, lane)); }

}  // namespace tvd
