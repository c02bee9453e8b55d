// Face-value arithmetic of MONC's TVD (ULTIMATE-QUICKEST) advection on gfx950.
//
// Each function computes the value on one face of one level for NS advected
// fields that share the same flows, from the limiter algebra of
//   components/tvdadvection/src/ultimateflux.F90:525-614 (limiter),
//   :616-676 (upwind stencils), :678-761 (transverse gradient terms),
//   :81-166 (x faces), :168-234 (y faces), :236-523 (z faces).
// The velocity-only quantities (outflow Courant sum and its reciprocal, face
// Courant number, transverse weights and upwind picks) are computed once per
// face and reused by every field; NS = 1 for a momentum field, whose flows are
// its own dual-grid interpolation.  Index conventions follow the reference:
// 1-based (k,j,i) = (z,y,x), z fastest, depth-2 halo in x and y.  The same
// templates serve the FP64 drop-in form and the FP32 form (T = double or
// float); metric vectors are always stored in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tvd_dual.hpp"

namespace tvd {

struct Dims {
  int nz, ny, nx;         // interior sizes; arrays are (nz, ny+4, nx+4)
  int64_t sj, si;         // strides of j and i in elements (nz, nz*(ny+4))
};

__device__ __forceinline__ int64_t at(const Dims& d, int k, int j, int i) {
  return (int64_t)(k - 1) + d.sj * (int64_t)(j - 1) + d.si * (int64_t)(i - 1);
}

// Raw buffer loads.  A field is addressed through a buffer resource (scalar
// registers) with a 32-bit per-lane offset and a 32-bit scalar offset: a wave
// of 64 levels at a wave-uniform column (j,i) puts (j,i) in the scalar offset
// and the level in the lane offset, and constant level shifts fold into the
// instruction's immediate offset, so neighbour loads cost no vector address
// arithmetic.  Fields must be smaller than 4 GiB (checked on the host).
template <typename T>
__device__ __forceinline__ T ldb(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 16)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  else if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): from_below gives
// lane L the value of lane L-1, from_above that of lane L+1; a lane shifted in
// from outside the wave reads 0.  Every lane of the wave must take part.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) { return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, true); }
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_shift(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp_i<CTRL>(__builtin_bit_cast(int, x)));
  } else {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = dpp_i<CTRL>((int)b), hi = dpp_i<CTRL>((int)(b >> 32));
    return __builtin_bit_cast(T, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  }
}
template <typename T>
__device__ __forceinline__ T from_below(T x) { return dpp_shift<0x138>(x); }
template <typename T>
__device__ __forceinline__ T from_above(T x) { return dpp_shift<0x130>(x); }
// lane `l`'s value, for every lane (wave-uniform)
template <typename T>
__device__ __forceinline__ T lane_value(T x, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
  } else {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(T, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  }
}

// A value every lane of the wave holds alike, kept in scalar registers: each
// 32-bit word through readfirstlane.  A value the compiler cannot prove
// uniform (one loaded from a table of members, say) otherwise stays in vector
// registers, and a buffer load whose scalar offset derives from it becomes a
// loop over the wave's distinct values.
template <typename S>
__device__ __forceinline__ S uniform(const S& x) {
  static_assert(sizeof(S) % 4 == 0, "whole 32-bit words");
  struct Words {
    int w[sizeof(S) / 4];
  };
  Words a = __builtin_bit_cast(Words, x);
#pragma unroll
  for (int n = 0; n < (int)(sizeof(S) / 4); n++) a.w[n] = __builtin_amdgcn_readfirstlane(a.w[n]);
  return __builtin_bit_cast(S, a);
}

// A pointer to global memory made uniform: its words through readfirstlane,
// and the result typed as a global (address space 1) pointer again, so that
// accesses through it stay global rather than generic (flat) instructions.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  typedef __attribute__((address_space(1))) T GT;
  const unsigned long long a = uniform((unsigned long long)p);
  return (T*)(GT*)a;
}

// gfx9 raw-buffer descriptor word 3 (32-bit data, no swizzle)
constexpr int BUFFER_RSRC_WORD3 = 0x00020000;

template <typename T>
struct Arr {
  __amdgpu_buffer_rsrc_t r;
  unsigned sjb, sib;  // byte strides of j and i
  __device__ __forceinline__ Arr() {}
  __device__ __forceinline__ Arr(const T* p, const Dims& d)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)0xffffffff, BUFFER_RSRC_WORD3)),
        sjb((unsigned)d.sj * (unsigned)sizeof(T)), sib((unsigned)d.si * (unsigned)sizeof(T)) {}
  // Level and row in the lane offset (one register per distinct row, level
  // shifts as immediates), column in the scalar offset (one scalar register
  // per distinct column).  i must be wave-uniform.
  __device__ __forceinline__ T operator()(int k, int j, int i) const {
    return ldb<T>(r, (unsigned)(k - 1) * (unsigned)sizeof(T) + (unsigned)(j - 1) * sjb, (unsigned)(i - 1) * sib);
  }
  __device__ __forceinline__ T lane(int k, int j, int i) const { return (*this)(k, j, i); }
};

// Vertical metrics as passed to ultflx.  For w the reference passes
// (rdz, rdzn, dzn) = (model rdzn, model rdz, model dz) and kdof = 1
// (tvdadvection.F90:881-894); the kernel builds this view accordingly.
struct VMet {
  const double* rdz;
  const double* rdzn;
  const double* dzn;
  const double* t1;  // tzc1 or tzd1
  const double* t2;  // tzc2 or tzd2
};

template <typename T>
__device__ __forceinline__ T vm(const double* v, int k) { return (T)v[k - 1]; }

template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
// In the working precisions max and min are one instruction each
// (v_max_f64, v_min_f32, ...) instead of a compare and a select per 32-bit
// half.  For numbers they give the value of a > b ? a : b (a < b ? a : b);
// at a tie of +0 and -0 the sign of the zero may differ, which no use here
// can observe (each result is added to a non-zero term or passed to fabs, or
// is a limiter ratio that multiplies a difference).  NaN operands aside.
// Written as the instructions themselves: the fmax/fmin builtins also quiet
// every operand first (IEEE mode), which costs an instruction per operand
// and measured 36-40 more registers in the scalar-pair kernels (spilling).
// A translation unit that defines TVD_MINMAX_SELECT keeps the compare and
// select: the persistent ensemble step measured 7.33 ms per step so against
// 8.33 ms with the instructions (the single-domain FP64 step 7.40 against
// 8.03 ms the other way).
#ifndef TVD_MINMAX_SELECT
__device__ __forceinline__ double tmax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double tmin(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float tmax(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float tmin(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
#endif
// Fortran SIGN(0.5,x): IEEE signed zero gives the negative branch for -0.0.
template <typename T>
__device__ __forceinline__ T half_sign(T x) { return copysign(T(0.5), x); }
template <typename T>
__device__ __forceinline__ bool sign_pos(T x) { return !signbit(x); }
// 1/x for the limiter's guarded denominators (x = delta + 1e-30 or
// S + 1e-30).  FP64: hardware reciprocal refined by two Newton steps (within
// an ulp of the correctly rounded quotient); FP32: the hardware reciprocal
// (1 ulp).  x = 0 never reaches a selected result: S >= 0, and a face whose
// delta + 1e-30 vanishes takes the upwind branch.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ Dual recip(Dual x) {
  const double r = recip(x.v);
  return Dual(r, -(r * r) * x.d);
}

// The reference's guard (ultimateflux.F90:548,551,576,579,604,607) adds 1e-30
// to the limiter denominators.  FP64 keeps it exactly; FP32 adds it with the
// denominator's sign, so a value of -1e-30 cannot cancel to zero.  The two
// differ only for |delta| below ~1e-23, far under FP32 field resolution.
__device__ __forceinline__ double guard(double x) { return x + 1.e-30; }
__device__ __forceinline__ float guard(float x) { return x + copysignf(1.e-30f, x); }
__device__ __forceinline__ Dual guard(Dual x) { return Dual(x.v + 1.e-30, x.d); }

template <typename T>
__device__ __forceinline__ T pos(T x) { return tmax(T(0.0), x); }
template <typename T>
__device__ __forceinline__ T negabs(T x) { return fabs(tmin(T(0.0), x)); }

// A flow field as ultflx sees it.  Scalars read the model array; momentum
// fields read its dual-grid interpolation (stencil.F90:147-169):
// 0.5*(f(k,j,i) + f(k+DK,j+DJ,i+DI)), raw f at k = nz when DK = 1.  The other
// raw-copy cases of the reference (shift past j = ny+4 or i = nx+4) are never
// reached: faces bounding interior cells read flows at j <= ny+3, i <= nx+3.
template <typename T, int DK, int DJ, int DI>
struct Flow {
  Arr<T> a;
  int nz;
  __device__ __forceinline__ Flow(const T* p, const Dims& d) : a(p, d), nz(d.nz) {}
  // (j,i) wave-uniform
  __device__ __forceinline__ T operator()(int k, int j, int i) const {
    T x = a(k, j, i);
    if constexpr (DK == 0 && DJ == 0 && DI == 0) {
      return x;
    } else if constexpr (DK == 1) {
      // the shifted load is clamped to level nz, where the raw value is used
      const int k1 = k < nz ? k + 1 : k;
      T y = a(k1, j + DJ, i + DI);
      return k < nz ? T(0.5) * (x + y) : x;
    } else {
      return T(0.5) * (x + a(k + DK, j + DJ, i + DI));
    }
  }
};

// NS advected fields (zf) read at the same positions.
template <typename T, int NS = 1>
struct Adv {
  Arr<T> a[NS];
  __device__ __forceinline__ T operator()(int n, int k, int j, int i) const { return a[n](k, j, i); }
  __device__ __forceinline__ T lane(int n, int k, int j, int i) const { return a[n].lane(k, j, i); }
};

// The arm of the limiter a face takes, in the terms of
// ultimateflux.F90:556-606: the face is fc (upwind branch, or ftemp = fncs),
// fd (ftemp = 1), the QUICKEST value ffs (ftemp = fnfs) or
// fu + (fc - fu)*hrcfl (ftemp = fnrefs).  With the arm fixed the face is linear
// in the advected field; the forward pass records it for the adjoint.
enum { LIM_C = 0, LIM_D = 1, LIM_Q = 2, LIM_R = 3 };
// The arm the selections of face() take, from the same operands.
template <typename T>
__device__ __forceinline__ int lim_class(T fcurvs, T fdels, T fnfs, T fncs, T fnrefs) {
  const T m1 = tmin(fnfs, fnrefs);
  const T m2 = tmin(m1, T(1.0));
  const int lim = fncs > m2 ? LIM_C : (m1 < T(1.0) ? (fnfs < fnrefs ? LIM_Q : LIM_R) : LIM_D);
  return fabs(fcurvs) >= fabs(fdels) ? LIM_C : lim;
}

// Limiter of ultimateflux.F90:556-606 (calculate_flux_in_x/_y): the
// velocity-only part, shared by every field advected by the same flows...
template <typename T>
struct HLim {
  T hrcfl, hcfl, cflcoef;
  __device__ __forceinline__ HLim(T vel, T ch, T dt, T sum_cfl_out) {
    const T r6 = T(1.0 / 6.0);
    T cflmod = fabs(vel * ch * dt);
    hrcfl = recip(sum_cfl_out + T(1.e-30));
    hcfl = T(0.5) * cflmod;
    cflcoef = r6 * (T(1.0) - cflmod * cflmod);
  }
  // ...and the per-field part; the upwind branch (|curvature| >= |delta|)
  // gives fc (ultimateflux.F90:104-105).
  __device__ __forceinline__ T face(T fu, T fc, T fd, T fgt1, T fgt2) const {
    T fdels = fd - fu;
    T fcurvs = fd - T(2.0) * fc + fu;
    T rfdels = recip(guard(fdels));
    T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * fcurvs - fgt1 - fgt2;
    T fnfs = (ffs - fu) * rfdels;
    T fncs = (fc - fu) * rfdels;
    T fnrefs = fncs * hrcfl;
    T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
    T lim = ftemp * fdels + fu;
    return fabs(fcurvs) >= fabs(fdels) ? fc : lim;
  }
  // face() and the arm it takes
  __device__ __forceinline__ T face_cls(T fu, T fc, T fd, T fgt1, T fgt2, int& cls) const {
    T fdels = fd - fu;
    T fcurvs = fd - T(2.0) * fc + fu;
    T rfdels = recip(guard(fdels));
    T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * fcurvs - fgt1 - fgt2;
    T fnfs = (ffs - fu) * rfdels;
    T fncs = (fc - fu) * rfdels;
    T fnrefs = fncs * hrcfl;
    T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
    T lim = ftemp * fdels + fu;
    cls = lim_class(fcurvs, fdels, fnfs, fncs, fnrefs);
    return fabs(fcurvs) >= fabs(fdels) ? fc : lim;
  }
  // the arm alone
  __device__ __forceinline__ int cls_of(T fu, T fc, T fd, T fgt1, T fgt2) const {
    int cls;
    (void)face_cls(fu, fc, fd, fgt1, fgt2, cls);
    return cls;
  }
};

// ultimateflux.F90:525-554 calculate_flux_in_w: velocity/metric part (k = the
// level whose top face) and per-field part.
template <typename T>
struct WLim {
  T hrcfl, hcfl, cflcoef, rdd, rdu;
  __device__ __forceinline__ WLim(int k, T wval, T dt, T sum_cfl_out, T rdc, T rdu_, const VMet& m) {
    const T r6 = T(1.0 / 6.0);
    T dd = vm<T>(m.dzn, k + 1);
    rdd = vm<T>(m.rdzn, k + 1);
    rdu = rdu_;
    T cflmod = fabs(wval * rdd * dt);
    hrcfl = recip(sum_cfl_out + T(1.e-30));
    hcfl = T(0.5) * cflmod;
    cflcoef = r6 * (T(1.0) - cflmod * cflmod) * (dd * dd * rdc);
  }
  __device__ __forceinline__ T face(T fu, T fc, T fd, T fgt1, T fgt2) const {
    T fcurvs = fd - T(2.0) * fc + fu;
    T fdels = fd - fu;
    T rfdels = recip(guard(fdels));
    T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * ((fd - fc) * rdd - (fc - fu) * rdu) - fgt1 - fgt2;
    T fnfs = (ffs - fu) * rfdels;
    T fncs = (fc - fu) * rfdels;
    T fnrefs = fncs * hrcfl;
    T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
    T lim = ftemp * fdels + fu;
    return fabs(fcurvs) >= fabs(fdels) ? fc : lim;
  }
  __device__ __forceinline__ T face_cls(T fu, T fc, T fd, T fgt1, T fgt2, int& cls) const {
    T fcurvs = fd - T(2.0) * fc + fu;
    T fdels = fd - fu;
    T rfdels = recip(guard(fdels));
    T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * ((fd - fc) * rdd - (fc - fu) * rdu) - fgt1 - fgt2;
    T fnfs = (ffs - fu) * rfdels;
    T fncs = (fc - fu) * rfdels;
    T fnrefs = fncs * hrcfl;
    T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
    T lim = ftemp * fdels + fu;
    cls = lim_class(fcurvs, fdels, fnfs, fncs, fnrefs);
    return fabs(fcurvs) >= fabs(fdels) ? fc : lim;
  }
  // the arm alone
  __device__ __forceinline__ int cls_of(T fu, T fc, T fd, T fgt1, T fgt2) const {
    int cls;
    (void)face_cls(fu, fc, fd, fgt1, fgt2, cls);
    return cls;
  }
};

// Right x face of column (j,i) at level k: ultimateflux.F90:102-165.
template <typename T, int NS, class FU, class FV, class FW>
__device__ __forceinline__ void face_x(int k, int j, int i, const Adv<T, NS>& F, const FU& U, const FV& V,
                                       const FW& W, const VMet& m, int kdof, T dt, T cx, T cy, int nz, T out[NS]) {
  const int kp1 = k + 1 < nz ? k + 1 : nz;
  const int kp0 = kp1 - 1;
  const T uk = U(k, j, i);
  const T sm = half_sign(uk);
  const T ap = T(0.5) + sm, an = T(0.5) - sm;
  // outflow Courant sum, :108-139 (strict U > 0 chooses the column)
  const bool upos = uk > T(0.0);
  const T w0 = W(k, j, i), w0m = W(k - 1, j, i), w1 = W(k, j, i + 1), w1m = W(k - 1, j, i + 1);
  const T v0 = V(k, j, i), v0m = V(k, j - 1, i), v1 = V(k, j, i + 1), v1m = V(k, j - 1, i + 1);
  T s = vm<T>(m.rdz, k) * (pos(upos ? w0 : w1) + negabs(upos ? w0m : w1m));
  s = s + cy * (pos(upos ? v0 : v1) + negabs(upos ? v0m : v1m));
  s = s + cx * (upos ? uk + negabs(U(k, j, i - 1)) : pos(U(k, j, i + 1)) - uk);
  s = s * dt;
  // :678-695 gradient term in z, two source columns
  const T sz = T(0.125) * (w0 + w0m + w1 + w1m);
  const bool pz = sign_pos(sz);
  const int ksu = pz ? k - 1 : kp1;
  const int ksl = pz ? k : kp0;
  const T cz = fabs(sz) * vm<T>(m.rdz, k + kdof) * dt;
  // :711-761 gradient term in y, "data_two" form
  const T sy = T(0.125) * (v1 + v0 + v1m + v0m);
  const int js = sign_pos(sy) ? j - 1 : j + 1;
  const T cyt = fabs(sy) * cy * dt;
  const HLim<T> lim(uk, cx, dt, s);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fa = F(n, k, j, i), fb = F(n, k, j, i + 1);
    const T fu = F(n, k, j, i - 1) * ap + F(n, k, j, i + 2) * an;
    const T fc = fa * ap + fb * an;
    const T fd = fb * ap + fa * an;
    const T tmp = (F(n, ksl, j, i) - F(n, ksu, j, i)) * ap + (F(n, ksl, j, i + 1) - F(n, ksu, j, i + 1)) * an;
    const T g1 = cz * tmp;
    const T t = F.lane(n, k, js, i) * ap + F.lane(n, k, js, i + 1) * an;
    const T g2 = cyt * (fc - t);
    out[n] = lim.face(fu, fc, fd, g1, g2);
  }
}

// North y face of row j at column i, level k: ultimateflux.F90:189-233
// (j >= 2 always here, so jperiod = 0).
template <typename T, int NS, class FU, class FV, class FW>
__device__ __forceinline__ void face_y(int k, int j, int i, const Adv<T, NS>& F, const FU& U, const FV& V,
                                       const FW& W, const VMet& m, int kdof, T dt, T cx, T cy, int nz, T out[NS]) {
  const int kp1 = k + 1 < nz ? k + 1 : nz;
  const int kp0 = kp1 - 1;
  const T vk = V(k, j, i);
  const T sm = half_sign(vk);
  const T ap = T(0.5) + sm, an = T(0.5) - sm;
  const bool vpos = sign_pos(vk);
  const int jj = vpos ? j : j + 1;  // j + jofset
  const T w0 = W(k, j, i), w0m = W(k - 1, j, i), w1 = W(k, j + 1, i), w1m = W(k - 1, j + 1, i);
  const T vn = V(k, j + 1, i), vs = V(k, j - 1, i);
  const T u0 = U(k, j, i), u0m = U(k, j, i - 1), u1 = U(k, j + 1, i), u1m = U(k, j + 1, i - 1);
  T s = vm<T>(m.rdz, k) * (pos(vpos ? w0 : w1) + negabs(vpos ? w0m : w1m));
  s = s + cy * (pos(vpos ? vk : vn) + negabs(vpos ? vs : vk));
  s = s + cx * (pos(vpos ? u0 : u1) + negabs(vpos ? u0m : u1m));
  s = s * dt;
  const T sz = T(0.125) * (w0 + w0m + w1 + w1m);
  const bool pz = sign_pos(sz);
  const int ksu = pz ? k - 1 : kp1;
  const int ksl = pz ? k : kp0;
  const T cz = fabs(sz) * vm<T>(m.rdz, k + kdof) * dt;
  const T sx = T(0.125) * (u0m + u0 + u1m + u1);
  const T posuf = T(0.5) + half_sign(sx);
  const T cxt = fabs(sx) * cx * dt;
  const HLim<T> lim(vk, cy, dt, s);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fa = F(n, k, j, i), fb = F(n, k, j + 1, i);
    const T fu = F(n, k, j - 1, i) * ap + F(n, k, j + 2, i) * an;
    const T fc = fa * ap + fb * an;
    const T fd = fb * ap + fa * an;
    const T g1 = cz * (F.lane(n, ksl, jj, i) - F.lane(n, ksu, jj, i));
    const T g2 = cxt * (fc - (F.lane(n, k, jj, i + 1) * (T(1.0) - posuf) + F.lane(n, k, jj, i - 1) * posuf));
    out[n] = lim.face(fu, fc, fd, g1, g2);
  }
}

// Transverse weights of the z faces (:711-761 single-source form and
// :697-709 with kpv = 1), shared by the fields.
template <typename T, class FU, class FV>
struct ZGrad {
  int js;
  T cyt, posuf, cxt;
  __device__ __forceinline__ ZGrad(int k, int j, int i, const FU& U, const FV& V, T dt, T cx, T cy) {
    const T sy = T(0.125) * (V(k, j - 1, i) + V(k, j, i) + V(k + 1, j - 1, i) + V(k + 1, j, i));
    js = sign_pos(sy) ? j - 1 : j + 1;
    cyt = fabs(sy) * cy * dt;
    const T sx = T(0.125) * (U(k, j, i - 1) + U(k, j, i) + U(k + 1, j, i - 1) + U(k + 1, j, i));
    posuf = T(0.5) + half_sign(sx);
    cxt = fabs(sx) * cx * dt;
  }
  template <int NS>
  __device__ __forceinline__ void apply(const Adv<T, NS>& F, int n, int kk, int j, int i, T fc, T& g1,
                                        T& g2) const {
    g1 = cyt * (fc - F.lane(n, kk, js, i));
    g2 = cxt * (fc - (F(n, kk, j, i + 1) * (T(1.0) - posuf) + F(n, kk, j, i - 1) * posuf));
  }
};

// Top face of level k (the reference's flux_z(k+1)) for 1 <= k <= nz-1.
// k = 1 with kmin = 1 (w): :307-411; k = nz-1: :413-523; otherwise :236-305.
// Scalars, u and v have kmin = 2 and flux_z(2) = 0 (:78).
template <typename T, int NS, class FU, class FV, class FW>
__device__ __forceinline__ void face_z(int k, int j, int i, const Adv<T, NS>& F, const FU& U, const FV& V,
                                       const FW& W, const VMet& m, int kdof, int kmin, T dt, T cx, T cy, int nz,
                                       T out[NS]) {
  if (k == 1) {
    if (kmin > 1) {
#pragma unroll
      for (int n = 0; n < NS; n++) out[n] = T(0.0);
      return;
    }
    const T wk = W(1, j, i);
    const bool up = wk >= T(0.0);
    T s, rdu, rdc;
    if (up) {
      s = wk * vm<T>(m.rdz, 1 + kdof);
      s = s + cy * (pos(V(1, j, i)) + negabs(V(1, j - 1, i)));
      s = s + cx * (pos(U(1, j, i)) + negabs(U(1, j, i - 1)));
      rdu = vm<T>(m.rdzn, 2);
      rdc = vm<T>(m.rdz, 1 + kdof);
    } else {
      s = vm<T>(m.rdz, 2 + kdof) * (pos(W(2, j, i)) - wk);
      s = s + cy * (pos(V(2, j, i)) + negabs(V(2, j - 1, i)));
      s = s + cx * (pos(U(2, j, i)) + negabs(U(2, j, i - 1)));
      rdu = vm<T>(m.rdzn, 3);
      rdc = vm<T>(m.rdz, 2 + kdof);
    }
    s = s * dt;
    const int kk = up ? 1 : 2;
    const ZGrad<T, FU, FV> zg(1, j, i, U, V, dt, cx, cy);
    const WLim<T> lim(1, wk, dt, s, rdc, rdu, m);
#pragma unroll
    for (int n = 0; n < NS; n++) {
      const T f1 = F(n, 1, j, i), f2 = F(n, 2, j, i);
      const T fu = up ? f1 : F(n, 3, j, i);
      const T fc = up ? f1 : f2;
      const T fd = up ? f2 : f1;
      T g1, g2;
      zg.apply(F, n, kk, j, i, fc, g1, g2);
      out[n] = lim.face(fu, fc, fd, g1, g2);
    }
    return;
  }
  if (k == nz - 1) {
    const T wk = W(k, j, i);
    const bool up = wk >= T(0.0);
    T s, rdu, rdc;
    if (up) {
      s = vm<T>(m.rdz, k + kdof) * (wk + negabs(W(k - 1, j, i)));
      s = s + cy * (pos(V(k, j, i)) + negabs(V(k, j - 1, i)));
      s = s + cx * (pos(U(k, j, i)) + negabs(U(k, j, i - 1)));
      rdu = vm<T>(m.rdzn, k);
      rdc = vm<T>(m.rdz, k + kdof);
    } else {
      const int kt = k + 1 + kdof < nz ? k + 1 + kdof : nz;
      s = -(wk * vm<T>(m.rdz, kt));
      s = s + cy * (pos(V(k + 1, j, i)) + negabs(V(k + 1, j - 1, i)));
      s = s + cx * (pos(U(k + 1, j, i)) + negabs(U(k + 1, j, i - 1)));
      rdu = vm<T>(m.rdzn, nz);
      rdc = vm<T>(m.rdz, kt);
    }
    s = s * dt;
    const int kk = up ? k : k + 1;
    const ZGrad<T, FU, FV> zg(k, j, i, U, V, dt, cx, cy);
    const WLim<T> lim(k, wk, dt, s, rdc, rdu, m);
#pragma unroll
    for (int n = 0; n < NS; n++) {
      const T fk = F(n, k, j, i), fk1 = F(n, k + 1, j, i);
      // W < 0: linear extrapolation to a level equidistant above the top
      const T fu = up ? F(n, k - 1, j, i) : T(2.0) * fk1 - fk;
      const T fc = up ? fk : fk1;
      const T fd = up ? fk1 : fk;
      T g1, g2;
      zg.apply(F, n, kk, j, i, fc, g1, g2);
      out[n] = lim.face(fu, fc, fd, g1, g2);
    }
    return;
  }
  // middle, 2 <= k <= nz-2
  const T wk = W(k, j, i);
  const T sm = half_sign(wk);
  const T ap = T(0.5) + sm, an = T(0.5) - sm;
  const int kk = sign_pos(wk) ? k : k + 1;  // k + kofset
  T s = vm<T>(m.rdz, kk) * (pos(W(kk, j, i)) + negabs(W(kk - 1, j, i)));
  s = s + cy * (pos(V(kk, j, i)) + negabs(V(kk, j - 1, i)));
  s = s + cx * (pos(U(kk, j, i)) + negabs(U(kk, j, i - 1)));
  s = s * dt;
  const ZGrad<T, FU, FV> zg(k, j, i, U, V, dt, cx, cy);
  const T rdu = vm<T>(m.rdzn, k) * ap + vm<T>(m.rdzn, k + 2) * an;
  const T rdc = vm<T>(m.rdz, k + kdof) * ap + vm<T>(m.rdz, k + 1 + kdof) * an;
  const WLim<T> lim(k, wk, dt, s, rdc, rdu, m);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fa = F(n, k, j, i), fb = F(n, k + 1, j, i);
    const T fu = F(n, k - 1, j, i) * ap + F(n, k + 2, j, i) * an;
    const T fc = F(n, kk, j, i);
    const T fd = fb * ap + fa * an;
    T g1, g2;
    zg.apply(F, n, kk, j, i, fc, g1, g2);
    out[n] = lim.face(fu, fc, fd, g1, g2);
  }
}

}  // namespace tvd
