// Face-value arithmetic of MONC's TVD (ULTIMATE-QUICKEST) advection on gfx950.
//
// Every function here computes ONE face value of one advected field F for one
// level, from the limiter algebra of
//   components/tvdadvection/src/ultimateflux.F90:525-614 (limiter),
//   :616-676 (upwind stencils), :678-761 (transverse gradient terms),
//   :81-166 (x faces), :168-234 (y faces), :236-523 (z faces),
// and the flux divergence of components/tvdadvection/src/tvdadvection.F90:911-960.
// Index conventions follow the reference: 1-based (k,j,i) = (z,y,x), z fastest,
// depth-2 halo in x and y.  The same templates serve the FP64 drop-in form and
// the FP32 form (T = double or float); metric vectors are always read as FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tvd {

struct Dims {
  int nz, ny, nx;         // interior sizes; arrays are (nz, ny+4, nx+4)
  int64_t sj, si;         // strides of j and i in elements (nz, nz*(ny+4))
};

__device__ __forceinline__ int64_t at(const Dims& d, int k, int j, int i) {
  return (int64_t)(k - 1) + d.sj * (int64_t)(j - 1) + d.si * (int64_t)(i - 1);
}

// Element (k,j,i) as column base + level.  When j and i are wave-uniform the
// base lives in scalar registers and every neighbour load is a scalar base
// plus the one 32-bit per-lane level offset.
template <typename T>
__device__ __forceinline__ T ld(const T* a, const Dims& d, int k, int j, int i) {
  const T* col = a + (d.sj * (int64_t)(j - 1) + d.si * (int64_t)(i - 1));
  return col[(unsigned)(k - 1)];
}

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
// Fortran SIGN(0.5,x): IEEE signed zero gives the negative branch for -0.0.
template <typename T>
__device__ __forceinline__ T half_sign(T x) { return copysign(T(0.5), x); }
template <typename T>
__device__ __forceinline__ bool sign_pos(T x) { return !signbit(x); }

// A flow field as ultflx sees it.  Scalars read the model array; momentum
// fields read its dual-grid interpolation (stencil.F90:147-169):
// 0.5*(f(k,j,i) + f(k+DK,j+DJ,i+DI)), raw f at k = nz when DK = 1.  The other
// raw-copy cases of the reference (shift past j = ny+4 or i = nx+4) are never
// reached: faces bounding interior cells read flows at j <= ny+3, i <= nx+3.
template <typename T, int DK, int DJ, int DI>
struct Flow {
  const T* a;
  Dims d;
  __device__ __forceinline__ T operator()(int k, int j, int i) const {
    T x = ld(a, d, k, j, i);
    if constexpr (DK == 0 && DJ == 0 && DI == 0) {
      return x;
    } else {
      if constexpr (DK == 1) {
        if (k == d.nz) return x;
      }
      return T(0.5) * (x + ld(a, d, k + DK, j + DJ, i + DI));
    }
  }
};

template <typename T>
struct Adv {  // the advected field zf
  const T* a;
  Dims d;
  __device__ __forceinline__ T operator()(int k, int j, int i) const { return ld(a, d, k, j, i); }
};

// ultimateflux.F90:556-606 calculate_flux_in_x / _y (ch = cx or cy)
template <typename T>
__device__ __forceinline__ T limit_h(T vel, T dt, T sum_cfl_out, T fdels, T fd, T fc, T fu, T fgt1, T fgt2, T ch,
                                     T fcurvs) {
  const T r6 = T(1.0) / T(6.0);
  T cflmod = fabs(vel * ch * dt);
  T hrcfl = T(1.0) / (sum_cfl_out + T(1.e-30));
  T hcfl = T(0.5) * cflmod;
  T cflcoef = r6 * (T(1.0) - cflmod * cflmod);
  T rfdels = T(1.0) / (fdels + T(1.e-30));
  T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * fcurvs - fgt1 - fgt2;
  T fnfs = (ffs - fu) * rfdels;
  T fncs = (fc - fu) * rfdels;
  T fnrefs = fncs * hrcfl;
  T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
  return ftemp * fdels + fu;
}

// ultimateflux.F90:525-554 calculate_flux_in_w (k = the level whose top face)
template <typename T>
__device__ __forceinline__ T limit_w(int k, T wval, T dt, T sum_cfl_out, T rdc, T fdels, T fd, T fc, T fu, T rdu,
                                     T fgt1, T fgt2, const VMet& m) {
  const T r6 = T(1.0) / T(6.0);
  T dd = vm<T>(m.dzn, k + 1), rdd = vm<T>(m.rdzn, k + 1);
  T cflmod = fabs(wval * rdd * dt);
  T hrcfl = T(1.0) / (sum_cfl_out + T(1.e-30));
  T hcfl = T(0.5) * cflmod;
  T cflcoef = r6 * (T(1.0) - cflmod * cflmod) * (dd * dd * rdc);
  T rfdels = T(1.0) / (fdels + T(1.e-30));
  T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * ((fd - fc) * rdd - (fc - fu) * rdu) - fgt1 - fgt2;
  T fnfs = (ffs - fu) * rfdels;
  T fncs = (fc - fu) * rfdels;
  T fnrefs = fncs * hrcfl;
  T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
  return ftemp * fdels + fu;
}

template <typename T>
__device__ __forceinline__ T pos(T x) { return tmax(T(0.0), x); }
template <typename T>
__device__ __forceinline__ T negabs(T x) { return fabs(tmin(T(0.0), x)); }

// Right x face of column (j,i) at level k: ultimateflux.F90:102-165.
template <typename T, class FU, class FV, class FW>
__device__ T face_x(int k, int j, int i, const Adv<T>& F, const FU& U, const FV& V, const FW& W, const VMet& m,
                    int kdof, T dt, T cx, T cy, int nz) {
  int kp1 = k + 1 < nz ? k + 1 : nz;
  int kp0 = kp1 - 1;
  T uk = U(k, j, i);
  T sm = half_sign(uk);
  T ap = T(0.5) + sm, an = T(0.5) - sm;
  T fu = F(k, j, i - 1) * ap + F(k, j, i + 2) * an;
  T fc = F(k, j, i) * ap + F(k, j, i + 1) * an;
  T fd = F(k, j, i + 1) * ap + F(k, j, i) * an;
  T fdels = fd - fu;
  T fcurvs = fd - T(2.0) * fc + fu;
  if (fabs(fcurvs) >= fabs(fdels)) return fc;
  T s;
  if (uk > T(0.0)) {
    s = vm<T>(m.rdz, k) * (pos(W(k, j, i)) + negabs(W(k - 1, j, i)));
    s = s + cy * (pos(V(k, j, i)) + negabs(V(k, j - 1, i)));
    s = s + cx * (uk + negabs(U(k, j, i - 1)));
  } else {
    s = vm<T>(m.rdz, k) * (pos(W(k, j, i + 1)) + negabs(W(k - 1, j, i + 1)));
    s = s + cy * (pos(V(k, j, i + 1)) + negabs(V(k, j - 1, i + 1)));
    s = s + cx * (pos(U(k, j, i + 1)) - uk);
  }
  s = s * dt;
  // :678-695 gradient term in z, two source columns
  T g1;
  {
    T sf = T(0.125) * (W(k, j, i) + W(k - 1, j, i) + W(k, j, i + 1) + W(k - 1, j, i + 1));
    bool p = sign_pos(sf);
    int ksu = p ? k - 1 : kp1;
    int ksl = p ? k : kp0;
    T tmp = (F(ksl, j, i) - F(ksu, j, i)) * ap + (F(ksl, j, i + 1) - F(ksu, j, i + 1)) * an;
    g1 = fabs(sf) * vm<T>(m.rdz, k + kdof) * dt * tmp;
  }
  // :711-761 gradient term in y, "data_two" form
  T g2;
  {
    T sf = T(0.125) * (V(k, j, i + 1) + V(k, j, i) + V(k, j - 1, i + 1) + V(k, j - 1, i));
    int js = sign_pos(sf) ? j - 1 : j + 1;
    T t = F(k, js, i) * ap + F(k, js, i + 1) * an;
    g2 = fabs(sf) * cy * dt * (fc - t);
  }
  return limit_h(uk, dt, s, fdels, fd, fc, fu, g1, g2, cx, fcurvs);
}

// North y face of row j at column i, level k: ultimateflux.F90:189-233
// (j >= 2 always here, so jperiod = 0).
template <typename T, class FU, class FV, class FW>
__device__ T face_y(int k, int j, int i, const Adv<T>& F, const FU& U, const FV& V, const FW& W, const VMet& m,
                    int kdof, T dt, T cx, T cy, int nz) {
  int kp1 = k + 1 < nz ? k + 1 : nz;
  int kp0 = kp1 - 1;
  T vk = V(k, j, i);
  T sm = half_sign(vk);
  T ap = T(0.5) + sm, an = T(0.5) - sm;
  int jo = sign_pos(vk) ? 0 : 1;
  T fu = F(k, j - 1, i) * ap + F(k, j + 2, i) * an;
  T fc = F(k, j, i) * ap + F(k, j + 1, i) * an;
  T fd = F(k, j + 1, i) * ap + F(k, j, i) * an;
  T fdels = fd - fu;
  T fcurvs = fd - T(2.0) * fc + fu;
  if (fabs(fcurvs) >= fabs(fdels)) return fc;
  int jj = j + jo;
  T s = vm<T>(m.rdz, k) * (pos(W(k, jj, i)) + negabs(W(k - 1, jj, i)));
  s = s + cy * (pos(V(k, jj, i)) + negabs(V(k, jj - 1, i)));
  s = s + cx * (pos(U(k, jj, i)) + negabs(U(k, jj, i - 1)));
  s = s * dt;
  T g1;
  {
    T sf = T(0.125) * (W(k, j, i) + W(k - 1, j, i) + W(k, j + 1, i) + W(k - 1, j + 1, i));
    bool p = sign_pos(sf);
    int ksu = p ? k - 1 : kp1;
    int ksl = p ? k : kp0;
    g1 = fabs(sf) * vm<T>(m.rdz, k + kdof) * dt * (F(ksl, jj, i) - F(ksu, jj, i));
  }
  T g2;
  {
    T sf = T(0.125) * (U(k, j, i - 1) + U(k, j, i) + U(k, j + 1, i - 1) + U(k, j + 1, i));
    T posuf = T(0.5) + half_sign(sf);
    g2 = fabs(sf) * cx * dt * (fc - (F(k, jj, i + 1) * (T(1.0) - posuf) + F(k, jj, i - 1) * posuf));
  }
  return limit_h(vk, dt, s, fdels, fd, fc, fu, g1, g2, cy, fcurvs);
}

// Transverse terms of the z faces (:711-761 single-source form, :697-709 kpv=1)
template <typename T, class FU, class FV>
__device__ __forceinline__ void zface_grads(int k, int j, int i, int kk, T fc, const Adv<T>& F, const FU& U,
                                            const FV& V, T dt, T cx, T cy, T& g1, T& g2) {
  {
    T sf = T(0.125) * (V(k, j - 1, i) + V(k, j, i) + V(k + 1, j - 1, i) + V(k + 1, j, i));
    int js = sign_pos(sf) ? j - 1 : j + 1;
    g1 = fabs(sf) * cy * dt * (fc - F(kk, js, i));
  }
  {
    T sf = T(0.125) * (U(k, j, i - 1) + U(k, j, i) + U(k + 1, j, i - 1) + U(k + 1, j, i));
    T posuf = T(0.5) + half_sign(sf);
    g2 = fabs(sf) * cx * dt * (fc - (F(kk, j, i + 1) * (T(1.0) - posuf) + F(kk, j, i - 1) * posuf));
  }
}

// Top face of level k (the reference's flux_z(k+1)) for 1 <= k <= nz-1.
// k = 1 with kmin = 1 (w): :307-411; k = nz-1: :413-523; otherwise :236-305.
// Scalars, u and v have kmin = 2 and flux_z(2) = 0 (:78).
template <typename T, class FU, class FV, class FW>
__device__ T face_z(int k, int j, int i, const Adv<T>& F, const FU& U, const FV& V, const FW& W, const VMet& m,
                    int kdof, int kmin, T dt, T cx, T cy, int nz) {
  if (k == 1) {
    if (kmin > 1) return T(0.0);
    T wk = W(1, j, i);
    T fu, fc, fd;
    bool up = wk >= T(0.0);
    if (up) {
      fu = F(1, j, i);
      fc = F(1, j, i);
      fd = F(2, j, i);
    } else {
      fu = F(3, j, i);
      fc = F(2, j, i);
      fd = F(1, j, i);
    }
    T fcurvs = fd - T(2.0) * fc + fu;
    T fdels = fd - fu;
    if (fabs(fcurvs) >= fabs(fdels)) return fc;
    T s, rdu, rdc, g1, g2;
    int kk;
    if (up) {
      s = wk * vm<T>(m.rdz, 1 + kdof);
      s = s + cy * (pos(V(1, j, i)) + negabs(V(1, j - 1, i)));
      s = s + cx * (pos(U(1, j, i)) + negabs(U(1, j, i - 1)));
      kk = 1;
      rdu = vm<T>(m.rdzn, 2);
      rdc = vm<T>(m.rdz, 1 + kdof);
    } else {
      s = vm<T>(m.rdz, 2 + kdof) * (pos(W(2, j, i)) - wk);
      s = s + cy * (pos(V(2, j, i)) + negabs(V(2, j - 1, i)));
      s = s + cx * (pos(U(2, j, i)) + negabs(U(2, j, i - 1)));
      kk = 2;
      rdu = vm<T>(m.rdzn, 3);
      rdc = vm<T>(m.rdz, 2 + kdof);
    }
    s = s * dt;
    zface_grads(1, j, i, kk, fc, F, U, V, dt, cx, cy, g1, g2);
    return limit_w(1, wk, dt, s, rdc, fdels, fd, fc, fu, rdu, g1, g2, m);
  }
  if (k == nz - 1) {
    T wk = W(k, j, i);
    bool up = wk >= T(0.0);
    T fu, fc, fd;
    if (up) {
      fu = F(k - 1, j, i);
      fc = F(k, j, i);
      fd = F(k + 1, j, i);
    } else {
      fu = T(2.0) * F(k + 1, j, i) - F(k, j, i);
      fc = F(k + 1, j, i);
      fd = F(k, j, i);
    }
    T fcurvs = fd - T(2.0) * fc + fu;
    T fdels = fd - fu;
    if (fabs(fcurvs) >= fabs(fdels)) return fc;
    T s, rdu, rdc, g1, g2;
    int kk;
    if (up) {
      s = vm<T>(m.rdz, k + kdof) * (wk + negabs(W(k - 1, j, i)));
      s = s + cy * (pos(V(k, j, i)) + negabs(V(k, j - 1, i)));
      s = s + cx * (pos(U(k, j, i)) + negabs(U(k, j, i - 1)));
      kk = k;
      rdu = vm<T>(m.rdzn, k);
      rdc = vm<T>(m.rdz, k + kdof);
    } else {
      int kt = k + 1 + kdof < nz ? k + 1 + kdof : nz;
      s = -(wk * vm<T>(m.rdz, kt));
      s = s + cy * (pos(V(k + 1, j, i)) + negabs(V(k + 1, j - 1, i)));
      s = s + cx * (pos(U(k + 1, j, i)) + negabs(U(k + 1, j, i - 1)));
      kk = k + 1;
      rdu = vm<T>(m.rdzn, nz);
      rdc = vm<T>(m.rdz, kt);
    }
    s = s * dt;
    zface_grads(k, j, i, kk, fc, F, U, V, dt, cx, cy, g1, g2);
    return limit_w(k, wk, dt, s, rdc, fdels, fd, fc, fu, rdu, g1, g2, m);
  }
  // middle, 2 <= k <= nz-2
  T wk = W(k, j, i);
  T sm = half_sign(wk);
  T ap = T(0.5) + sm, an = T(0.5) - sm;
  int ko = sign_pos(wk) ? 0 : 1;
  T fu = F(k - 1, j, i) * ap + F(k + 2, j, i) * an;
  T fc = F(k + ko, j, i);
  T fd = F(k + 1, j, i) * ap + F(k, j, i) * an;
  T fcurvs = fd - T(2.0) * fc + fu;
  T fdels = fd - fu;
  if (fabs(fcurvs) >= fabs(fdels)) return fc;
  int kk = k + ko;
  T s = vm<T>(m.rdz, kk) * (pos(W(kk, j, i)) + negabs(W(kk - 1, j, i)));
  s = s + cy * (pos(V(kk, j, i)) + negabs(V(kk, j - 1, i)));
  s = s + cx * (pos(U(kk, j, i)) + negabs(U(kk, j, i - 1)));
  s = s * dt;
  T g1, g2;
  zface_grads(k, j, i, kk, fc, F, U, V, dt, cx, cy, g1, g2);
  T rdu = vm<T>(m.rdzn, k) * ap + vm<T>(m.rdzn, k + 2) * an;
  T rdc = vm<T>(m.rdz, k + kdof) * ap + vm<T>(m.rdz, k + 1 + kdof) * an;
  return limit_w(k, wk, dt, s, rdc, fdels, fd, fc, fu, rdu, g1, g2, m);
}

}  // namespace tvd
