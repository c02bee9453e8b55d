// Kernels of the TVD advection path: periodic halo fill, the tendency sweep
// (drop-in `s += ...` form) and the fused forward step.
#pragma once
#include "tvd_faces.hpp"

namespace tvd {

enum FieldKind { KIND_U = 0, KIND_V = 1, KIND_W = 2, KIND_SCALAR = 3 };
// MODE_ACCUM: s += tendency (the drop-in timestep callback); MODE_STEP: the
// fused forward-Euler step; MODE_LEAP: the fused MONC step of an
// advection-only run (stepfields + swap_smooth: leapfrog with the Robert
// filter, or the forward first step with the classic swap).
enum Mode { MODE_ACCUM = 0, MODE_STEP = 1, MODE_LEAP = 2, MODE_REC = 3 };
// MODE_REC: MODE_STEP that also records each face's limiter arm (LIM_*) for
// the adjoint, one byte per face owner (k, j, i) and field in the field's
// layout: bits 0-1 the right x face, 2-3 the north y face, 4-5 the top z face.

// Largest number of fields one launch advects (u, v, w, th and up to 12
// scalars); the host splits longer scalar lists over several launches.
constexpr int MAX_FIELDS = 16;

template <typename T>
struct FieldSet {
  int nf;                  // fields in this launch
  int kind[MAX_FIELDS];    // KIND_U, KIND_V, KIND_W or KIND_SCALAR
  const T* F[MAX_FIELDS];  // advected field (prognostic or z-field)
  const T* f[MAX_FIELDS];  // current prognostic (MODE_STEP update base)
  T* out[MAX_FIELDS];      // MODE_ACCUM: source field; MODE_STEP: next state
  T dt[MAX_FIELDS];        // ultflx dt of the field
  T* zf[MAX_FIELDS];       // MODE_ACCUM: flux_z copy of the field (tvd_dgs_terms), or null
  const T* z[MAX_FIELDS];  // MODE_LEAP: current z-field (time level t-1)
  T* out2[MAX_FIELDS];     // MODE_LEAP: next z-field
  T gal[MAX_FIELDS];       // MODE_LEAP: galilean shift of the field (ugal, vgal; 0 otherwise)
  unsigned char* rec[MAX_FIELDS];  // MODE_REC: limiter arms of the field's faces
};

template <typename T>
struct LaunchArgs {
  Dims d;
  const T* u;
  const T* v;
  const T* w;
  VMet mc;  // scalars, u, v
  VMet mw;  // w (swapped metric roles)
  T dtm, cx, cy;
  // MODE_LEAP: Robert filter coefficient tsmth, c1 = 1 - 2 tsmth, 2 dtm (all
  // formed in FP64), and whether this is the forward first step
  T tsmth, c1, dtm2;
  int forward;
};

// stepfields.F90:597-627 (with perform_timesmooth_for_field, :522-533, for
// the scalars when centred) followed by swap_smooth (swapsmooth.F90) for one
// point: the current field fc and z-field zc, the tendency s (0 where the
// sources stay clear) -> the next field fn and z-field zn.
//   forward (first step):  e = gal + fc (u, v) or fc; zs = e + dtm s;
//                          classic swap: x = e + zs, fn = x*1 + e*(-1), zn = x - fn;
//   centred, u and v:      sav = zc + gal, e = gal + fc, zs = gal + zc + 2dtm s,
//                          fn = zs, zn = c1 e + tsmth (zs + sav - gal);
//   centred, w:            sav = zc, zs = zc + 2dtm s, fn = zs, zn = c1 fc + tsmth (zs + sav);
//   centred, scalars:      sm = c1 fc + tsmth zc, zs = zc + 2dtm s, fn = zs, zn = sm + tsmth zs.
template <typename T, int KIND>
__device__ __forceinline__ void leap_update(const LaunchArgs<T>& L, T gal, T fc, T zc, T s, T& fn, T& zn) {
  constexpr bool FLOW = KIND == KIND_U || KIND == KIND_V;
  if (L.forward) {
    const T e = FLOW ? gal + fc : fc;
    const T zs = e + L.dtm * s;
    const T x = e + zs;
    fn = x - e;
    zn = x - fn;
  } else if constexpr (KIND == KIND_SCALAR) {
    const T sm = L.c1 * fc + L.tsmth * zc;
    const T zs = zc + L.dtm2 * s;
    fn = zs;
    zn = sm + L.tsmth * zs;
  } else if constexpr (KIND == KIND_W) {
    const T zs = zc + L.dtm2 * s;
    fn = zs;
    zn = L.c1 * fc + L.tsmth * (zs + zc);
  } else {
    const T sav = zc + gal;
    const T e = gal + fc;
    const T zs = (gal + zc) + L.dtm2 * s;
    fn = zs;
    zn = L.c1 * e + L.tsmth * ((zs + sav) - gal);
  }
}

// Wrapped index of a halo position for a depth-2 periodic halo:
// 1 -> n+1, 2 -> n+2, n+3 -> 3, n+4 -> 4 (halocommunication.F90:1121-1168).
__device__ __forceinline__ int wrap(int x, int n) {
  return x <= 2 ? x + n : (x >= n + 3 ? x - n : x);
}

// Flows of a field kind: raw for scalars, dual-interpolated along the primal
// dimension of u (x), v (y) or w (z) for momentum.
template <typename T, int KIND>
struct Flows {
  static constexpr int DK = KIND == KIND_W ? 1 : 0;
  static constexpr int DJ = KIND == KIND_V ? 1 : 0;
  static constexpr int DI = KIND == KIND_U ? 1 : 0;
  using F = Flow<T, DK, DJ, DI>;
  F U, V, W;
  __device__ __forceinline__ Flows(const LaunchArgs<T>& L) : U(L.u, L.d), V(L.v, L.d), W(L.w, L.d) {}
};

// Flux-divergence weights of one cell (tvdadvection.F90:925-959), shared by
// the fields advected by the same flows.
template <typename T>
struct DivW {
  T vs, vn, wb, wt, t1, t2, ul, ur;
  template <class FL>
  __device__ __forceinline__ DivW(const FL& fl, const VMet& m, int k, int j, int i, int nz) {
    vs = fl.V(k, j - 1, i);
    vn = fl.V(k, j, i);
    wb = fl.W(k - 1, j, i);
    ul = fl.U(k, j, i - 1);
    ur = fl.U(k, j, i);
    t1 = vm<T>(m.t1, k);
    const bool below_top = k < nz;
    wt = below_top ? fl.W(k, j, i) : T(0);
    t2 = below_top ? vm<T>(m.t2, k) : T(0);
  }
  // s += three terms in the reference's order; `top` = level nz of u, v and
  // scalars, which is overwritten rather than accumulated (:942-959)
  __device__ __forceinline__ T apply(bool top, T s0, T cx, T cy, T fxl, T fxr, T fys, T fyn, T fzb,
                                     T fzt) const {
    T s;
    if (top) {
      s = (vs * fys - vn * fyn) * cy;
      s = s + T(4.0) * t1 * wb * fzb;
      s = s + (ul * fxl - ur * fxr) * cx;
    } else {
      s = s0 + (vs * fys - vn * fyn) * cy;
      s = s + T(4.0) * (wb * fzb * t1 - wt * fzt * t2);
      s = s + (ul * fxl - ur * fxr) * cx;
    }
    return s;
  }
};

// MODE_STEP output: out = f + dtm*s at the cell and at every halo position
// that is its periodic image (ny, nx >= 4, so each image is unique).
template <typename T>
__device__ __forceinline__ void store_step(const Dims& d, T* o, const T* f, T dtm, T s, int k, int j, int i) {
  const int64_t p = at(d, k, j, i);
  const T val = f[p] + dtm * s;
  const int jh = (j <= 4) ? j + d.ny : (j >= d.ny + 1 ? j - d.ny : 0);
  const int ih = (i <= 4) ? i + d.nx : (i >= d.nx + 1 ? i - d.nx : 0);
  o[p] = val;
  if (jh) o[at(d, k, jh, i)] = val;
  if (ih) o[at(d, k, j, ih)] = val;
  if (jh && ih) o[at(d, k, jh, ih)] = val;
}

// ---------------------------------------------------------------------------
// Cell kernel: one thread per cell, the six bounding faces recomputed per
// cell.  Lanes along z (coalesced 64-level runs), 4 rows of y per block, one x
// column per block row.  Used for columns taller than the slab kernel allows.
template <typename T, int KIND, int NS, int MODE>
__device__ __forceinline__ void cell_group(const LaunchArgs<T>& L, const FieldSet<T>& S, int n0, int k, int j,
                                           int i) {
  const Flows<T, KIND> fl(L);
  const VMet m = KIND == KIND_W ? L.mw : L.mc;
  const int nz = L.d.nz;
  const int kdof = KIND == KIND_W ? 1 : 0;
  const int kmin = KIND == KIND_W ? 1 : 2;
  const int klast = KIND == KIND_W ? nz - 1 : nz;
  const T dt = S.dt[n0];
  Adv<T, NS> F{};
#pragma unroll
  for (int n = 0; n < NS; n++) F.a[n] = Arr<T>(S.F[n0 + n], L.d);
  const bool cell = k >= 2 && k <= klast;
  T fxl[NS], fxr[NS], fys[NS], fyn[NS], fzb[NS], fzt[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) fxl[n] = fxr[n] = fys[n] = fyn[n] = fzb[n] = fzt[n] = T(0);
  const bool top = KIND != KIND_W && k == nz;
  if (cell) {
    face_x(k, j, i - 1, F, fl.U, fl.V, fl.W, m, kdof, dt, L.cx, L.cy, nz, fxl);
    face_x(k, j, i, F, fl.U, fl.V, fl.W, m, kdof, dt, L.cx, L.cy, nz, fxr);
    face_y(k, j - 1, i, F, fl.U, fl.V, fl.W, m, kdof, dt, L.cx, L.cy, nz, fys);
    face_y(k, j, i, F, fl.U, fl.V, fl.W, m, kdof, dt, L.cx, L.cy, nz, fyn);
    face_z(k - 1, j, i, F, fl.U, fl.V, fl.W, m, kdof, kmin, dt, L.cx, L.cy, nz, fzb);
    if (!top) face_z(k, j, i, F, fl.U, fl.V, fl.W, m, kdof, kmin, dt, L.cx, L.cy, nz, fzt);
  }
  const int64_t p = at(L.d, k, j, i);
  if (MODE == MODE_ACCUM && S.zf[n0]) {  // flux_z(k): the face below level k
    if (cell) {
      S.zf[n0][p] = fzb[0];
    } else if (KIND == KIND_W && k == nz) {
      T fz[NS];
      face_z(k - 1, j, i, F, fl.U, fl.V, fl.W, m, kdof, kmin, dt, L.cx, L.cy, nz, fz);
      S.zf[n0][p] = fz[0];
    }
  }
  if (cell) {  // the weights read flows at k-1: only for cells (k >= 2)
    const DivW<T> dw(fl, m, k, j, i, nz);
#pragma unroll
    for (int n = 0; n < NS; n++) {
      if (MODE == MODE_ACCUM)
        S.out[n0 + n][p] = dw.apply(top, S.out[n0 + n][p], L.cx, L.cy, fxl[n], fxr[n], fys[n], fyn[n], fzb[n],
                                    fzt[n]);
      else
        store_step(L.d, S.out[n0 + n], S.f[n0 + n], L.dtm,
                   dw.apply(top, T(0), L.cx, L.cy, fxl[n], fxr[n], fys[n], fyn[n], fzb[n], fzt[n]), k, j, i);
    }
  } else if (MODE == MODE_STEP) {
#pragma unroll
    for (int n = 0; n < NS; n++) store_step(L.d, S.out[n0 + n], S.f[n0 + n], L.dtm, T(0), k, j, i);
  }
}

// MODE_ACCUM: levels 2..nz (2..nz-1 for w), out += tendency, level nz
// overwritten.  MODE_STEP: levels 1..nz, out = f + dtm * tendency (0 where the
// reference leaves s untouched: k = 1, and k = nz for w), halo images too.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) cell_kernel(LaunchArgs<T> L, FieldSet<T> S) {
  const int k0 = MODE == MODE_STEP ? 1 : 2;
  const int k = k0 + blockIdx.x * 64 + threadIdx.x;
  const int j = 3 + blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.y);  // wave-uniform row
  const int i = 3 + blockIdx.z;
  if (j > L.d.ny + 2 || k > L.d.nz) return;
  for (int n = 0; n < S.nf; n++) {
    switch (S.kind[n]) {
      case KIND_U: cell_group<T, KIND_U, 1, MODE>(L, S, n, k, j, i); break;
      case KIND_V: cell_group<T, KIND_V, 1, MODE>(L, S, n, k, j, i); break;
      case KIND_W: cell_group<T, KIND_W, 1, MODE>(L, S, n, k, j, i); break;
      default: cell_group<T, KIND_SCALAR, 1, MODE>(L, S, n, k, j, i); break;
    }
  }
}

// ---------------------------------------------------------------------------
// Periodic fill of the depth-2 x/y halo of one field, corners included.  One
// thread per (k, halo column); the halo columns are the 4 rows j in {1,2,ny+3,
// ny+4} over every i, then the 4 columns i in {1,2,nx+3,nx+4} over interior j.
// Block (bx, by) of the fill: 64 levels x 4 halo columns.
template <typename T>
__device__ __forceinline__ void halo_block(T* f, const Dims& d, int bx, int64_t by) {
  const int k = 1 + bx * 64 + (int)threadIdx.x;
  const int64_t h = by * 4 + threadIdx.y;
  const int nx4 = d.nx + 4;
  const int64_t nrow = 4 * (int64_t)nx4;
  const int64_t nh = nrow + 4 * (int64_t)d.ny;
  if (k > d.nz || h >= nh) return;
  int j, i;
  if (h < nrow) {
    const int r = (int)(h / nx4);
    i = 1 + (int)(h % nx4);
    j = r < 2 ? 1 + r : d.ny + 1 + r;
  } else {
    const int64_t q = h - nrow;
    const int c = (int)(q / d.ny);
    j = 3 + (int)(q % d.ny);
    i = c < 2 ? 1 + c : d.nx + 1 + c;
  }
  f[at(d, k, j, i)] = f[at(d, k, wrap(j, d.ny), wrap(i, d.nx))];
}

template <typename T>
__global__ void __launch_bounds__(256) halo_kernel(T* f, Dims d) {
  halo_block<T>(f, d, (int)blockIdx.x, (int64_t)blockIdx.y);
}

// FP64 weighted sum over the interior, sum_k w(k) sum_{j,i} f(k,j,i), in two
// stages whose order of additions is fixed by the grid alone (the result does
// not depend on scheduling): WSUM_BLOCKS blocks of 64 levels x 4 rows each take
// the column groups (4 rows x 1 column) b, b + WSUM_BLOCKS, ... and write one
// partial; one block of 256 threads then adds the partials.  Scratch: nz
// weights, then WSUM_BLOCKS partials, then the result.
constexpr int WSUM_BLOCKS = 1024;

__device__ __forceinline__ double block_sum_256(double acc, double* red) {
  const int t = (int)(threadIdx.x + blockDim.x * threadIdx.y);
  red[t] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  return red[0];
}

template <typename T>
__global__ void __launch_bounds__(256) wsum_partial_kernel(const T* f, Dims d, const double* w, double* partial) {
  __shared__ double red[256];
  const int groups_y = (d.ny + 3) / 4;
  const int64_t ngroups = (int64_t)groups_y * d.nx;
  double acc = 0.0;
  for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const int i = 3 + (int)(g / groups_y);
    const int j = 3 + (int)(g % groups_y) * 4 + (int)threadIdx.y;
    if (j <= d.ny + 2)
      for (int k = 1 + (int)threadIdx.x; k <= d.nz; k += 64) acc += w[k - 1] * (double)f[at(d, k, j, i)];
  }
  const double s = block_sum_256(acc, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.x] = s;
}

template <typename P>
__global__ void __launch_bounds__(256) wsum_final_kernel(const P* partial, int n, P* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int q = (int)threadIdx.x; q < n; q += 256) acc += partial[q];
  const double s = block_sum_256(acc, red);
  if (threadIdx.x == 0) out[0] = s;
}

}  // namespace tvd
