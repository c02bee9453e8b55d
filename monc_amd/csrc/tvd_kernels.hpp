// Kernels of the TVD advection path: periodic halo fill, the per-cell tendency
// (drop-in `s += ...` form) and the fused forward step.
#pragma once
#include "tvd_faces.hpp"

namespace tvd {

enum FieldKind { KIND_U = 0, KIND_V = 1, KIND_W = 2, KIND_SCALAR = 3 };
enum Mode { MODE_ACCUM = 0, MODE_STEP = 1 };

template <typename T>
struct CellArgs {
  Dims d;
  const T* u;       // flows (current fields)
  const T* v;
  const T* w;
  const T* F;       // advected field (field or z-field)
  const T* f;       // current field (MODE_STEP: base of the update)
  T* out;           // MODE_ACCUM: s-field (read-modify-write); MODE_STEP: new field
  VMet m;
  T dt;             // ultflx dt (dtm or 2*dtm)
  T dtm;            // update step (MODE_STEP)
  T cx, cy;
};

// Wrapped index of a halo position for a depth-2 periodic halo:
// 1 -> n+1, 2 -> n+2, n+3 -> 3, n+4 -> 4 (halocommunication.F90:1121-1168).
__device__ __forceinline__ int wrap(int x, int n) {
  return x <= 2 ? x + n : (x >= n + 3 ? x - n : x);
}

// The tendency of one interior cell (k,j,i) of one field: the six faces that
// bound it and the three-term divergence of tvdadvection.F90:925-959.
template <typename T, int KIND>
__device__ __forceinline__ T cell_tendency(const CellArgs<T>& a, int k, int j, int i, T s0) {
  constexpr int DK = KIND == KIND_W ? 1 : 0;
  constexpr int DJ = KIND == KIND_V ? 1 : 0;
  constexpr int DI = KIND == KIND_U ? 1 : 0;
  const Flow<T, DK, DJ, DI> U{a.u, a.d}, V{a.v, a.d}, W{a.w, a.d};
  const Adv<T> F{a.F, a.d};
  const int nz = a.d.nz;
  const int kdof = KIND == KIND_W ? 1 : 0;
  const int kmin = KIND == KIND_W ? 1 : 2;
  T fxl = face_x(k, j, i - 1, F, U, V, W, a.m, kdof, a.dt, a.cx, a.cy, nz);
  T fxr = face_x(k, j, i, F, U, V, W, a.m, kdof, a.dt, a.cx, a.cy, nz);
  T fys = face_y(k, j - 1, i, F, U, V, W, a.m, kdof, a.dt, a.cx, a.cy, nz);
  T fyn = face_y(k, j, i, F, U, V, W, a.m, kdof, a.dt, a.cx, a.cy, nz);
  T fzb = face_z(k - 1, j, i, F, U, V, W, a.m, kdof, kmin, a.dt, a.cx, a.cy, nz);
  T s;
  if (KIND != KIND_W && k == nz) {
    // top row is overwritten, not accumulated (:942-959)
    s = (V(k, j - 1, i) * fys - V(k, j, i) * fyn) * a.cy;
    s = s + T(4.0) * vm<T>(a.m.t1, k) * W(k - 1, j, i) * fzb;
    s = s + (U(k, j, i - 1) * fxl - U(k, j, i) * fxr) * a.cx;
  } else {
    T fzt = face_z(k, j, i, F, U, V, W, a.m, kdof, kmin, a.dt, a.cx, a.cy, nz);
    s = s0 + (V(k, j - 1, i) * fys - V(k, j, i) * fyn) * a.cy;
    s = s + T(4.0) * (W(k - 1, j, i) * fzb * vm<T>(a.m.t1, k) - W(k, j, i) * fzt * vm<T>(a.m.t2, k));
    s = s + (U(k, j, i - 1) * fxl - U(k, j, i) * fxr) * a.cx;
  }
  return s;
}

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
};

template <typename T, int KIND, int MODE>
__device__ __forceinline__ void do_field(const LaunchArgs<T>& L, const FieldSet<T>& S, int n, int k, int j, int i) {
  CellArgs<T> a;
  a.d = L.d;
  a.u = L.u;
  a.v = L.v;
  a.w = L.w;
  a.F = S.F[n];
  a.f = S.f[n];
  a.out = S.out[n];
  a.m = KIND == KIND_W ? L.mw : L.mc;
  a.dt = S.dt[n];
  a.dtm = L.dtm;
  a.cx = L.cx;
  a.cy = L.cy;
  const int nz = L.d.nz;
  const int klast = KIND == KIND_W ? nz - 1 : nz;
  if (MODE == MODE_ACCUM) {
    if (k < 2 || k > klast) return;
    const int64_t p = at(a.d, k, j, i);
    a.out[p] = cell_tendency<T, KIND>(a, k, j, i, a.out[p]);
  } else {
    T s = T(0.0);
    if (k >= 2 && k <= klast) s = cell_tendency<T, KIND>(a, k, j, i, T(0.0));
    const T val = a.f[at(a.d, k, j, i)] + a.dtm * s;
    const int ny = a.d.ny, nx = a.d.nx;
    // the cell itself and every halo position that maps onto it (ny, nx >= 4)
    const int jh = (j <= 4) ? j + ny : (j >= ny + 1 ? j - ny : 0);
    const int ih = (i <= 4) ? i + nx : (i >= nx + 1 ? i - nx : 0);
    a.out[at(a.d, k, j, i)] = val;
    if (jh) a.out[at(a.d, k, jh, i)] = val;
    if (ih) a.out[at(a.d, k, j, ih)] = val;
    if (jh && ih) a.out[at(a.d, k, jh, ih)] = val;
  }
}

// One thread per cell of the interior, every field of the set: lanes along z
// (coalesced 64-level runs), 4 rows of y per block, one x column per block row.
// MODE_ACCUM: levels 2..nz (2..nz-1 for w), out(k,j,i) += tendency, level nz
// overwritten (tvdadvection.F90:942-959).
// MODE_STEP: levels 1..nz, out = f + dtm * tendency with tendency 0 where the
// reference leaves s untouched (k = 1, and k = nz for w), and the periodic halo
// images of the updated cell are written too.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) cell_kernel(LaunchArgs<T> L, FieldSet<T> S) {
  const int k0 = MODE == MODE_STEP ? 1 : 2;
  const int k = k0 + blockIdx.x * 64 + threadIdx.x;
  const int j = 3 + blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.y);  // wave-uniform row
  const int i = 3 + blockIdx.z;
  if (j > L.d.ny + 2 || k > L.d.nz) return;
  for (int n = 0; n < S.nf; n++) {
    switch (S.kind[n]) {
      case KIND_U: do_field<T, KIND_U, MODE>(L, S, n, k, j, i); break;
      case KIND_V: do_field<T, KIND_V, MODE>(L, S, n, k, j, i); break;
      case KIND_W: do_field<T, KIND_W, MODE>(L, S, n, k, j, i); break;
      default: do_field<T, KIND_SCALAR, MODE>(L, S, n, k, j, i); break;
    }
  }
}

// ---------------------------------------------------------------------------
// Slab kernel: every face computed once.
//
// A block owns TY cell rows j0..j0+TY-1 over all levels (NCH chunks of 64
// levels, lanes along z) and marches along an x strip.  Each thread (level k,
// row j) computes per step i only the faces its cell owns: the right x face
// (kept in a register as the next step's left face), the north y face and the
// top z face.  The south y face comes from the row below through LDS (an extra
// face-only wave row j0-1 supplies it for the first row), the bottom z face
// from the lane below by a wave shuffle (across chunks through LDS).
template <typename T, int NCH, int TY>
struct SlabLds {
  T yn[2][TY + 1][64 * NCH];  // north y faces, double-buffered by step parity
  T zt[2][TY + 1][NCH];       // top z face of the last level of each chunk
};

template <typename T, int KIND, int MODE, int NCH, int TY>
__device__ __forceinline__ void slab_field(const LaunchArgs<T>& L, const FieldSet<T>& S, int n, int i_begin,
                                           int i_end, SlabLds<T, NCH, TY>& lds) {
  constexpr int DK = KIND == KIND_W ? 1 : 0;
  constexpr int DJ = KIND == KIND_V ? 1 : 0;
  constexpr int DI = KIND == KIND_U ? 1 : 0;
  const Flow<T, DK, DJ, DI> U{L.u, L.d}, V{L.v, L.d}, W{L.w, L.d};
  const Adv<T> F{S.F[n], L.d};
  const VMet m = KIND == KIND_W ? L.mw : L.mc;
  const int nz = L.d.nz, ny = L.d.ny, nx = L.d.nx;
  const int kdof = KIND == KIND_W ? 1 : 0;
  const int kmin = KIND == KIND_W ? 1 : 2;
  const int kmax = KIND == KIND_W ? nz - 1 : nz;  // last level with x/y faces and a cell
  const T dt = S.dt[n];
  // a wave is 64 levels of one (chunk, row): c and r are wave-uniform
  const int lane = threadIdx.x;
  const int c = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int r = __builtin_amdgcn_readfirstlane(threadIdx.z);
  const int k = 1 + 64 * c + lane;
  const int j = 3 + (int)blockIdx.x * TY - 1 + r;
  const bool row_ok = j <= ny + 2;
  const bool cell_row = r > 0 && row_ok;
  const bool hface = row_ok && k >= 2 && k <= kmax;  // level has x/y faces
  const bool cell = cell_row && k >= 2 && k <= kmax;
  const bool zface = cell_row && k <= nz - 1;
  T fxl = T(0);
  if (cell_row && hface) fxl = face_x(k, j, i_begin - 1, F, U, V, W, m, kdof, dt, L.cx, L.cy, nz);
  for (int i = i_begin; i < i_end; i++) {
    const int b = i & 1;
    T fyn = T(0), fxr = T(0), fzt = T(0);
    // the fences keep the scheduler from interleaving the three face
    // computations, which bounds the live registers to one face's worth
    if (hface) fyn = face_y(k, j, i, F, U, V, W, m, kdof, dt, L.cx, L.cy, nz);
    __builtin_amdgcn_sched_barrier(0);
    if (cell_row && hface) fxr = face_x(k, j, i, F, U, V, W, m, kdof, dt, L.cx, L.cy, nz);
    __builtin_amdgcn_sched_barrier(0);
    if (zface) fzt = face_z(k, j, i, F, U, V, W, m, kdof, kmin, dt, L.cx, L.cy, nz);
    __builtin_amdgcn_sched_barrier(0);
    lds.yn[b][r][64 * c + lane] = fyn;
    if (lane == 63) lds.zt[b][r][c] = fzt;
    __syncthreads();
    T fzb = __shfl_up(fzt, 1);
    if (lane == 0) fzb = c > 0 ? lds.zt[b][r][c - 1] : T(0);
    if (cell_row && k <= nz) {
      T s;
      const int64_t p = at(L.d, k, j, i);
      if (cell) {
        const T fys = lds.yn[b][r - 1][64 * c + lane];
        const T s0 = MODE == MODE_ACCUM ? S.out[n][p] : T(0);
        if (KIND != KIND_W && k == nz) {
          s = (V(k, j - 1, i) * fys - V(k, j, i) * fyn) * L.cy;
          s = s + T(4.0) * vm<T>(m.t1, k) * W(k - 1, j, i) * fzb;
          s = s + (U(k, j, i - 1) * fxl - U(k, j, i) * fxr) * L.cx;
        } else {
          s = s0 + (V(k, j - 1, i) * fys - V(k, j, i) * fyn) * L.cy;
          s = s + T(4.0) * (W(k - 1, j, i) * fzb * vm<T>(m.t1, k) - W(k, j, i) * fzt * vm<T>(m.t2, k));
          s = s + (U(k, j, i - 1) * fxl - U(k, j, i) * fxr) * L.cx;
        }
      } else {
        s = T(0);
      }
      if (MODE == MODE_ACCUM) {
        if (cell) S.out[n][p] = s;
      } else {
        const T val = S.f[n][p] + L.dtm * s;
        const int jh = (j <= 4) ? j + ny : (j >= ny + 1 ? j - ny : 0);
        const int ih = (i <= 4) ? i + nx : (i >= nx + 1 ? i - nx : 0);
        T* o = S.out[n];
        o[p] = val;
        if (jh) o[at(L.d, k, jh, i)] = val;
        if (ih) o[at(L.d, k, j, ih)] = val;
        if (jh && ih) o[at(L.d, k, jh, ih)] = val;
      }
    }
    fxl = fxr;
  }
  __syncthreads();  // the next field reuses the LDS buffers
}

template <typename T, int MODE, int NCH, int TY>
__global__ void __launch_bounds__(64 * NCH * (TY + 1)) slab_kernel(LaunchArgs<T> L, FieldSet<T> S, int strip) {
  __shared__ SlabLds<T, NCH, TY> lds;
  const int i_begin = 3 + (int)blockIdx.y * strip;
  const int i_end = min(i_begin + strip, L.d.nx + 3);
  for (int n = 0; n < S.nf; n++) {
    switch (S.kind[n]) {
      case KIND_U: slab_field<T, KIND_U, MODE, NCH, TY>(L, S, n, i_begin, i_end, lds); break;
      case KIND_V: slab_field<T, KIND_V, MODE, NCH, TY>(L, S, n, i_begin, i_end, lds); break;
      case KIND_W: slab_field<T, KIND_W, MODE, NCH, TY>(L, S, n, i_begin, i_end, lds); break;
      default: slab_field<T, KIND_SCALAR, MODE, NCH, TY>(L, S, n, i_begin, i_end, lds); break;
    }
  }
}

// Periodic fill of the depth-2 x/y halo of one field, corners included.  One
// thread per (k, halo column); the halo columns are the 4 rows j in {1,2,ny+3,
// ny+4} over every i, then the 4 columns i in {1,2,nx+3,nx+4} over interior j.
template <typename T>
__global__ void __launch_bounds__(256) halo_kernel(T* f, Dims d) {
  const int k = 1 + blockIdx.x * 64 + threadIdx.x;
  const int64_t h = (int64_t)blockIdx.y * 4 + threadIdx.y;
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

}  // namespace tvd
