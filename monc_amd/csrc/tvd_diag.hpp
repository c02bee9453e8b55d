// Tendency diagnostics of the TVD component (tvdadvection.F90:1064-1188):
// the 3-D tendency of one call, tend = s_after - s_before over the interior
// columns (or s_after*r(k) - s_before*r(k) for the tabs diagnostic), and its
// per-level total over the interior columns, accumulated in FP64.
#pragma once
#include "tvd_kernels.hpp"

namespace tvd {

// One thread per interior cell: lanes along z, 4 rows per block, one column
// per block row.  out has the reference's tend_3d layout (nz, ny, nx).
template <typename T>
__global__ void __launch_bounds__(256) tend3d_kernel(const T* after, const T* before, const double* r, T* out,
                                                     Dims d) {
  // products rounded separately, as the reference forms them (no fused
  // multiply-add: at an untouched level after == before and v must be 0)
#pragma clang fp contract(off)
  const int k = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 3 + blockIdx.y * 4 + threadIdx.y;
  const int i = 3 + blockIdx.z;
  if (k > d.nz || j > d.ny + 2) return;
  const int64_t p = at(d, k, j, i);
  T v;
  if (r) {
    const T rk = (T)r[k - 1];
    v = after[p] * rk - before[p] * rk;
  } else {
    v = after[p] - before[p];
  }
  out[(k - 1) + (int64_t)d.nz * ((j - 3) + (int64_t)d.ny * (i - 3))] = v;
}

// Per-level partial totals over one row: partial[(j-1)*nz + k-1] =
// sum_i t(k, j, i), FP64, for the (nz, ny, nx) tendency layout.
template <typename T>
__global__ void __launch_bounds__(64) profile_rows_kernel(const T* t, int nz, int ny, int nx, double* partial) {
  const int k = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (k > nz) return;
  const int64_t sj = nz, si = (int64_t)nz * ny;
  double acc = 0.0;
  for (int i = 1; i <= nx; i++) acc += (double)t[(k - 1) + sj * (j - 1) + si * (i - 1)];
  partial[(int64_t)(j - 1) * nz + (k - 1)] = acc;
}

// out[k-1] = sum_j partial[(j-1)*nz + k-1], in row order.
__global__ void __launch_bounds__(64) profile_sum_kernel(const double* partial, int nz, int ny, double* out) {
  const int k = 1 + blockIdx.x * 64 + threadIdx.x;
  if (k > nz) return;
  double acc = 0.0;
  for (int j = 0; j < ny; j++) acc += partial[(int64_t)j * nz + (k - 1)];
  out[k - 1] = acc;
}

}  // namespace tvd
