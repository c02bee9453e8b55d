// Batched ensembles of uneven-size TVD domains (the C ABI's monc_tvd_ensemble_*).
//
// Every member is one local domain of its own size (nz, ny, nx), spacing and
// metric vectors, advecting u, v, w, th and q with forward-Euler steps of its
// own dtm (monc_tvd_forward_steps, per member).  All members live in one
// device pool per state variant: member m's field f starts at element
// off[m] + f * fstride[m], each field in the reference's (z, y, x) layout with
// its depth-2 halo, every field 256-byte aligned.
//
// Forward step: ONE persistent launch for the whole ensemble.  The tile list
// holds (member, kernel kind, tile row, tile column) for every march tile of
// every member -- u, v, w and the (th, q) pair -- largest members first; the
// blocks of a grid sized to fill the GPU once take tiles from a device counter
// until the list is exhausted, and each tile is the march kernel's tile
// (tvd_march.hpp) on that member's fields.  The last block to finish resets
// the counter, so the launch can be replayed without a memset.  Two steps
// (state pool A -> B -> A) are captured in one hipGraph per starting pool and
// replayed.  The tile shape (waves per block, x columns per tile) is one of
// the compiled shapes, picked by an autotuner that times the whole step.
//
// Adjoint step, halo fill and the adjoint's periodic fold: persistent
// grid-stride launches over the blocks of every member (a prefix table of
// per-member block counts maps a global block to its member).
// max and min as compare and select in this translation unit (faster for the
// persistent step kernel; see tvd_faces.hpp)
#define TVD_MINMAX_SELECT 1
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/monc_tvd.h"
#include "tvd_adjoint.hpp"
#include "tvd_host.hpp"
#include "tvd_march.hpp"

using namespace tvd;

namespace {

constexpr int ENS_FIELDS = 5;    // u, v, w, th, q
constexpr int ENS_NKMAX = 256;   // tallest column: 4 sub-columns of 64 levels
constexpr int ENS_ALIGN = 64;    // field alignment in elements (256 B in FP32)
constexpr int ENS_WAVES[3] = {4, 8, 12};
constexpr int ENS_STRIPS[4] = {8, 16, 32, 64};
constexpr int ENS_ADJ_STRIP = 16;  // owner columns per march adjoint tile

// One member as the device sees it.
struct EnsMember {
  Dims d;
  long long off;      // element offset of the member's u in a pool
  long long fstride;  // elements from one field of the member to the next
  long long met;      // offset of the member's MET_N metric vectors
  double dtm, cx, cy;
  int nch;            // 64-level sub-columns
};

template <typename T>
__device__ __forceinline__ LaunchArgs<T> member_args(const EnsMember& M, const T* state, const double* met) {
  LaunchArgs<T> L;
  L.d = M.d;
  const T* b = state + M.off;
  L.u = b;
  L.v = b + M.fstride;
  L.w = b + 2 * M.fstride;
  L.mc = vmet_from(met + M.met, M.d.nz, false);
  L.mw = vmet_from(met + M.met, M.d.nz, true);
  L.dtm = (T)M.dtm;
  L.cx = (T)M.cx;
  L.cy = (T)M.cy;
  return L;
}

// Kinds of march tile: u, v, w, and the (th, q) pair.
enum { TILE_U = 0, TILE_V = 1, TILE_W = 2, TILE_SCALARS = 3 };

template <typename T, int KIND, int NS, int WAVES>
__device__ __forceinline__ void ens_tile(const EnsMember& M, const double* met, const T* in, T* out, int f0, int ty,
                                         int tx, int strip, MarchLds<T, 2, WAVES, ENS_NKMAX>& lds) {
  const LaunchArgs<T> L = member_args<T>(M, in, met);
  FieldSet<T> S;
  S.nf = NS;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    S.kind[n] = KIND;
    S.F[n] = in + M.off + (long long)(f0 + n) * M.fstride;
    S.f[n] = S.F[n];
    S.out[n] = out + M.off + (long long)(f0 + n) * M.fstride;
    S.dt[n] = (T)M.dtm;
    S.zf[n] = nullptr;
  }
  // one kernel holds every field kind, so its register budget is set by the
  // largest; in FP32 the scalar pair's windows fit within it, the registers
  // of vertical neighbours by wave shifts do not (measured slower)
  constexpr bool WIN = !(sizeof(T) == 8 && NS == 2), VD = false;
  march_tile<T, KIND, NS, MODE_STEP, WAVES, ENS_NKMAX, 2, WIN, VD>(L, S, M.nch, ty, tx, strip, lds);
}

// Waves per SIMD the compiler keeps registers for (as the single-domain
// march kernel): two or more blocks share a CU, so one block's loads overlap
// another's arithmetic across the per-column barrier.
template <int ES, int WAVES>
constexpr int ens_wpe() {
  return ES == 4 ? 5 : 3;
}

// The persistent forward step: tiles are taken from ctr[0] until the list is
// exhausted; ctr[1] counts finished blocks, and the last one resets both.
template <typename T, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(ens_wpe<(int)sizeof(T), WAVES>())))
ens_step_kernel(const EnsMember* mem, const double* met, const int4* tiles, int ntiles, unsigned* ctr, const T* in,
                T* out, int strip) {
  __shared__ MarchLds<T, 2, WAVES, ENS_NKMAX> lds;
  __shared__ int next;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  for (;;) {
    if (tid == 0) next = (int)atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(next);
    // every thread has read `next` before thread 0 writes it again: the tile
    // below passes at least two barriers, and a block that leaves the loop
    // never writes it again
    if (t >= ntiles) break;
    const int4 e = tiles[t];
    const EnsMember& M = mem[e.x];
    switch (e.y) {
      case TILE_U: ens_tile<T, KIND_U, 1, WAVES>(M, met, in, out, 0, e.z, e.w, strip, lds); break;
      case TILE_V: ens_tile<T, KIND_V, 1, WAVES>(M, met, in, out, 1, e.z, e.w, strip, lds); break;
      case TILE_W: ens_tile<T, KIND_W, 1, WAVES>(M, met, in, out, 2, e.z, e.w, strip, lds); break;
      default: ens_tile<T, KIND_SCALAR, 2, WAVES>(M, met, in, out, 3, e.z, e.w, strip, lds); break;
    }
  }
  if (tid == 0) {
    if (atomicAdd(&ctr[1], 1u) == gridDim.x - 1) {
      atomicExch(&ctr[0], 0u);
      atomicExch(&ctr[1], 0u);
    }
  }
}

// Member owning global block b: prefix[m] <= b < prefix[m+1].
__device__ __forceinline__ int member_of(const long long* prefix, int nmem, long long b) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

enum { OP_HALO = 0, OP_FOLD_ROWS = 1, OP_FOLD_COLS = 2, OP_ADJOINT = 3, OP_COUNT = 4 };

// Blocks of each grid-stride operation per member and field (256 threads:
// 64 levels x 4 of the second index).
__host__ __device__ inline long long op_blocks_per_field(int op, int nz, int ny, int nx, int* n2) {
  const int nbx = (nz + 63) / 64;
  int m2;
  switch (op) {
    case OP_HALO: m2 = (4 * (nx + 4) + 4 * ny + 3) / 4; break;
    case OP_FOLD_ROWS: m2 = nx + 4; break;
    case OP_FOLD_COLS: m2 = ny; break;
    default:  // one work item per march adjoint tile (64 levels x 4 owner rows x 16 owner columns)
      m2 = (ny + 1 + ADJ_ROWS - 1) / ADJ_ROWS * ((nx + 1 + ENS_ADJ_STRIP - 1) / ENS_ADJ_STRIP);
      break;
  }
  if (n2) *n2 = m2;
  return (long long)nbx * m2;
}

// Halo fill (OP_HALO) or fold (OP_FOLD_*) of every field of every member.
template <typename T, int OP>
__global__ void __launch_bounds__(256) ens_grid_kernel(const EnsMember* mem, const long long* prefix, int nmem, T* pool) {
  const long long total = prefix[nmem];
  for (long long b = blockIdx.x; b < total; b += gridDim.x) {
    const int m = member_of(prefix, nmem, b);
    const EnsMember& M = mem[m];
    // per-member counts fit in 32 bits: no 64-bit division per block
    const int per = (int)op_blocks_per_field(OP, M.d.nz, M.d.ny, M.d.nx, nullptr);
    int r = (int)(b - prefix[m]);
    const int f = r / per;
    r -= f * per;
    const int nbx = (M.d.nz + 63) / 64;
    const int by = r / nbx, bx = r - by * nbx;
    T* fp = pool + M.off + (long long)f * M.fstride;
    if constexpr (OP == OP_HALO) halo_block<T>(fp, M.d, bx, by);
    else if constexpr (OP == OP_FOLD_ROWS) fold_rows_block<T>(fp, M.d, bx, by);
    else fold_cols_block<T>(fp, M.d, bx, by);
  }
}

// Adjoint of one forward step of every member: one work item is one march
// adjoint tile (adjoint_tile: 64 levels x 4 owner rows x 16 owner columns,
// LDS accumulation) of one field of one member.
template <typename T>
__global__ void __launch_bounds__(64 * ADJ_ROWS) ens_adjoint_kernel(const EnsMember* mem, const double* met,
                                                                    const long long* prefix, int nmem,
                                                                    const T* state, const T* lin, T* lam) {
  __shared__ AdjLds<T, ADJ_ROWS> lds;
  const long long total = prefix[nmem];
  for (long long b = blockIdx.x; b < total; b += gridDim.x) {
    const int m = member_of(prefix, nmem, b);
    const EnsMember& M = mem[m];
    const int nbx = (M.d.nz + 63) / 64, nty = (M.d.ny + 1 + ADJ_ROWS - 1) / ADJ_ROWS;
    const int per = (int)op_blocks_per_field(OP_ADJOINT, M.d.nz, M.d.ny, M.d.nx, nullptr);
    int r = (int)(b - prefix[m]);
    const int f = r / per;
    r -= f * per;
    const int band = r % nbx;
    r /= nbx;
    const int ty = r % nty, tx = r / nty;
    const LaunchArgs<T> L = member_args<T>(M, state, met);
    AdjArgs<T, 1> A{};
    const long long fo = M.off + (long long)f * M.fstride;
    A.F[0] = state + fo;
    A.lin[0] = lin + fo;
    A.lf[0] = lam + fo;
    A.lu = lam + M.off;
    A.lv = lam + M.off + M.fstride;
    A.lw = lam + M.off + 2 * M.fstride;
    A.dtm = (T)M.dtm;
    switch (f) {
      case 0: adjoint_tile<T, KIND_U, ADJ_ROWS>(L, A, band, ty, tx, ENS_ADJ_STRIP, lds); break;
      case 1: adjoint_tile<T, KIND_V, ADJ_ROWS>(L, A, band, ty, tx, ENS_ADJ_STRIP, lds); break;
      case 2: adjoint_tile<T, KIND_W, ADJ_ROWS>(L, A, band, ty, tx, ENS_ADJ_STRIP, lds); break;
      default: adjoint_tile<T, KIND_SCALAR, ADJ_ROWS>(L, A, band, ty, tx, ENS_ADJ_STRIP, lds); break;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) to_prec_kernel(T* dst, const double* src, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t] = (T)src[t];
}
template <typename T>
__global__ void __launch_bounds__(256) to_double_kernel(double* dst, const T* src, long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) dst[t] = (double)src[t];
}

}  // namespace

struct monc_tvd_ensemble {
  int nmem, precision;
  size_t es;
  std::vector<EnsMember> mem;   // host copy of the member table
  std::vector<long long> cells;  // interior cells per member
  long long pool_elems;
  EnsMember* dmem;
  double* dmet;
  void* pool[2];                 // state ping-pong
  int cur;                       // pool holding the current state
  void* adj;                     // adjoint accumulators
  void* adj_in;                  // incoming adjoint of a step
  long long* dprefix[OP_COUNT];
  long long nblocks[OP_COUNT];
  hipStream_t stream;            // caller's stream (null: legacy default)
  hipStream_t cap;               // library stream used only for graph capture
  int waves, strip, use_graph;
  int4* dtiles;
  int ntiles;
  int grid;                      // blocks of the persistent step launch
  unsigned* dctr;
  hipGraphExec_t graph[2];       // two steps starting from pool 0 / pool 1
  int cus;
};

namespace {

using Ens = monc_tvd_ensemble;

int check_ens(Ens* e) {
  if (!e) return fail(MONC_TVD_ERR_ARG, "null ensemble");
  return MONC_TVD_OK;
}

bool shape_ok(const Ens* e, int waves) {
  for (const EnsMember& m : e->mem)
    if (m.nch > waves) return false;
  return true;
}

void drop_graphs(Ens* e) {
  for (int p = 0; p < 2; p++)
    if (e->graph[p]) {
      (void)hipGraphExecDestroy(e->graph[p]);
      e->graph[p] = nullptr;
    }
}

int step_grid(Ens* e, int waves);

// Tile list of a shape, largest members first.
int build_tiles(Ens* e, int waves, int strip) {
  std::vector<int> order(e->nmem);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [e](int a, int b) { return e->cells[a] > e->cells[b]; });
  std::vector<int4> t;
  for (int m : order) {
    const EnsMember& M = e->mem[m];
    const int R = waves / M.nch;
    const int ty_n = (M.d.ny + R - 1) / R, tx_n = (M.d.nx + strip - 1) / strip;
    for (int kind = 0; kind < 4; kind++)
      for (int tx = 0; tx < tx_n; tx++)
        for (int ty = 0; ty < ty_n; ty++) t.push_back(make_int4(m, kind, ty, tx));
  }
  if (e->dtiles) HIP_TRY(hipFree(e->dtiles));
  e->dtiles = nullptr;
  HIP_TRY(dev_alloc(&e->dtiles, t.size() * sizeof(int4)));
  HIP_TRY(hipMemcpy(e->dtiles, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
  e->ntiles = (int)t.size();
  e->waves = waves;
  e->strip = strip;
  drop_graphs(e);
  return step_grid(e, waves);
}

// Blocks of the persistent launch: as many as fit on the GPU at once.
template <typename T, int WAVES>
int step_occupancy(int* per_cu) {
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ens_step_kernel<T, WAVES>, 64 * WAVES, 0));
  return MONC_TVD_OK;
}

int step_grid(Ens* e, int waves) {
  int per_cu = 0, rc;
  if (e->precision == MONC_TVD_FP64)
    rc = waves == 4 ? step_occupancy<double, 4>(&per_cu) : waves == 8 ? step_occupancy<double, 8>(&per_cu)
                                                                      : step_occupancy<double, 12>(&per_cu);
  else
    rc = waves == 4 ? step_occupancy<float, 4>(&per_cu) : waves == 8 ? step_occupancy<float, 8>(&per_cu)
                                                                     : step_occupancy<float, 12>(&per_cu);
  if (rc) return rc;
  if (per_cu < 1) return fail(MONC_TVD_ERR_HIP, "the %d-wave ensemble step kernel does not fit on a CU", waves);
  e->grid = std::max(1, std::min(per_cu * e->cus, e->ntiles));
  return MONC_TVD_OK;
}

template <typename T, int WAVES>
int launch_step_w(Ens* e, hipStream_t s, const void* in, void* out) {
  hipLaunchKernelGGL((ens_step_kernel<T, WAVES>), dim3(e->grid), dim3(64, WAVES, 1), 0, s, e->dmem, e->dmet,
                     e->dtiles, e->ntiles, e->dctr, (const T*)in, (T*)out, e->strip);
  HIP_TRY(hipGetLastError());
  return MONC_TVD_OK;
}

template <typename T>
int launch_step(Ens* e, hipStream_t s, const void* in, void* out) {
  switch (e->waves) {
    case 4: return launch_step_w<T, 4>(e, s, in, out);
    case 8: return launch_step_w<T, 8>(e, s, in, out);
    default: return launch_step_w<T, 12>(e, s, in, out);
  }
}

int step_once(Ens* e, hipStream_t s, const void* in, void* out) {
  return e->precision == MONC_TVD_FP64 ? launch_step<double>(e, s, in, out) : launch_step<float>(e, s, in, out);
}

int capture_graphs(Ens* e) {
  for (int p = 0; p < 2; p++) {
    if (e->graph[p]) continue;
    HIP_TRY(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
    int rc = step_once(e, e->cap, e->pool[p], e->pool[1 - p]);
    if (!rc) rc = step_once(e, e->cap, e->pool[1 - p], e->pool[p]);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ce != hipSuccess) return fail(MONC_TVD_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&e->graph[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      e->graph[p] = nullptr;
      return fail(MONC_TVD_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
    }
  }
  return MONC_TVD_OK;
}

int grid_for(const Ens* e, long long blocks) {
  return (int)std::max(1LL, std::min(blocks, (long long)e->cus * 8));
}

template <typename T, int OP>
int launch_grid_op(Ens* e, void* pool) {
  hipLaunchKernelGGL((ens_grid_kernel<T, OP>), dim3(grid_for(e, e->nblocks[OP])), dim3(64, 4, 1), 0, e->stream,
                     e->dmem, e->dprefix[OP], e->nmem, (T*)pool);
  HIP_TRY(hipGetLastError());
  return MONC_TVD_OK;
}

template <typename T>
int halo_impl(Ens* e) {
  return launch_grid_op<T, OP_HALO>(e, e->pool[e->cur]);
}

template <typename T>
int fold_impl(Ens* e) {
  int rc = launch_grid_op<T, OP_FOLD_ROWS>(e, e->adj);
  return rc ? rc : launch_grid_op<T, OP_FOLD_COLS>(e, e->adj);
}

template <typename T>
int adjoint_impl(Ens* e, const void* state) {
  HIP_TRY(hipMemcpyAsync(e->adj_in, e->adj, (size_t)e->pool_elems * e->es, hipMemcpyDeviceToDevice, e->stream));
  hipLaunchKernelGGL(ens_adjoint_kernel<T>, dim3(grid_for(e, e->nblocks[OP_ADJOINT])), dim3(64, ADJ_ROWS, 1), 0,
                     e->stream,
                     e->dmem, e->dmet, e->dprefix[OP_ADJOINT], e->nmem, (const T*)(state ? state : e->pool[e->cur]),
                     (const T*)e->adj_in, (T*)e->adj);
  HIP_TRY(hipGetLastError());
  return fold_impl<T>(e);
}

int upload_members(Ens* e) {
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(e->dmem, e->mem.data(), e->mem.size() * sizeof(EnsMember), hipMemcpyHostToDevice));
  return MONC_TVD_OK;
}

int check_mfv(Ens* e, int member, int field, int variant) {
  if (member < 0 || member >= e->nmem) return fail(MONC_TVD_ERR_ARG, "member %d out of range [0,%d)", member, e->nmem);
  if (field < 0 || field >= ENS_FIELDS) return fail(MONC_TVD_ERR_ARG, "field id %d out of range [0,5)", field);
  if (variant != MONC_TVD_PROG && variant != MONC_TVD_ADJOINT)
    return fail(MONC_TVD_ERR_ARG, "ensemble variant %d is not PROG or ADJOINT", variant);
  return MONC_TVD_OK;
}

void* variant_pool(Ens* e, int variant) { return variant == MONC_TVD_ADJOINT ? e->adj : e->pool[e->cur]; }

}  // namespace

extern "C" {

int monc_tvd_ensemble_create(int n_members, const int* nz, const int* ny, const int* nx, const double* cx,
                             const double* cy, const double* metrics, int precision, monc_tvd_ensemble** out) {
  if (!out) return fail(MONC_TVD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (n_members < 1) return fail(MONC_TVD_ERR_ARG, "n_members %d < 1", n_members);
  if (!nz || !ny || !nx || !cx || !cy || !metrics) return fail(MONC_TVD_ERR_ARG, "null member array");
  if (precision != MONC_TVD_FP64 && precision != MONC_TVD_FP32)
    return fail(MONC_TVD_ERR_ARG, "unknown precision %d", precision);
  const size_t es = precision == MONC_TVD_FP64 ? 8 : 4;
  std::vector<EnsMember> mem(n_members);
  std::vector<long long> cells(n_members);
  long long off = 0, met = 0;
  for (int m = 0; m < n_members; m++) {
    if (nz[m] < 4 || ny[m] < 4 || nx[m] < 4)
      return fail(MONC_TVD_ERR_ARG, "member %d: grid %dx%dx%d too small (need nz, ny, nx >= 4)", m, nz[m], ny[m], nx[m]);
    if (nz[m] > ENS_NKMAX)
      return fail(MONC_TVD_ERR_ARG, "member %d: nz %d above the ensemble's %d levels", m, nz[m], ENS_NKMAX);
    if (!(cx[m] > 0.0) || !(cy[m] > 0.0)) return fail(MONC_TVD_ERR_ARG, "member %d: cx, cy must be positive", m);
    const long long n_elem = (long long)nz[m] * (ny[m] + 4) * (nx[m] + 4);
    if ((unsigned long long)n_elem * es >= (1ULL << 32) - (1ULL << 20))
      return fail(MONC_TVD_ERR_ARG, "member %d: a field must stay under 4 GiB", m);
    EnsMember& M = mem[m];
    M.d.nz = nz[m];
    M.d.ny = ny[m];
    M.d.nx = nx[m];
    M.d.sj = nz[m];
    M.d.si = (int64_t)nz[m] * (ny[m] + 4);
    M.fstride = (n_elem + ENS_ALIGN - 1) / ENS_ALIGN * ENS_ALIGN;
    M.off = off;
    M.met = met;
    M.dtm = 0.0;
    M.cx = cx[m];
    M.cy = cy[m];
    M.nch = (nz[m] + 63) / 64;
    off += ENS_FIELDS * M.fstride;
    met += MET_N * (long long)nz[m];
    cells[m] = (long long)nz[m] * ny[m] * nx[m];
  }
  Ens* e = new Ens();
  e->nmem = n_members;
  e->precision = precision;
  e->es = es;
  e->mem = mem;
  e->cells = cells;
  e->pool_elems = off;
  e->dmem = nullptr;
  e->dmet = nullptr;
  e->pool[0] = e->pool[1] = e->adj = e->adj_in = nullptr;
  e->cur = 0;
  for (int op = 0; op < OP_COUNT; op++) e->dprefix[op] = nullptr;
  e->stream = nullptr;
  e->cap = nullptr;
  e->dtiles = nullptr;
  e->ntiles = 0;
  e->grid = 1;
  e->dctr = nullptr;
  e->graph[0] = e->graph[1] = nullptr;
  e->use_graph = 1;
  auto bail = [e](int rc) {
    monc_tvd_ensemble_destroy(e);
    return rc;
  };
  int dev = 0;
  hipError_t he = hipGetDevice(&dev);
  if (he == hipSuccess) he = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking);
  const size_t bytes = (size_t)off * es;
  for (void** p : {&e->pool[0], &e->pool[1], &e->adj, &e->adj_in})
    if (he == hipSuccess) he = dev_alloc(p, bytes);
  for (void* p : {e->pool[0], e->pool[1], e->adj, e->adj_in})
    if (he == hipSuccess) he = hipMemset(p, 0, bytes);
  if (he == hipSuccess) he = dev_alloc(&e->dmem, sizeof(EnsMember) * n_members);
  if (he == hipSuccess) he = hipMemcpy(e->dmem, mem.data(), sizeof(EnsMember) * n_members, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = dev_alloc(&e->dmet, sizeof(double) * met);
  if (he == hipSuccess) he = hipMemcpy(e->dmet, metrics, sizeof(double) * met, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = dev_alloc(&e->dctr, 2 * sizeof(unsigned));
  if (he == hipSuccess) he = hipMemset(e->dctr, 0, 2 * sizeof(unsigned));
  for (int op = 0; op < OP_COUNT && he == hipSuccess; op++) {
    std::vector<long long> pre(n_members + 1, 0);
    for (int m = 0; m < n_members; m++)
      pre[m + 1] = pre[m] + ENS_FIELDS * op_blocks_per_field(op, nz[m], ny[m], nx[m], nullptr);
    e->nblocks[op] = pre[n_members];
    he = dev_alloc(&e->dprefix[op], sizeof(long long) * (n_members + 1));
    if (he == hipSuccess)
      he = hipMemcpy(e->dprefix[op], pre.data(), sizeof(long long) * (n_members + 1), hipMemcpyHostToDevice);
  }
  if (he != hipSuccess) return bail(fail(MONC_TVD_ERR_HIP, "ensemble set-up: %s", hipGetErrorString(he)));
  // default shape: 8 waves per block if every member's column fits, else 12
  const int waves = shape_ok(e, 8) ? 8 : 12;
  if (int rc = build_tiles(e, waves, 16)) return bail(rc);
  *out = e;
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_destroy(monc_tvd_ensemble* e) {
  if (!e) return MONC_TVD_OK;
  (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  for (void* p : {e->pool[0], e->pool[1], e->adj, e->adj_in, (void*)e->dmem, (void*)e->dmet, (void*)e->dtiles,
                  (void*)e->dctr})
    if (p) (void)hipFree(p);
  for (int op = 0; op < OP_COUNT; op++)
    if (e->dprefix[op]) (void)hipFree(e->dprefix[op]);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  delete e;
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_layout(monc_tvd_ensemble* e, int member, long long* offset, long long* field_stride) {
  if (int rc = check_ens(e)) return rc;
  if (member < -1 || member >= e->nmem) return fail(MONC_TVD_ERR_ARG, "member %d out of range", member);
  if (member == -1) {
    if (offset) *offset = e->pool_elems;
    if (field_stride) *field_stride = 0;
    return MONC_TVD_OK;
  }
  if (offset) *offset = e->mem[member].off;
  if (field_stride) *field_stride = e->mem[member].fstride;
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_set_field(monc_tvd_ensemble* e, int member, int field, int variant, const double* host) {
  if (int rc = check_ens(e)) return rc;
  if (int rc = check_mfv(e, member, field, variant)) return rc;
  if (!host) return fail(MONC_TVD_ERR_ARG, "null host pointer");
  const EnsMember& M = e->mem[member];
  const long long n = (long long)M.d.nz * (M.d.ny + 4) * (M.d.nx + 4);
  char* dst = (char*)variant_pool(e, variant) + (size_t)(M.off + field * M.fstride) * e->es;
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->precision == MONC_TVD_FP64) {
    HIP_TRY(hipMemcpy(dst, host, n * sizeof(double), hipMemcpyHostToDevice));
    return MONC_TVD_OK;
  }
  double* tmp = nullptr;
  HIP_TRY(dev_alloc(&tmp, n * sizeof(double)));
  hipError_t he = hipMemcpy(tmp, host, n * sizeof(double), hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    hipLaunchKernelGGL(to_prec_kernel<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, (float*)dst,
                       (const double*)tmp, n);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (he != hipSuccess) return fail(MONC_TVD_ERR_HIP, "set_field: %s", hipGetErrorString(he));
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_get_field(monc_tvd_ensemble* e, int member, int field, int variant, double* host) {
  if (int rc = check_ens(e)) return rc;
  if (int rc = check_mfv(e, member, field, variant)) return rc;
  if (!host) return fail(MONC_TVD_ERR_ARG, "null host pointer");
  const EnsMember& M = e->mem[member];
  const long long n = (long long)M.d.nz * (M.d.ny + 4) * (M.d.nx + 4);
  const char* src = (const char*)variant_pool(e, variant) + (size_t)(M.off + field * M.fstride) * e->es;
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (e->precision == MONC_TVD_FP64) {
    HIP_TRY(hipMemcpy(host, src, n * sizeof(double), hipMemcpyDeviceToHost));
    return MONC_TVD_OK;
  }
  double* tmp = nullptr;
  HIP_TRY(dev_alloc(&tmp, n * sizeof(double)));
  hipLaunchKernelGGL(to_double_kernel<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, tmp,
                     (const float*)src, n);
  hipError_t he = hipGetLastError();
  if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
  if (he == hipSuccess) he = hipMemcpy(host, tmp, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  if (he != hipSuccess) return fail(MONC_TVD_ERR_HIP, "get_field: %s", hipGetErrorString(he));
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_copy_pool(monc_tvd_ensemble* e, int variant, void* dst, const void* src) {
  if (int rc = check_ens(e)) return rc;
  if (variant != MONC_TVD_PROG && variant != MONC_TVD_ADJOINT)
    return fail(MONC_TVD_ERR_ARG, "ensemble variant %d is not PROG or ADJOINT", variant);
  if ((dst == nullptr) == (src == nullptr)) return fail(MONC_TVD_ERR_ARG, "give exactly one of dst and src");
  void* pool = variant_pool(e, variant);
  HIP_TRY(hipMemcpyAsync(dst ? dst : pool, dst ? pool : src, (size_t)e->pool_elems * e->es, hipMemcpyDeviceToDevice,
                         e->stream));
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_set_dt(monc_tvd_ensemble* e, const double* dtm) {
  if (int rc = check_ens(e)) return rc;
  if (!dtm) return fail(MONC_TVD_ERR_ARG, "null dtm array");
  for (int m = 0; m < e->nmem; m++) {
    if (!(dtm[m] == dtm[m])) return fail(MONC_TVD_ERR_ARG, "member %d: dtm is NaN", m);
    e->mem[m].dtm = dtm[m];
  }
  return upload_members(e);
}

int monc_tvd_ensemble_set_stream(monc_tvd_ensemble* e, void* hip_stream) {
  if (int rc = check_ens(e)) return rc;
  e->stream = (hipStream_t)hip_stream;
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_synchronize(monc_tvd_ensemble* e) {
  if (int rc = check_ens(e)) return rc;
  HIP_TRY(hipStreamSynchronize(e->stream));
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_halo_fill(monc_tvd_ensemble* e) {
  if (int rc = check_ens(e)) return rc;
  return e->precision == MONC_TVD_FP64 ? halo_impl<double>(e) : halo_impl<float>(e);
}

int monc_tvd_ensemble_set_tiling(monc_tvd_ensemble* e, int waves, int strip, int use_graph) {
  if (int rc = check_ens(e)) return rc;
  if (waves != 4 && waves != 8 && waves != 12) return fail(MONC_TVD_ERR_ARG, "waves %d is not 4, 8 or 12", waves);
  if (strip < 1 || strip > 4096) return fail(MONC_TVD_ERR_ARG, "strip %d out of range [1,4096]", strip);
  if (!shape_ok(e, waves)) return fail(MONC_TVD_ERR_ARG, "a member's column needs more than %d waves", waves);
  e->use_graph = use_graph ? 1 : 0;
  if (waves == e->waves && strip == e->strip) return MONC_TVD_OK;
  HIP_TRY(hipStreamSynchronize(e->stream));
  return build_tiles(e, waves, strip);
}

int monc_tvd_ensemble_forward_steps(monc_tvd_ensemble* e, int nsteps) {
  if (int rc = check_ens(e)) return rc;
  if (nsteps < 0) return fail(MONC_TVD_ERR_ARG, "nsteps %d < 0", nsteps);
  int s = 0;
  if (e->use_graph && nsteps >= 2) {
    if (int rc = capture_graphs(e)) return rc;
    for (; s + 2 <= nsteps; s += 2) HIP_TRY(hipGraphLaunch(e->graph[e->cur], e->stream));
  }
  for (; s < nsteps; s++) {
    if (int rc = step_once(e, e->stream, e->pool[e->cur], e->pool[1 - e->cur])) return rc;
    e->cur ^= 1;
  }
  return MONC_TVD_OK;
}

int monc_tvd_ensemble_adjoint_step(monc_tvd_ensemble* e, const void* state) {
  if (int rc = check_ens(e)) return rc;
  return e->precision == MONC_TVD_FP64 ? adjoint_impl<double>(e, state) : adjoint_impl<float>(e, state);
}

int monc_tvd_ensemble_adjoint_fold(monc_tvd_ensemble* e) {
  if (int rc = check_ens(e)) return rc;
  return e->precision == MONC_TVD_FP64 ? fold_impl<double>(e) : fold_impl<float>(e);
}

int monc_tvd_ensemble_autotune(monc_tvd_ensemble* e, int reps, int* chosen) {
  if (int rc = check_ens(e)) return rc;
  if (reps < 1 || reps > 100) return fail(MONC_TVD_ERR_ARG, "reps %d out of range [1,100]", reps);
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  float best = 1e30f;
  int bw = e->waves, bs = e->strip;
  int rc = MONC_TVD_OK;
  for (int w : ENS_WAVES) {
    if (!shape_ok(e, w)) continue;
    for (int st : ENS_STRIPS) {
      if ((rc = build_tiles(e, w, st))) break;
      float ms = 0.f;
      hipError_t he = hipSuccess;
      // one untimed launch, then reps timed ones, all from the current state
      // into the spare pool: the state is not advanced
      for (int it = 0; it <= reps && !rc; it++) {
        if (it == 1) he = hipEventRecord(a, e->stream);
        rc = step_once(e, e->stream, e->pool[e->cur], e->pool[1 - e->cur]);
      }
      if (rc) break;
      if (he == hipSuccess) he = hipEventRecord(b, e->stream);
      if (he == hipSuccess) he = hipEventSynchronize(b);
      if (he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
      if (he != hipSuccess) {
        rc = fail(MONC_TVD_ERR_HIP, "autotune timing: %s", hipGetErrorString(he));
        break;
      }
      if (ms < best) {
        best = ms;
        bw = w;
        bs = st;
      }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  if ((rc = build_tiles(e, bw, bs))) return rc;
  if (chosen) {
    chosen[0] = bw;
    chosen[1] = bs;
  }
  return MONC_TVD_OK;
}

}  // extern "C"
