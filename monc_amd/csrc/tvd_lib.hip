// libmonc_tvd: the C ABI of include/monc_tvd.h over the gfx950 kernels of
// tvd_kernels.hpp.  Host code owns the device buffers, the metric vectors and
// the launch shapes; every entry point validates its arguments and reports
// failures through a status code and a thread-local message.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/monc_tvd.h"
#include "tvd_adjoint.hpp"
#include "tvd_diag.hpp"
#include "tvd_host.hpp"
#include "tvd_kernels.hpp"
#include "tvd_march.hpp"
#include "tvd_stage.hpp"

using namespace tvd;

namespace {
thread_local std::string g_err;
}  // namespace

namespace tvd {
int g_fail_allocs = 0;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace tvd

namespace {

constexpr int NVAR = 5;  // PROG, ZFIELD, SOURCE, ADJOINT, TANGENT

// Tangent-linear support: pack (value, tangent) pairs and split them again.
__global__ void __launch_bounds__(256) pack_dual_kernel(Dual* out, const double* v, const double* dv, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) out[t] = Dual(v[t], dv[t]);
}
__global__ void __launch_bounds__(256) unpack_dual_kernel(double* v, double* dv, const Dual* in, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) {
    const Dual x = in[t];
    v[t] = x.v;
    dv[t] = x.d;
  }
}


}  // namespace

struct monc_tvd_ctx {
  int nz, ny, nx, n_q, n_tracers, advect, precision;
  std::vector<char> active;    // per field: the reference's %active (th and q fields are gated by it)
  double cx, cy;
  size_t n_elem;     // elements per field, nz*(ny+4)*(nx+4)
  size_t elem_size;  // 8 or 4
  double* dmet;      // MET_N vectors of length nz, device
  int nfields;
  std::vector<void*> buf;      // [field*NVAR + variant]
  std::vector<char> owned;     // library-allocated?
  std::vector<void*> bound;    // caller pointer of a bound PROG buffer (restored after stepping)
  std::vector<void*> bound_z;  // caller pointer of a bound ZFIELD buffer (restored after stepping)
  std::vector<void*> spare;    // library-owned second buffer of each prognostic (stepping)
  std::vector<void*> spare_z;  // library-owned second buffer of each z-field (leapfrog stepping)
  std::vector<void*> adj_tmp;  // copy of the incoming adjoint of each field (adjoint step)
  std::vector<const unsigned char*> rec;  // caller buffer of the limiter arms the adjoint step reads, per field
  std::vector<void*> dual[2];  // packed (value, tangent) state, double-buffered (tangent steps)
  hipStream_t stream;
  int kernel;                  // MONC_TVD_KERNEL_MARCH, _CELL or _STAGE
  int rows[4], strip[4];       // tile shape per kind (u, v, w, scalars): cell rows and x columns
                               // per block (0: defaults)
  int kind_kernel[4];          // per kind: MONC_TVD_KERNEL_MARCH or _STAGE (forward step)
  int buffer_ok;               // fields small enough for 32-bit buffer offsets
  // optional per-kernel timing: HIP events recorded around every launch on
  // the context stream, folded into per-kind totals on query
  int profiling;
  struct Timed { int kind; hipEvent_t a, b; };
  std::vector<Timed> pending;
  double prof_ms[MONC_TVD_NKERNEL_KINDS];
  long long prof_n[MONC_TVD_NKERNEL_KINDS];
  // tendency diagnostics per component (TVD, pw_advection, th_advection)
  // and field: s before the call (shared), tend_3d (nz,ny,nx), per-level
  // totals (FP64), flux_z copies (TVD); tabs from th and rprefrcp
  int diag_flags;
  double* rprefrcp;            // device, nz (or null)
  std::vector<void*> diag_before, diag_zf;
  std::vector<void*> diag_3d[MONC_TVD_NCOMPONENTS];
  std::vector<double*> diag_prof[MONC_TVD_NCOMPONENTS];
  std::vector<char> diag_have[MONC_TVD_NCOMPONENTS];  // tend_3d/profile recorded
  void* tabs_3d[MONC_TVD_NCOMPONENTS];
  double* tabs_prof[MONC_TVD_NCOMPONENTS];
  int tabs_have[MONC_TVD_NCOMPONENTS];
  double* diag_partial;        // ny*nz FP64 scratch of the level totals
  double* thadv_tab;           // device, 2*nz: dthref, thref of th_advection (allocated on first use)
  double* wsum_scratch;        // device, weights + partials + result of monc_tvd_weighted_sum (first use)
};

namespace {

// Event pair around a launch when profiling is on.
struct LaunchTimer {
  monc_tvd_ctx* c;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  LaunchTimer(monc_tvd_ctx* c_, int kind_) : c(c_), kind(kind_) {
    if (!c->profiling) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(a, c->stream);
  }
  ~LaunchTimer() {
    if (!c->profiling || !a || !b) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({kind, a, b});
  }
};

int check_ctx(monc_tvd_ctx* c) {
  if (!c) return fail(MONC_TVD_ERR_ARG, "null context");
  return MONC_TVD_OK;
}

int check_fv(monc_tvd_ctx* c, int field, int variant) {
  if (field < 0 || field >= c->nfields) return fail(MONC_TVD_ERR_ARG, "field id %d out of range [0,%d)", field, c->nfields);
  if (variant < 0 || variant >= NVAR) return fail(MONC_TVD_ERR_ARG, "variant %d out of range", variant);
  return MONC_TVD_OK;
}

void*& slot(monc_tvd_ctx* c, int field, int variant) { return c->buf[(size_t)field * NVAR + variant]; }

int ensure(monc_tvd_ctx* c, int field, int variant) {
  void*& p = slot(c, field, variant);
  if (p) return MONC_TVD_OK;
  HIP_TRY(dev_alloc(&p, c->n_elem * c->elem_size));
  HIP_TRY(hipMemsetAsync(p, 0, c->n_elem * c->elem_size, c->stream));
  c->owned[(size_t)field * NVAR + variant] = 1;
  return MONC_TVD_OK;
}

Dims dims_of(const monc_tvd_ctx* c) {
  Dims d;
  d.nz = c->nz;
  d.ny = c->ny;
  d.nx = c->nx;
  d.sj = c->nz;
  d.si = (int64_t)c->nz * (c->ny + 4);
  return d;
}

VMet vmet_of(const monc_tvd_ctx* c, bool is_w) { return vmet_from(c->dmet, c->nz, is_w); }

template <typename T>
LaunchArgs<T> launch_args(monc_tvd_ctx* c, double dtm) {
  LaunchArgs<T> L;
  L.d = dims_of(c);
  L.u = (const T*)slot(c, MONC_TVD_U, MONC_TVD_PROG);
  L.v = (const T*)slot(c, MONC_TVD_V, MONC_TVD_PROG);
  L.w = (const T*)slot(c, MONC_TVD_W, MONC_TVD_PROG);
  L.mc = vmet_of(c, false);
  L.mw = vmet_of(c, true);
  L.dtm = (T)dtm;
  L.cx = (T)c->cx;
  L.cy = (T)c->cy;
  L.tsmth = T(0);
  L.c1 = T(1);
  L.dtm2 = (T)(2.0 * dtm);
  L.forward = 0;
  return L;
}

int kind_of(int field) {
  return field == MONC_TVD_U ? KIND_U : field == MONC_TVD_V ? KIND_V : field == MONC_TVD_W ? KIND_W : KIND_SCALAR;
}

// March kernel (kernel 0): NCH waves per row, R cell rows per block, on a 1-D
// grid of tiles_y * tiles_x tiles (the kernel deals tiles to XCDs).  WPE, the
// waves per SIMD the compiler keeps registers for, lets two or more blocks
// share a CU, so that one block's loads overlap another's arithmetic across
// the per-column barrier.
template <int ES, int NCH, int R>
constexpr int march_wpe() {
  return ES == 4 ? (NCH * R >= 16 ? 8 : 6) : (NCH * R >= 16 ? 4 : 3);
}

template <typename T, int KIND, int NS, int MODE, int NCH, int R>
void launch_march(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& S, int strip) {
  constexpr int WPE = pair_wide<T, NS, NCH, R>() ? 2 : march_wpe<(int)sizeof(T), NCH, R>();
  const int tiles_y = (c->ny + R - 1) / R, tiles_x = (c->nx + strip - 1) / strip;
  dim3 block(64, NCH * R, 1);
  hipLaunchKernelGGL((march_kernel<T, KIND, NS, MODE, NCH, R, WPE>), dim3(tiles_y * tiles_x), block, 0, c->stream,
                     L, S, strip, tiles_y);
}

// Cell rows per block of the march kernel compiled for each column height
// (NCH sub-columns of 64 levels); the first entry is the default.
constexpr int MARCH_ROWS[4][4] = {{12, 8, 4, 0}, {6, 4, 2, 0}, {4, 2, 0, 0}, {3, 2, 4, 1}};

// Default shape per precision and column height, from timings of the forward
// step on MI355X (64, 128 and 256 levels over 256^2 and 512^2 columns): rows
// per block, and x columns per block chosen so that a launch has at least
// 1024 blocks (4 per CU) when the domain allows it.
int default_rows(int precision, int nch, int kind) {
  static const int fp32[4] = {8, 4, 4, 3}, fp64[4] = {8, 2, 2, 3};
  if (precision == MONC_TVD_FP64 && kind == 3 && nch == 4) return 2;  // the FP64 scalars' wide shape
  return (precision == MONC_TVD_FP64 ? fp64 : fp32)[nch - 1];
}

// Small domains take narrower strips until the launch has 256 blocks (one
// per CU): at 64^3 the step went from 0.28 to 0.055 ms (FP64) and 0.17 to
// 0.036 ms (FP32) in strips of 2 columns instead of 16.
int default_strip(int ny, int nx, int rows, bool wide) {
  const long tiles_y = (ny + rows - 1) / rows;
  if (wide && tiles_y * ((nx + 63) / 64) >= 1024) return 64;
  if (tiles_y * ((nx + 31) / 32) >= 1024) return 32;
  int strip = 16;
  while (strip > 1 && tiles_y * ((nx + strip - 1) / strip) < 256) strip /= 2;
  return strip;
}

bool march_rows_ok(int nch, int rows) {
  if (nch < 1 || nch > 4) return false;
  for (int r : MARCH_ROWS[nch - 1])
    if (r && r == rows) return true;
  return false;
}

template <typename T, int KIND, int NS, int MODE>
void launch_shape(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& S, int nch, int kernel, int rows,
                  int strip) {
  if constexpr (MODE == MODE_LEAP || MODE == MODE_REC) {  // compiled for the default shapes only
    switch (nch * 100 + rows) {
      case 108: launch_march<T, KIND, NS, MODE, 1, 8>(c, L, S, strip); break;
      case 204: launch_march<T, KIND, NS, MODE, 2, 4>(c, L, S, strip); break;
      case 202: launch_march<T, KIND, NS, MODE, 2, 2>(c, L, S, strip); break;
      case 304: launch_march<T, KIND, NS, MODE, 3, 4>(c, L, S, strip); break;
      case 302: launch_march<T, KIND, NS, MODE, 3, 2>(c, L, S, strip); break;
      default: launch_march<T, KIND, NS, MODE, 4, 3>(c, L, S, strip); break;
    }
  } else {
    switch (nch * 100 + rows) {
      case 112: launch_march<T, KIND, NS, MODE, 1, 12>(c, L, S, strip); break;
      case 108: launch_march<T, KIND, NS, MODE, 1, 8>(c, L, S, strip); break;
      case 104: launch_march<T, KIND, NS, MODE, 1, 4>(c, L, S, strip); break;
      case 206: launch_march<T, KIND, NS, MODE, 2, 6>(c, L, S, strip); break;
      case 204: launch_march<T, KIND, NS, MODE, 2, 4>(c, L, S, strip); break;
      case 202: launch_march<T, KIND, NS, MODE, 2, 2>(c, L, S, strip); break;
      case 304: launch_march<T, KIND, NS, MODE, 3, 4>(c, L, S, strip); break;
      case 302: launch_march<T, KIND, NS, MODE, 3, 2>(c, L, S, strip); break;
      case 404: launch_march<T, KIND, NS, MODE, 4, 4>(c, L, S, strip); break;
      case 402: launch_march<T, KIND, NS, MODE, 4, 2>(c, L, S, strip); break;
      case 403: launch_march<T, KIND, NS, MODE, 4, 3>(c, L, S, strip); break;
      case 401: launch_march<T, KIND, NS, MODE, 4, 1>(c, L, S, strip); break;
    }
  }
}

// Launches over a field set: per field kind (u, v, w, scalars in pairs) the
// slab kernel for columns of up to 256 levels; the one-cell-per-thread kernel
// (all fields in one launch) beyond that or when forced.
// Stage kernel (MONC_TVD_KERNEL_STAGE): the step of one field kind with u, v,
// w (and the kind's scalars) staged in LDS, over 64-level bands x R rows x x
// strips.  The register budget (waves per SIMD the compiler plans for) is set
// by the kind's demand, not by the LDS: the FP32 flow kinds fit 80 registers
// (6 waves, two 11-row blocks per CU; w takes 90 unbounded and one block),
// the FP32 scalars 128 (4 waves), FP64 168 (3 waves: the 12 waves of an
// 11-row tile) and the FP64 scalar pair in tiles of up to 8 rows 256 (2
// waves); a smaller budget spills.
template <typename T, int KIND, int NS, int R>
constexpr int stage_wpe() {
  if (sizeof(T) == 4) return KIND == KIND_SCALAR ? 4 : 6;
  return KIND == KIND_SCALAR && NS == 2 && R < 11 ? 2 : (R == 15 ? 4 : 3);
}

template <typename T, int KIND, int NS, int R, int MODE>
void launch_stage(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& G, int strip) {
  const int nbands = (c->nz + 63) / 64, tiles_y = (c->ny + R - 1) / R, tiles_x = (c->nx + strip - 1) / strip;
  hipLaunchKernelGGL((stage_kernel<T, KIND, NS, R, stage_wpe<T, KIND, NS, R>(), MODE>),
                     dim3(nbands * tiles_y * tiles_x), dim3(64, R + 1, 1), 0, c->stream, L, G, strip, nbands,
                     tiles_y);
}

// Rows per block of the stage kernel (15 only for the flow kinds: the FP64
// scalar pair's ring of 16 planes of 19 rows would not fit the LDS).  In
// FP64, 15 rows is 16 waves, 4 per SIMD, within 128 registers and 139 KB of
// LDS; at 512x512x256 it runs u in 1.54 ms per launch in strips of 32 columns
// (2240 blocks), 1.63 in strips of 64 (1120 blocks, a short last round of
// one block per CU) and 1.92 in strips of 128, against 1.55 for 11 rows in
// strips of 128; the autotuner tries them all.
constexpr int STAGE_ROWS[5] = {4, 6, 8, 11, 15};

bool stage_shape_compiled(int precision, int kind, int rows) {
  (void)precision;
  return rows != 15 || kind != KIND_SCALAR;
}

// Default kernel per field kind for the forward step, from timings at
// 512x512x256 on MI355X (the autotuner's pick there):
//   * FP64: u, v and w staged in 11-row tiles of 128 columns (1.51-1.58 ms
//     per launch against 2.0-2.1 ms marching) and the (th, q) pair in 11-row
//     tiles of 128 columns (2.24 ms: 149 KB of LDS with v and w two planes
//     deep, 12 waves at 160 registers; 2.53-2.56 in 8-row tiles of 9 waves,
//     2.71 for the march kernel's 2 x 64 shape);
//   * FP32: u, v and w staged in 11-row tiles of 32 columns (0.75-0.77 ms
//     per launch against 0.88-0.94 marching; at most 80 registers and 55 KB
//     of LDS, two blocks per CU); the (th, q) pair marches (staged, 1.44
//     against 1.42 ms).
// Small domains keep the march kernel: the staged form runs one or two blocks
// per CU, so it needs at least 512 blocks to fill the chip; in FP64 the strip
// is the longest of 128, 64 and 32 columns that still gives them.  Returns the
// strip, or 0 for the march kernel.
int stage_default_rows(int precision, int kind) {
  (void)precision;
  (void)kind;
  return 11;
}

int stage_default_strip(const monc_tvd_ctx* c, int kind) {
  const bool fp64 = c->precision == MONC_TVD_FP64;
  if (!fp64 && kind == KIND_SCALAR) return 0;
  const int rows = stage_default_rows(c->precision, kind);
  const long tiles = (long)((c->nz + 63) / 64) * ((c->ny + rows - 1) / rows);
  if (!fp64) return tiles * ((c->nx + 31) / 32) >= 512 ? 32 : 0;
  for (int strip : {128, 64, 32})
    if (tiles * ((c->nx + strip - 1) / strip) >= 512) return strip;
  return 0;
}

bool stage_rows_ok(int rows) { return rows == 4 || rows == 6 || rows == 8 || rows == 11 || rows == 15; }

// A field group the stage kernel takes: a flow field, or one or two scalars
// with the same timestep, fields under 4 GiB; MODE_STEP and MODE_REC with
// each field stepped from the field it advects (a flow advecting itself),
// MODE_ACCUM without z-face diagnostics, MODE_LEAP (in both, a flow kind may
// advect another field, its z-field in a centred step).
template <typename T, int MODE>
bool stage_fits(const monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& G) {
  if (!c->buffer_ok || G.nf < 1 || G.nf > 2) return false;
  for (int n = 0; n < G.nf; n++) {
    if (G.dt[n] != G.dt[0] || G.kind[n] != G.kind[0]) return false;
    if ((MODE == MODE_STEP || MODE == MODE_REC) && G.f[n] != G.F[n]) return false;
    if (MODE == MODE_ACCUM && G.zf[n]) return false;
  }
  switch (G.kind[0]) {
    case KIND_U: return G.nf == 1 && (G.F[0] == L.u || MODE == MODE_ACCUM || MODE == MODE_LEAP);
    case KIND_V: return G.nf == 1 && (G.F[0] == L.v || MODE == MODE_ACCUM || MODE == MODE_LEAP);
    case KIND_W: return G.nf == 1 && (G.F[0] == L.w || MODE == MODE_ACCUM || MODE == MODE_LEAP);
    default: return true;
  }
}

template <typename T, int KIND, int NS, int MODE>
void launch_stage_rows(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& G, int rows, int strip) {
  switch (rows) {
    case 6: launch_stage<T, KIND, NS, 6, MODE>(c, L, G, strip); break;
    case 8: launch_stage<T, KIND, NS, 8, MODE>(c, L, G, strip); break;
    case 11: launch_stage<T, KIND, NS, 11, MODE>(c, L, G, strip); break;
    case 15:
      if constexpr (KIND != KIND_SCALAR && NS == 0) {
        launch_stage<T, KIND, NS, 15, MODE>(c, L, G, strip);
        break;
      }
      [[fallthrough]];
    default: launch_stage<T, KIND, NS, 4, MODE>(c, L, G, strip); break;
  }
}

// A flow kind advecting its own staged flow (NS = 0), or, in the drop-in call
// and the leapfrog step, another field (its z-field in a centred step),
// staged as a fourth array (NS = 1; at most 11 rows: 16 planes of 19 rows
// would not fit the LDS in FP64).
template <typename T, int KIND, int MODE>
void launch_stage_flow(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& G, const T* flow, int rows,
                       int strip) {
  if (G.F[0] == flow) launch_stage_rows<T, KIND, 0, MODE>(c, L, G, rows, strip);
  else if constexpr (MODE == MODE_ACCUM || MODE == MODE_LEAP) launch_stage_rows<T, KIND, 1, MODE>(c, L, G, rows, strip);
}

template <typename T, int MODE = MODE_STEP>
void launch_stage_group(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& G, int kk, int rows,
                        int strip) {
  switch (G.kind[0]) {
    case KIND_U: launch_stage_flow<T, KIND_U, MODE>(c, L, G, L.u, rows, strip); break;
    case KIND_V: launch_stage_flow<T, KIND_V, MODE>(c, L, G, L.v, rows, strip); break;
    case KIND_W: launch_stage_flow<T, KIND_W, MODE>(c, L, G, L.w, rows, strip); break;
    default:
      if (G.nf == 2) launch_stage_rows<T, KIND_SCALAR, 2, MODE>(c, L, G, rows, strip);
      else launch_stage_rows<T, KIND_SCALAR, 1, MODE>(c, L, G, rows, strip);
  }
  (void)kk;
}

template <typename T, int MODE>
int launch_set(monc_tvd_ctx* c, const LaunchArgs<T>& L, const FieldSet<T>& S) {
  const int nch = (c->nz + 63) / 64;
  const bool march_fits = nch <= 4 && c->buffer_ok;
  if constexpr (MODE == MODE_LEAP) {
    if (!march_fits)
      return fail(MONC_TVD_ERR_ARG, "leapfrog stepping needs columns of at most 256 levels and fields under 4 GiB");
  } else if constexpr (MODE == MODE_REC) {
    if (!march_fits)
      return fail(MONC_TVD_ERR_ARG,
                  "recording limiter arms needs columns of at most 256 levels and fields under 4 GiB");
  } else if (c->kernel == MONC_TVD_KERNEL_CELL || !march_fits) {
    const int nk = MODE == MODE_STEP ? c->nz : c->nz - 1;
    dim3 block(64, 4, 1);
    dim3 grid((nk + 63) / 64, (c->ny + 3) / 4, c->nx);
    LaunchTimer timer(c, 5);
    hipLaunchKernelGGL((cell_kernel<T, MODE>), grid, block, 0, c->stream, L, S);
    HIP_TRY(hipGetLastError());
    return MONC_TVD_OK;
  }
  for (int n = 0; n < S.nf;) {
    FieldSet<T> G{};
    const int pair = S.kind[n] == KIND_SCALAR && n + 1 < S.nf && S.kind[n + 1] == KIND_SCALAR && S.dt[n + 1] == S.dt[n];
    G.nf = pair ? 2 : 1;
    for (int m = 0; m < G.nf; m++) {
      G.kind[m] = S.kind[n + m];
      G.F[m] = S.F[n + m];
      G.f[m] = S.f[n + m];
      G.out[m] = S.out[n + m];
      G.dt[m] = S.dt[n + m];
      G.zf[m] = S.zf[n + m];
      G.z[m] = S.z[n + m];
      G.out2[m] = S.out2[n + m];
      G.gal[m] = S.gal[n + m];
      G.rec[m] = S.rec[n + m];
    }
    const int kk = pair || S.kind[n] == KIND_SCALAR ? 3 : S.kind[n];
    // the recording step runs the stage kernel where the forward step does (so
    // that the two write the same values), else the march kernel in its
    // default shape; the leapfrog step too
    const bool leap = MODE == MODE_LEAP || MODE == MODE_REC;
    const int kernel = leap ? MONC_TVD_KERNEL_MARCH : c->kernel;
    // the leapfrog step keeps the general shape: its FP64 pair measured
    // slower in the wide one (6.26 against 6.42 Gcell/s)
    // a kind set to the stage kernel that does not fit it (a drop-in call
    // with z-face diagnostics) marches in the default shape: its rows and
    // strip are the stage's
    const bool own = !leap && c->kind_kernel[kk] != MONC_TVD_KERNEL_STAGE;
    const int rows = own && c->rows[kk] > 0 && march_rows_ok(nch, c->rows[kk])
                         ? c->rows[kk]
                         : default_rows(c->precision, nch, leap ? -1 : kk);
    const bool wide = c->precision == MONC_TVD_FP64 && kk == 3 && nch == 4 && rows <= 2;
    const int strip = own && c->strip[kk] > 0 ? c->strip[kk] : default_strip(c->ny, c->nx, rows, wide);
    LaunchTimer timer(c, kk);
    if constexpr (MODE == MODE_STEP || MODE == MODE_ACCUM || MODE == MODE_REC || MODE == MODE_LEAP) {
      const int dstrip = stage_default_strip(c, S.kind[n]);
      const bool dflt = c->kind_kernel[kk] == MONC_TVD_KERNEL_MARCH && !c->rows[kk] && !c->strip[kk] &&
                        c->kernel == MONC_TVD_KERNEL_MARCH && dstrip > 0;
      if ((c->kind_kernel[kk] == MONC_TVD_KERNEL_STAGE || dflt) && stage_fits<T, MODE>(c, L, G)) {
        const int srows = dflt ? stage_default_rows(c->precision, S.kind[n]) : (c->rows[kk] > 0 ? c->rows[kk] : 6);
        const int sstrip = dflt ? dstrip : (c->strip[kk] > 0 ? c->strip[kk] : 64);
        launch_stage_group<T, MODE>(c, L, G, kk, srows, sstrip);
        HIP_TRY(hipGetLastError());
        n += G.nf;
        continue;
      }
    }
    switch (S.kind[n]) {
      case KIND_U: launch_shape<T, KIND_U, 1, MODE>(c, L, G, nch, kernel, rows, strip); break;
      case KIND_V: launch_shape<T, KIND_V, 1, MODE>(c, L, G, nch, kernel, rows, strip); break;
      case KIND_W: launch_shape<T, KIND_W, 1, MODE>(c, L, G, nch, kernel, rows, strip); break;
      default:
        if (pair) launch_shape<T, KIND_SCALAR, 2, MODE>(c, L, G, nch, kernel, rows, strip);
        else launch_shape<T, KIND_SCALAR, 1, MODE>(c, L, G, nch, kernel, rows, strip);
    }
    HIP_TRY(hipGetLastError());
    n += G.nf;
  }
  return MONC_TVD_OK;
}

template <typename T>
int launch_halo(monc_tvd_ctx* c, T* f) {
  Dims d = dims_of(c);
  const int64_t nh = 4 * (int64_t)(c->nx + 4) + 4 * (int64_t)c->ny;
  dim3 block(64, 4, 1);
  dim3 grid((c->nz + 63) / 64, (unsigned)((nh + 3) / 4), 1);
  LaunchTimer timer(c, 4);
  hipLaunchKernelGGL((halo_kernel<T>), grid, block, 0, c->stream, f, d);
  HIP_TRY(hipGetLastError());
  return MONC_TVD_OK;
}

// Fields advected this call, in the reference's order u, v, w, th, q(:),
// tracers (tvdadvection.F90:591-594): th only when active (:706), each q field
// only when active (:661), the flow fields and tracers always (:603-646,
// :684-694).
std::vector<int> advected_fields(const monc_tvd_ctx* c) {
  std::vector<int> f;
  if (c->advect & MONC_TVD_ADVECT_FLOW) {
    f.push_back(MONC_TVD_U);
    f.push_back(MONC_TVD_V);
    f.push_back(MONC_TVD_W);
  }
  if ((c->advect & MONC_TVD_ADVECT_THETA) && c->active[MONC_TVD_TH]) f.push_back(MONC_TVD_TH);
  if (c->advect & MONC_TVD_ADVECT_Q)
    for (int n = 0; n < c->n_q; n++)
      if (c->active[MONC_TVD_Q(n)]) f.push_back(MONC_TVD_Q(n));
  if (c->advect & MONC_TVD_ADVECT_TRACER)
    for (int n = 0; n < c->n_tracers; n++) f.push_back(MONC_TVD_TRACER(c->n_q, n));
  return f;
}

// Tracers have no tendency diagnostics and no tvd_dgs_terms face copies in
// the reference (tvdadvection.F90:679-695, 1064-1188).
bool is_tracer(const monc_tvd_ctx* c, int field) { return field >= 4 + c->n_q; }

// th_advection (thadvection.F90:124-211): at every column the registry hands
// it -- halo columns included, as the component does not skip them -- and
// levels k = 2..nz, the vertical advection of the reference state
//   sth -= tzc1(k)*2*dthref(k-1)*w(k-1) + tzc2(k)*2*dthref(k)*w(k)
// (anelastic only, :165-179) and then the advection of the mean
// baroclinicity (:186-211)
//   anelastic:  sth += thref(k)*(fcoriol/G) * ((v+vgal)*rate_x - (u+ugal)*rate_y)
//   otherwise:  sth += (v+vgal)*mfx - (u+ugal)*mfy,  mf = rate*thref0*(fcoriol/G)
// with the factors of initialisation_callback (:60-77) formed on the host in
// FP64.  One thread per (level, column), z fastest; a wave covers levels
// 1..64 (65..128, ...) so that its loads and stores fill whole cache lines
// (level 1 idles).
template <typename T>
__global__ void __launch_bounds__(256) thadv_kernel(T* sth, const T* u, const T* v, const T* w, const double* tzc1,
                                                    const double* tzc2, const double* dthref, const double* thref,
                                                    Dims d, int anelastic, int baro, double fcog, double rate_x,
                                                    double rate_y, double mfx, double mfy, T ugal, T vgal) {
  const int k = 1 + (int)blockIdx.x * 64 + (int)threadIdx.x;
  const int col = (int)blockIdx.y * 4 + (int)threadIdx.y;  // < 2^31, checked on the host
  if (k < 2 || k > d.nz || col >= (d.ny + 4) * (d.nx + 4)) return;
  const int jb = d.ny + 4, i0 = col / jb;
  const int j = 1 + (col - i0 * jb), i = 1 + i0;
  const int64_t o = at(d, k, j, i);
  T s = sth[o];
  if (anelastic) {
    const T s1 = (T)(tzc1[k - 1] * 2.0 * dthref[k - 2]), s2 = (T)(tzc2[k - 1] * 2.0 * dthref[k - 1]);
    s = s - (s1 * w[o - 1] + s2 * w[o]);
  }
  if (baro) {
    const T vv = v[o] + vgal, uu = u[o] + ugal;
    if (anelastic) s = s + (T)(thref[k - 1] * fcog) * (vv * (T)rate_x - uu * (T)rate_y);
    else s = s + (vv * (T)mfx - uu * (T)mfy);
  }
  sth[o] = s;
}

// Diagnostics around one call of a component (TVD, pw_advection or
// th_advection): snapshot the s-fields it changes before it
// (save_precomponent_tendencies), allocate the z-face buffers (TVD only).
int diag_before_call(monc_tvd_ctx* c, const std::vector<int>& fields, int comp) {
  if (!c->diag_flags) return MONC_TVD_OK;
  const size_t bytes = c->n_elem * c->elem_size;
  for (int fl : fields) {
    if (is_tracer(c, fl)) continue;
    if (c->diag_flags & MONC_TVD_DIAG_TENDENCY) {
      if (!c->diag_before[fl]) HIP_TRY(dev_alloc(&c->diag_before[fl], bytes));
      HIP_TRY(hipMemcpyAsync(c->diag_before[fl], slot(c, fl, MONC_TVD_SOURCE), bytes, hipMemcpyDeviceToDevice,
                             c->stream));
    }
    if (comp == MONC_TVD_COMPONENT_TVD && (c->diag_flags & MONC_TVD_DIAG_ZFACE) && !c->diag_zf[fl]) {
      HIP_TRY(dev_alloc(&c->diag_zf[fl], bytes));
      HIP_TRY(hipMemsetAsync(c->diag_zf[fl], 0, bytes, c->stream));
    }
  }
  return MONC_TVD_OK;
}

// tend_3d = after - before and its per-level totals (compute_component_tendencies).
template <typename T>
int tend_and_profile(monc_tvd_ctx* c, const T* after, const T* before, const double* r, T* t3d, double* prof) {
  const Dims d = dims_of(c);
  hipLaunchKernelGGL((tend3d_kernel<T>), dim3((c->nz + 63) / 64, (c->ny + 3) / 4, c->nx), dim3(64, 4, 1), 0,
                     c->stream, after, before, r, t3d, d);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((profile_rows_kernel<T>), dim3((c->nz + 63) / 64, c->ny, 1), dim3(64, 1, 1), 0, c->stream,
                     (const T*)t3d, c->nz, c->ny, c->nx, c->diag_partial);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(profile_sum_kernel, dim3((c->nz + 63) / 64, 1, 1), dim3(64, 1, 1), 0, c->stream,
                     (const double*)c->diag_partial, c->nz, c->ny, prof);
  HIP_TRY(hipGetLastError());
  return MONC_TVD_OK;
}

template <typename T>
int diag_after_call(monc_tvd_ctx* c, const std::vector<int>& fields, int comp) {
  if (!(c->diag_flags & MONC_TVD_DIAG_TENDENCY)) return MONC_TVD_OK;
  const size_t n3d = (size_t)c->nz * c->ny * c->nx;
  if (!c->diag_partial) HIP_TRY(dev_alloc(&c->diag_partial, sizeof(double) * (size_t)c->nz * c->ny));
  int rc;
  for (int fl : fields) {
    if (is_tracer(c, fl)) continue;
    void*& t3d = c->diag_3d[comp][fl];
    double*& prof = c->diag_prof[comp][fl];
    if (!t3d) HIP_TRY(dev_alloc(&t3d, n3d * sizeof(T)));
    if (!prof) HIP_TRY(dev_alloc(&prof, sizeof(double) * c->nz));
    if ((rc = tend_and_profile<T>(c, (const T*)slot(c, fl, MONC_TVD_SOURCE), (const T*)c->diag_before[fl], nullptr,
                                  (T*)t3d, prof)))
      return rc;
    c->diag_have[comp][fl] = 1;
    if (fl == MONC_TVD_TH && c->rprefrcp) {
      if (!c->tabs_3d[comp]) HIP_TRY(dev_alloc(&c->tabs_3d[comp], n3d * sizeof(T)));
      if (!c->tabs_prof[comp]) HIP_TRY(dev_alloc(&c->tabs_prof[comp], sizeof(double) * c->nz));
      if ((rc = tend_and_profile<T>(c, (const T*)slot(c, fl, MONC_TVD_SOURCE), (const T*)c->diag_before[fl],
                                    c->rprefrcp, (T*)c->tabs_3d[comp], c->tabs_prof[comp])))
        return rc;
      c->tabs_have[comp] = 1;
    }
  }
  return MONC_TVD_OK;
}

template <typename T>
int thadv_impl(monc_tvd_ctx* c, int anelastic, const double* dthref, int baro, const double* thref, double thref0,
               double fcoriol, double g, double rate_x, double rate_y, double ugal, double vgal) {
  const int nz = c->nz;
  if (!c->thadv_tab) HIP_TRY(dev_alloc(&c->thadv_tab, 2 * (size_t)nz * sizeof(double)));
  if (anelastic)
    HIP_TRY(hipMemcpyAsync(c->thadv_tab, dthref, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (baro && anelastic)
    HIP_TRY(hipMemcpyAsync(c->thadv_tab + nz, thref, nz * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const double fcog = fcoriol / g;
  const double mfx = rate_x * thref0 * fcog, mfy = rate_y * thref0 * fcog;
  const Dims d = dims_of(c);
  const long long ncol = (long long)(c->ny + 4) * (c->nx + 4);
  if (ncol > 0x7ffffff0LL) return fail(MONC_TVD_ERR_ARG, "too many columns for th_advection's grid");
  // th's tendency diagnostics when th is active (thadvection.F90:84-88, 152-159)
  std::vector<int> diag_fields;
  if (c->active[MONC_TVD_TH]) diag_fields.push_back(MONC_TVD_TH);
  int rc;
  if ((rc = diag_before_call(c, diag_fields, MONC_TVD_COMPONENT_THADV))) return rc;
  if (anelastic || baro) {
    dim3 block(64, 4, 1), grid((nz + 63) / 64, (unsigned)((ncol + 3) / 4), 1);
    LaunchTimer timer(c, 4);
    hipLaunchKernelGGL((thadv_kernel<T>), grid, block, 0, c->stream, (T*)slot(c, MONC_TVD_TH, MONC_TVD_SOURCE),
                       (const T*)slot(c, MONC_TVD_U, MONC_TVD_PROG), (const T*)slot(c, MONC_TVD_V, MONC_TVD_PROG),
                       (const T*)slot(c, MONC_TVD_W, MONC_TVD_PROG), c->dmet + MET_TZC1 * nz,
                       c->dmet + MET_TZC2 * nz, c->thadv_tab, c->thadv_tab + nz, d, anelastic, baro, fcog, rate_x,
                       rate_y, mfx, mfy, (T)ugal, (T)vgal);
    HIP_TRY(hipGetLastError());
  }
  return diag_after_call<T>(c, diag_fields, MONC_TVD_COMPONENT_THADV);
}

// pw_advection (pwadvection.F90:258-610): the centred second-order scheme,
// one thread per interior cell (halo columns are skipped, :307), levels
// 2..nz, all three directions active.  su, sv, sw and sth are overwritten
// (the reference assigns them), sq(n) accumulates; l_toplevel = .true.
// (:21): at k = nz su, sv and the scalars drop the W(nz) term and sw is left
// alone.  tcx = 0.25/dx, tcy = 0.25/dy; cx, cy = 1/dx, 1/dy.  Expressions
// keep the reference's operation order.
template <typename T>
struct PwArgs {
  const T* f[MAX_FIELDS];
  T* s[MAX_FIELDS];
  const double *tzc1, *tzc2, *tzd1, *tzd2;
  double cx, cy, tcx, tcy;
  int nf, flow, th, q;
};

// A block is 64 levels x 4 rows; each thread marches a strip of PW_STRIP x
// columns of its (level, row) and carries the flows two columns share in
// registers (u, v, w at i-1..i+1; v(j-1), w(k-1) at i..i+1; u(j+1), u(k+1)
// at i-1..i), so a column loads 17 flow values instead of ~33.  Waves cover
// levels 1..64, 65..128, ... so that accesses fill whole cache lines (level 1
// idles).  Every expression keeps the reference's operation order.
constexpr int PW_STRIP = 16;
template <typename T>
__global__ void __launch_bounds__(256) pw_kernel(PwArgs<T> P, Dims d) {
  const int k = 1 + (int)blockIdx.x * 64 + (int)threadIdx.x;
  const int j = 3 + (int)blockIdx.y * 4 + (int)threadIdx.y;
  if (k < 2 || k > d.nz || j > d.ny + 2) return;
  const int i0 = 3 + (int)blockIdx.z * PW_STRIP, i1 = min(i0 + PW_STRIP, d.nx + 3);
  const bool top = k == d.nz;
  const int kp = top ? k : k + 1;  // level k+1 (only read below the top; in bounds there too)
  const T *u = P.f[0], *v = P.f[1], *w = P.f[2];
  auto F = [&](const T* a, int kk, int jj, int ii) { return a[at(d, kk, jj, ii)]; };
  const T t1 = (T)P.tzc1[k - 1], t2 = (T)P.tzc2[k - 1];
  const T d1 = top ? T(0) : (T)P.tzd1[k - 1], d2 = top ? T(0) : (T)P.tzd2[k - 1];
  const T tcx = (T)P.tcx, tcy = (T)P.tcy;
  const T cxh = (T)(P.cx * 0.5), cyh = (T)(P.cy * 0.5);
  T u_m = F(u, k, j, i0 - 1), u_0 = F(u, k, j, i0);
  T v_m = F(v, k, j, i0 - 1), v_0 = F(v, k, j, i0);
  T w_m = F(w, k, j, i0 - 1), w_0 = F(w, k, j, i0);
  T vs_0 = F(v, k, j - 1, i0), wd_0 = F(w, k - 1, j, i0);
  T un_m = F(u, k, j + 1, i0 - 1), uk_m = F(u, kp, j, i0 - 1);
  for (int i = i0; i < i1; i++) {
    const T u_p = F(u, k, j, i + 1), v_p = F(v, k, j, i + 1), w_p = F(w, k, j, i + 1);
    const T vs_p = F(v, k, j - 1, i + 1), wd_p = F(w, k - 1, j, i + 1);
    const T un_0 = F(u, k, j + 1, i), uk_0 = F(u, kp, j, i);
    const int64_t o = at(d, k, j, i);
    if (P.flow) {
      const T us = F(u, k, j - 1, i), ud = F(u, k - 1, j, i);
      const T vn = F(v, k, j + 1, i), vd = F(v, k - 1, j, i), vk = F(v, kp, j, i), vsk = F(v, kp, j - 1, i);
      const T wn = F(w, k, j + 1, i), wnd = F(w, k - 1, j + 1, i), wk = F(w, kp, j, i), ws = F(w, k, j - 1, i);
      T a = tcx * (u_m * (u_0 + u_m) - u_p * (u_0 + u_p));
      a = a + tcy * (us * (vs_0 + vs_p) - un_0 * (v_0 + v_p));
      if (!top) a = a + (t1 * ud * (wd_0 + wd_p) - t2 * uk_0 * (w_0 + w_p));
      else a = a + t1 * ud * (wd_0 + wd_p);
      P.s[0][o] = a;
      T b = tcy * (vs_0 * (v_0 + vs_0) - vn * (v_0 + vn));
      b = b + tcx * (v_m * (u_m + un_m) - v_p * (u_0 + un_0));
      if (!top) b = b + (t1 * vd * (wd_0 + wnd) - t2 * vk * (w_0 + wn));
      else b = b + t1 * vd * (wd_0 + wnd);
      P.s[1][o] = b;
      if (!top) {
        T c = (d1 * wd_0 * (w_0 + wd_0) - d2 * wk * (w_0 + wk));
        c = c + tcx * (w_m * (u_m + uk_m) - w_p * (u_0 + uk_0));
        c = c + tcy * (ws * (vs_0 + vsk) - wn * (v_0 + vk));
        P.s[2][o] = c;
      }
    }
    // scalars: sth is assigned (the reference's own `sth = cx*...`, :407 and
    // :437), sq(n) accumulates
    for (int n = 3; n < P.nf; n++) {
      if (n == 3 ? !P.th : !P.q) continue;
      const T* f = P.f[n];
      T s = n == 3 ? T(0) : P.s[n][o];
      s = s + cxh * (u_m * F(f, k, j, i - 1) - u_0 * F(f, k, j, i + 1));
      s = s + cyh * (vs_0 * F(f, k, j - 1, i) - v_0 * F(f, k, j + 1, i));
      if (!top) s = s + T(2.0) * (t1 * wd_0 * F(f, k - 1, j, i) - t2 * w_0 * F(f, kp, j, i));
      else s = s + t1 * T(2.0) * wd_0 * F(f, k - 1, j, i);
      P.s[n][o] = s;
    }
    u_m = u_0;
    u_0 = u_p;
    v_m = v_0;
    v_0 = v_p;
    w_m = w_0;
    w_0 = w_p;
    vs_0 = vs_p;
    wd_0 = wd_p;
    un_m = un_0;
    uk_m = uk_0;
  }
}

template <typename T>
int pw_impl(monc_tvd_ctx* c, int flow, int th, int q, double tcx, double tcy) {
  PwArgs<T> P{};
  P.nf = 4 + c->n_q;  // pw_advection's q loop covers number_q_fields (pwadvection.F90:333), no tracers
  for (int n = 0; n < P.nf; n++) {
    P.f[n] = (const T*)slot(c, n, MONC_TVD_PROG);
    P.s[n] = (T*)slot(c, n, MONC_TVD_SOURCE);
  }
  P.tzc1 = c->dmet + MET_TZC1 * c->nz;
  P.tzc2 = c->dmet + MET_TZC2 * c->nz;
  P.tzd1 = c->dmet + MET_TZD1 * c->nz;
  P.tzd2 = c->dmet + MET_TZD2 * c->nz;
  P.cx = c->cx;
  P.cy = c->cy;
  P.tcx = tcx;
  P.tcy = tcy;
  P.flow = flow;
  P.th = th && c->active[MONC_TVD_TH];  // advect_th_field: only when th%active (pwadvection.F90:403)
  P.q = q;
  const Dims d = dims_of(c);
  // tendency diagnostics of the fields pw_advection changes
  // (pwadvection.F90:94-120, 310-318)
  std::vector<int> diag_fields;
  if (P.flow)
    for (int f = MONC_TVD_U; f <= MONC_TVD_W; f++) diag_fields.push_back(f);
  if (P.th) diag_fields.push_back(MONC_TVD_TH);
  if (P.q)
    for (int n = 0; n < c->n_q; n++) diag_fields.push_back(MONC_TVD_Q(n));
  int rc;
  if ((rc = diag_before_call(c, diag_fields, MONC_TVD_COMPONENT_PW))) return rc;
  {
    dim3 block(64, 4, 1), grid((c->nz + 63) / 64, (c->ny + 3) / 4, (c->nx + PW_STRIP - 1) / PW_STRIP);
    LaunchTimer timer(c, 4);
    hipLaunchKernelGGL((pw_kernel<T>), grid, block, 0, c->stream, P, d);
    HIP_TRY(hipGetLastError());
  }
  return diag_after_call<T>(c, diag_fields, MONC_TVD_COMPONENT_PW);
}

int need(monc_tvd_ctx* c, int field, int variant) {
  if (!slot(c, field, variant))
    return fail(MONC_TVD_ERR_STATE, "field %d variant %d is needed but was never set or bound", field, variant);
  return MONC_TVD_OK;
}

template <typename T>
int tendencies_impl(monc_tvd_ctx* c, double dtm, int field_stepping, int momentum_stepping, int scalar_stepping) {
  int rc;
  for (int fl : {MONC_TVD_U, MONC_TVD_V, MONC_TVD_W})
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
  const int adv_var = field_stepping == MONC_TVD_FORWARD_STEPPING ? MONC_TVD_PROG : MONC_TVD_ZFIELD;
  const double dt_mom = momentum_stepping == MONC_TVD_FORWARD_STEPPING ? dtm : dtm * 2.0;
  const double dt_sca = scalar_stepping == MONC_TVD_FORWARD_STEPPING ? dtm : dtm * 2.0;
  std::vector<int> fields = advected_fields(c);
  for (int fl : fields) {
    if ((rc = need(c, fl, adv_var)) || (rc = need(c, fl, MONC_TVD_SOURCE))) return rc;
  }
  if ((rc = diag_before_call(c, fields, MONC_TVD_COMPONENT_TVD))) return rc;
  const LaunchArgs<T> L = launch_args<T>(c, dtm);
  for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
    FieldSet<T> S{};
    S.nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
    for (int n = 0; n < S.nf; n++) {
      const int fl = fields[first + n];
      S.kind[n] = kind_of(fl);
      S.F[n] = (const T*)slot(c, fl, adv_var);
      S.f[n] = (const T*)slot(c, fl, MONC_TVD_PROG);
      S.out[n] = (T*)slot(c, fl, MONC_TVD_SOURCE);
      S.dt[n] = (T)(fl <= MONC_TVD_W ? dt_mom : dt_sca);
      S.zf[n] = (c->diag_flags & MONC_TVD_DIAG_ZFACE) && !is_tracer(c, fl) ? (T*)c->diag_zf[fl] : nullptr;
    }
    if ((rc = launch_set<T, MODE_ACCUM>(c, L, S))) return rc;
  }
  return diag_after_call<T>(c, fields, MONC_TVD_COMPONENT_TVD);
}

template <typename T>
int forward_steps_impl(monc_tvd_ctx* c, double dtm, int nsteps) {
  int rc;
  for (int fl : {MONC_TVD_U, MONC_TVD_V, MONC_TVD_W})
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
  std::vector<int> fields = advected_fields(c);
  for (int fl : fields) {
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
    if (!c->spare[fl]) HIP_TRY(dev_alloc(&c->spare[fl], c->n_elem * c->elem_size));
  }
  // the flows are read by every field's update, so no field may be updated in
  // place: each step writes the spare buffers, then the roles swap.
  for (int s = 0; s < nsteps; s++) {
    const LaunchArgs<T> L = launch_args<T>(c, dtm);
    for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
      FieldSet<T> S{};
      S.nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
      for (int n = 0; n < S.nf; n++) {
        const int fl = fields[first + n];
        S.kind[n] = kind_of(fl);
        S.F[n] = (const T*)slot(c, fl, MONC_TVD_PROG);
        S.f[n] = S.F[n];
        S.out[n] = (T*)c->spare[fl];
        S.dt[n] = (T)dtm;
      }
      if ((rc = launch_set<T, MODE_STEP>(c, L, S))) return rc;
    }
    for (int fl : fields) std::swap(slot(c, fl, MONC_TVD_PROG), c->spare[fl]);
  }
  // a caller-bound prognostic keeps holding the state after the call
  for (int fl : fields) {
    void* b = c->bound[fl];
    if (b && slot(c, fl, MONC_TVD_PROG) != b) {
      HIP_TRY(hipMemcpyAsync(b, slot(c, fl, MONC_TVD_PROG), c->n_elem * c->elem_size, hipMemcpyDeviceToDevice,
                             c->stream));
      std::swap(slot(c, fl, MONC_TVD_PROG), c->spare[fl]);
    }
  }
  return MONC_TVD_OK;
}

// One forward step from caller buffers to caller buffers (the recompute of
// a checkpointed rollout keeps every state without copies).
template <typename T>
int forward_into_impl(monc_tvd_ctx* c, double dtm, void* const* in, void* const* out) {
  std::vector<int> fields = advected_fields(c);
  for (int f = 0; f < c->nfields; f++)
    if (!in[f] || !out[f]) return fail(MONC_TVD_ERR_ARG, "null buffer for field %d", f);
  // every written buffer must be distinct from every buffer the step reads
  // (the flows u, v, w are read by every field's update) and from the others
  for (int fl : fields) {
    for (int g = 0; g < c->nfields; g++)
      if (out[fl] == in[g]) return fail(MONC_TVD_ERR_ARG, "output of field %d aliases the input of field %d", fl, g);
    for (int g : fields)
      if (g != fl && out[fl] == out[g]) return fail(MONC_TVD_ERR_ARG, "fields %d and %d share an output", fl, g);
  }
  LaunchArgs<T> L = launch_args<T>(c, dtm);
  L.u = (const T*)in[MONC_TVD_U];
  L.v = (const T*)in[MONC_TVD_V];
  L.w = (const T*)in[MONC_TVD_W];
  for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
    FieldSet<T> S{};
    S.nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
    for (int n = 0; n < S.nf; n++) {
      const int fl = fields[first + n];
      if (in[fl] == out[fl]) return fail(MONC_TVD_ERR_ARG, "field %d: the step cannot run in place", fl);
      S.kind[n] = kind_of(fl);
      S.F[n] = (const T*)in[fl];
      S.f[n] = S.F[n];
      S.out[n] = (T*)out[fl];
      S.dt[n] = (T)dtm;
    }
    if (int rc = launch_set<T, MODE_STEP>(c, L, S)) return rc;
  }
  return MONC_TVD_OK;
}

// forward_into_impl that also records every face's limiter arm into rec[f]
// (one byte per point of field f's layout) for the adjoint of the step.
template <typename T>
int forward_record_impl(monc_tvd_ctx* c, double dtm, void* const* in, void* const* out, unsigned char* const* rec) {
  std::vector<int> fields = advected_fields(c);
  for (int f = 0; f < c->nfields; f++)
    if (!in[f] || !out[f]) return fail(MONC_TVD_ERR_ARG, "null buffer for field %d", f);
  for (int fl : fields) {
    if (!rec[fl]) return fail(MONC_TVD_ERR_ARG, "null record buffer for advected field %d", fl);
    for (int g = 0; g < c->nfields; g++)
      if (out[fl] == in[g]) return fail(MONC_TVD_ERR_ARG, "output of field %d aliases the input of field %d", fl, g);
    for (int g : fields)
      if (g != fl && (out[fl] == out[g] || rec[fl] == rec[g]))
        return fail(MONC_TVD_ERR_ARG, "fields %d and %d share an output", fl, g);
  }
  LaunchArgs<T> L = launch_args<T>(c, dtm);
  L.u = (const T*)in[MONC_TVD_U];
  L.v = (const T*)in[MONC_TVD_V];
  L.w = (const T*)in[MONC_TVD_W];
  for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
    FieldSet<T> S{};
    S.nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
    for (int n = 0; n < S.nf; n++) {
      const int fl = fields[first + n];
      if (in[fl] == out[fl]) return fail(MONC_TVD_ERR_ARG, "field %d: the step cannot run in place", fl);
      S.kind[n] = kind_of(fl);
      S.F[n] = (const T*)in[fl];
      S.f[n] = S.F[n];
      S.out[n] = (T*)out[fl];
      S.dt[n] = (T)dtm;
      S.rec[n] = rec[fl];
    }
    if (int rc = launch_set<T, MODE_REC>(c, L, S)) return rc;
  }
  return MONC_TVD_OK;
}

// The MONC timestep of an advection-only run, resident: per step the TVD
// tendencies from cleared sources (momentum and scalar stepping CENTRED, so
// ultflx takes dt = 2 dtm, steppingdirection.F90:27-32; the advected field is
// the prognostic on a forward step and the z-field otherwise,
// tvdadvection.F90:739-749), then stepfields and swap_smooth fused into the
// same pass (leap_update).  New fields and z-fields go to spare buffers with
// their periodic halo images, and the roles swap after every step.
template <typename T>
int leapfrog_impl(monc_tvd_ctx* c, double dtm, double tsmth, double ugal, double vgal, int nsteps, int first_forward) {
  int rc;
  const int all = MONC_TVD_ADVECT_FLOW | MONC_TVD_ADVECT_THETA | MONC_TVD_ADVECT_Q |
                  (c->n_tracers ? MONC_TVD_ADVECT_TRACER : 0);
  if ((c->advect & all) != all || (int)advected_fields(c).size() != c->nfields)
    return fail(MONC_TVD_ERR_ARG,
                "leapfrog stepping steps every field: create the context with all advect flags and every field active");
  std::vector<int> fields = advected_fields(c);
  for (int fl : fields) {
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
    if (!first_forward && (rc = need(c, fl, MONC_TVD_ZFIELD))) return rc;
    if ((rc = ensure(c, fl, MONC_TVD_ZFIELD))) return rc;
    if (!c->spare[fl]) HIP_TRY(dev_alloc(&c->spare[fl], c->n_elem * c->elem_size));
    if (!c->spare_z[fl]) HIP_TRY(dev_alloc(&c->spare_z[fl], c->n_elem * c->elem_size));
  }
  for (int s = 0; s < nsteps; s++) {
    const bool forward = first_forward && s == 0;
    LaunchArgs<T> L = launch_args<T>(c, dtm);
    L.tsmth = (T)tsmth;
    L.c1 = (T)(1.0 - 2.0 * tsmth);
    L.forward = forward ? 1 : 0;
    for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
      FieldSet<T> S{};
      S.nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
      for (int n = 0; n < S.nf; n++) {
        const int fl = fields[first + n];
        S.kind[n] = kind_of(fl);
        S.f[n] = (const T*)slot(c, fl, MONC_TVD_PROG);
        S.z[n] = (const T*)slot(c, fl, MONC_TVD_ZFIELD);
        S.F[n] = forward ? S.f[n] : S.z[n];
        S.out[n] = (T*)c->spare[fl];
        S.out2[n] = (T*)c->spare_z[fl];
        S.dt[n] = (T)(2.0 * dtm);
        S.gal[n] = (T)(fl == MONC_TVD_U ? ugal : fl == MONC_TVD_V ? vgal : 0.0);
      }
      if ((rc = launch_set<T, MODE_LEAP>(c, L, S))) return rc;
    }
    for (int fl : fields) {
      std::swap(slot(c, fl, MONC_TVD_PROG), c->spare[fl]);
      std::swap(slot(c, fl, MONC_TVD_ZFIELD), c->spare_z[fl]);
    }
  }
  // caller-bound buffers keep holding the state after the call
  for (int fl : fields) {
    for (int v : {MONC_TVD_PROG, MONC_TVD_ZFIELD}) {
      void* b = v == MONC_TVD_PROG ? c->bound[fl] : c->bound_z[fl];
      void*& sp = v == MONC_TVD_PROG ? c->spare[fl] : c->spare_z[fl];
      if (b && slot(c, fl, v) != b) {
        HIP_TRY(hipMemcpyAsync(b, slot(c, fl, v), c->n_elem * c->elem_size, hipMemcpyDeviceToDevice, c->stream));
        std::swap(slot(c, fl, v), sp);
      }
    }
  }
  return MONC_TVD_OK;
}

// Tile-shape autotuner: for each kernel kind of the forward step (u, v, w,
// the scalar group) every compiled march shape (rows per block) and strip
// length is timed with HIP events on the context's own fields, writing the
// step's result into the spare buffers (the state is not advanced), and the
// fastest shape per kind is kept.
template <typename T>
int autotune_impl(monc_tvd_ctx* c, double dtm, int reps, int* chosen) {
  int rc;
  const int nch = (c->nz + 63) / 64;
  if (nch > 4 || !c->buffer_ok) return fail(MONC_TVD_ERR_ARG, "the march kernel does not run this context");
  for (int fl : {MONC_TVD_U, MONC_TVD_V, MONC_TVD_W})
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
  std::vector<int> fields = advected_fields(c);
  for (int fl : fields) {
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
    if (!c->spare[fl]) HIP_TRY(dev_alloc(&c->spare[fl], c->n_elem * c->elem_size));
  }
  const LaunchArgs<T> L = launch_args<T>(c, dtm);
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  static const int strips[] = {8, 16, 32, 64};
  for (size_t n = 0; n < fields.size();) {
    FieldSet<T> G{};
    const int fl0 = fields[n];
    const int kind = kind_of(fl0);
    const bool pair = kind == KIND_SCALAR && n + 1 < fields.size();
    G.nf = pair ? 2 : 1;
    for (int m = 0; m < G.nf; m++) {
      const int fl = fields[n + m];
      G.kind[m] = kind;
      G.F[m] = (const T*)slot(c, fl, MONC_TVD_PROG);
      G.f[m] = G.F[m];
      G.out[m] = (T*)c->spare[fl];
      G.dt[m] = (T)dtm;
    }
    const int kk = kind == KIND_SCALAR ? 3 : kind;
    float best = 1e30f;
    int best_rows = 0, best_strip = 0, best_kernel = MONC_TVD_KERNEL_MARCH;
    // the stage kernel's shapes where it takes this field group
    if (stage_fits<T, MODE_STEP>(c, L, G)) {
      for (int rows : STAGE_ROWS) {
        if (rows == 4 || !stage_shape_compiled(c->precision, kind, rows)) continue;
        for (int strip : {32, 64, 128}) {
          float ms = 0.f;
          for (int it = 0; it <= reps; it++) {
            if (it == 1) HIP_TRY(hipEventRecord(a, c->stream));
            launch_stage_group<T>(c, L, G, kk, rows, strip);
            HIP_TRY(hipGetLastError());
          }
          HIP_TRY(hipEventRecord(b, c->stream));
          HIP_TRY(hipEventSynchronize(b));
          HIP_TRY(hipEventElapsedTime(&ms, a, b));
          if (ms < best) {
            best = ms;
            best_rows = rows;
            best_strip = strip;
            best_kernel = MONC_TVD_KERNEL_STAGE;
          }
        }
      }
    }
    for (int rows : MARCH_ROWS[nch - 1]) {
      if (!rows) continue;
      for (int strip : strips) {
        float ms = 0.f;
        for (int it = 0; it <= reps; it++) {
          if (it == 1) HIP_TRY(hipEventRecord(a, c->stream));
          switch (kind) {
            case KIND_U: launch_shape<T, KIND_U, 1, MODE_STEP>(c, L, G, nch, MONC_TVD_KERNEL_MARCH, rows, strip); break;
            case KIND_V: launch_shape<T, KIND_V, 1, MODE_STEP>(c, L, G, nch, MONC_TVD_KERNEL_MARCH, rows, strip); break;
            case KIND_W: launch_shape<T, KIND_W, 1, MODE_STEP>(c, L, G, nch, MONC_TVD_KERNEL_MARCH, rows, strip); break;
            default:
              if (pair) launch_shape<T, KIND_SCALAR, 2, MODE_STEP>(c, L, G, nch, MONC_TVD_KERNEL_MARCH, rows, strip);
              else launch_shape<T, KIND_SCALAR, 1, MODE_STEP>(c, L, G, nch, MONC_TVD_KERNEL_MARCH, rows, strip);
          }
          HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipEventRecord(b, c->stream));
        HIP_TRY(hipEventSynchronize(b));
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        if (ms < best) {
          best = ms;
          best_rows = rows;
          best_strip = strip;
          best_kernel = MONC_TVD_KERNEL_MARCH;
        }
      }
    }
    c->rows[kk] = best_rows;
    c->strip[kk] = best_strip;
    c->kind_kernel[kk] = best_kernel;
    n += G.nf;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->kernel = MONC_TVD_KERNEL_MARCH;
  if (chosen)
    for (int q = 0; q < 4; q++) {
      chosen[3 * q] = c->kind_kernel[q];
      chosen[3 * q + 1] = c->rows[q];
      chosen[3 * q + 2] = c->strip[q];
    }
  return MONC_TVD_OK;
}

// Adjoint of the periodic fill on every allocated adjoint field: rows of all
// fields in one launch, then columns.
template <typename T>
int fold_halo(monc_tvd_ctx* c) {
  FieldPtrs<T> P{};
  int nf = 0;
  for (int f = 0; f < c->nfields; f++) {
    void* p = slot(c, f, MONC_TVD_ADJOINT);
    if (!p) continue;
    if (nf == MAX_FIELDS) {  // more fields than one launch holds
      Dims d = dims_of(c);
      dim3 block(64, 4, 1);
      dim3 g1((c->nz + 63) / 64, (unsigned)((4 * (c->nx + 4) + 3) / 4), nf);
      hipLaunchKernelGGL((fold_rows_kernel<T>), g1, block, 0, c->stream, P, d);
      dim3 g2((c->nz + 63) / 64, (unsigned)((4 * c->ny + 3) / 4), nf);
      hipLaunchKernelGGL((fold_cols_kernel<T>), g2, block, 0, c->stream, P, d);
      HIP_TRY(hipGetLastError());
      nf = 0;
    }
    P.f[nf++] = (T*)p;
  }
  if (!nf) return MONC_TVD_OK;
  Dims d = dims_of(c);
  dim3 block(64, 4, 1);
  dim3 g1((c->nz + 63) / 64, (unsigned)((4 * (c->nx + 4) + 3) / 4), nf);
  hipLaunchKernelGGL((fold_rows_kernel<T>), g1, block, 0, c->stream, P, d);
  HIP_TRY(hipGetLastError());
  dim3 g2((c->nz + 63) / 64, (unsigned)((4 * c->ny + 3) / 4), nf);
  hipLaunchKernelGGL((fold_cols_kernel<T>), g2, block, 0, c->stream, P, d);
  HIP_TRY(hipGetLastError());
  return MONC_TVD_OK;
}

template <typename T>
int adjoint_step_impl(monc_tvd_ctx* c, double dtm) {
  int rc;
  std::vector<int> fields = advected_fields(c);
  for (int fl : {MONC_TVD_U, MONC_TVD_V, MONC_TVD_W}) {
    if ((rc = need(c, fl, MONC_TVD_PROG))) return rc;
    if ((rc = ensure(c, fl, MONC_TVD_ADJOINT))) return rc;  // flows receive adjoint from every field
  }
  for (int fl : fields) {
    if ((rc = need(c, fl, MONC_TVD_PROG)) || (rc = need(c, fl, MONC_TVD_ADJOINT))) return rc;
    if (!c->adj_tmp[fl]) HIP_TRY(dev_alloc(&c->adj_tmp[fl], c->n_elem * c->elem_size));
  }
  // the incoming adjoint is read while the accumulators (which start from it:
  // the identity part of X' = X + dtm T(X)) receive the transposed Jacobian
  const int64_t words = (int64_t)(c->n_elem * c->elem_size / 4);
  bool aligned = true;
  for (int fl : fields)
    aligned = aligned && ((uintptr_t)slot(c, fl, MONC_TVD_ADJOINT) % 16 == 0);
  for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
    const int nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
    if (!aligned) {
      for (int n = 0; n < nf; n++)
        HIP_TRY(hipMemcpyAsync(c->adj_tmp[fields[first + n]], slot(c, fields[first + n], MONC_TVD_ADJOINT),
                               c->n_elem * c->elem_size, hipMemcpyDeviceToDevice, c->stream));
      continue;
    }
    FieldPtrs<T> P{};
    for (int n = 0; n < nf; n++) {
      P.f[n] = (T*)c->adj_tmp[fields[first + n]];
      P.src[n] = (const T*)slot(c, fields[first + n], MONC_TVD_ADJOINT);
    }
    const unsigned nb = (unsigned)std::min<int64_t>((words / 4 + 255) / 256 + 1, 4096);
    hipLaunchKernelGGL((copy_fields_kernel<T>), dim3(nb, nf, 1), dim3(256), 0, c->stream, P, words);
    HIP_TRY(hipGetLastError());
  }
  const LaunchArgs<T> L = launch_args<T>(c, dtm);
  // march adjoint: 64-level bands x 4 owner rows x strips of 4 owner columns.
  // Short strips keep the backward sweep of a localised loss short: its few
  // non-zero tiles each march their strip serially (a 64-step rollout at
  // 256x256x128 runs its backward 1.6x faster than with 16 columns).
  constexpr int ADJ_STRIP = 4;
  dim3 block(64, ADJ_ROWS, 1);
  dim3 grid((c->nz + 63) / 64, (c->ny + 1 + ADJ_ROWS - 1) / ADJ_ROWS, (c->nx + 1 + ADJ_STRIP - 1) / ADJ_STRIP);
  T* const lu = (T*)slot(c, MONC_TVD_U, MONC_TVD_ADJOINT);
  T* const lv = (T*)slot(c, MONC_TVD_V, MONC_TVD_ADJOINT);
  T* const lw = (T*)slot(c, MONC_TVD_W, MONC_TVD_ADJOINT);
  // the limiter arms the forward step recorded, when every advected field
  // has them; otherwise the adjoint re-derives them from the state
  bool rec = true;
  for (int fl : fields) rec = rec && c->rec[fl];
  auto args1 = [&](int fl) {
    AdjArgs<T, 1> A{};
    A.F[0] = (const T*)slot(c, fl, MONC_TVD_PROG);
    A.lin[0] = (const T*)c->adj_tmp[fl];
    A.lf[0] = (T*)slot(c, fl, MONC_TVD_ADJOINT);
    A.rec[0] = c->rec[fl];
    A.lu = lu, A.lv = lv, A.lw = lw;
    A.dtm = (T)dtm;
    return A;
  };
  auto launch1 = [&](int kind, const AdjArgs<T, 1>& A) {
#define MONC_ADJ1(K, R)                                                                                      \
  hipLaunchKernelGGL((adjoint_march_kernel<T, K, 1, R>), grid, block, 0, c->stream, L, A, ADJ_STRIP)
    if (rec) {
      if (kind == KIND_U) MONC_ADJ1(KIND_U, true);
      else if (kind == KIND_V) MONC_ADJ1(KIND_V, true);
      else if (kind == KIND_W) MONC_ADJ1(KIND_W, true);
      else MONC_ADJ1(KIND_SCALAR, true);
    } else {
      if (kind == KIND_U) MONC_ADJ1(KIND_U, false);
      else if (kind == KIND_V) MONC_ADJ1(KIND_V, false);
      else if (kind == KIND_W) MONC_ADJ1(KIND_W, false);
      else MONC_ADJ1(KIND_SCALAR, false);
    }
#undef MONC_ADJ1
  };
  std::vector<int> scalars;
  for (int fl : fields) {
    const int kind = kind_of(fl);
    if (kind == KIND_SCALAR) {
      scalars.push_back(fl);
      continue;
    }
    LaunchTimer timer(c, kind);
    launch1(kind, args1(fl));
    HIP_TRY(hipGetLastError());
  }
  // scalars in pairs sharing their flows' recompute and reverse (th with the
  // first q, ...), an odd one alone
  for (size_t n = 0; n < scalars.size(); n += 2) {
    LaunchTimer timer(c, KIND_SCALAR);
    if (n + 1 < scalars.size()) {
      const AdjArgs<T, 1> A0 = args1(scalars[n]), A1 = args1(scalars[n + 1]);
      AdjArgs<T, 2> A;
      A.F[0] = A0.F[0], A.F[1] = A1.F[0];
      A.lin[0] = A0.lin[0], A.lin[1] = A1.lin[0];
      A.lf[0] = A0.lf[0], A.lf[1] = A1.lf[0];
      A.rec[0] = A0.rec[0], A.rec[1] = A1.rec[0];
      A.lu = lu, A.lv = lv, A.lw = lw;
      A.dtm = (T)dtm;
      if (rec)
        hipLaunchKernelGGL((adjoint_march_kernel<T, KIND_SCALAR, 2, true>), grid, block, 0, c->stream, L, A,
                           ADJ_STRIP);
      else
        hipLaunchKernelGGL((adjoint_march_kernel<T, KIND_SCALAR, 2, false>), grid, block, 0, c->stream, L, A,
                           ADJ_STRIP);
    } else {
      launch1(KIND_SCALAR, args1(scalars[n]));
    }
    HIP_TRY(hipGetLastError());
  }
  return fold_halo<T>(c);
}

int tangent_steps_impl(monc_tvd_ctx* c, double dtm, int nsteps) {
  int rc;
  std::vector<int> fields = advected_fields(c);
  std::vector<int> state = {MONC_TVD_U, MONC_TVD_V, MONC_TVD_W};
  for (int fl : fields)
    if (fl > MONC_TVD_W) state.push_back(fl);
  const int64_t n = (int64_t)c->n_elem;
  const unsigned nb = (unsigned)((n + 255) / 256);
  for (int fl : state) {
    if ((rc = need(c, fl, MONC_TVD_PROG)) || (rc = need(c, fl, MONC_TVD_TANGENT))) return rc;
    for (int b = 0; b < 2; b++)
      if (!c->dual[b][fl]) HIP_TRY(dev_alloc(&c->dual[b][fl], c->n_elem * sizeof(Dual)));
    hipLaunchKernelGGL(pack_dual_kernel, dim3(nb), dim3(256), 0, c->stream, (Dual*)c->dual[0][fl],
                       (const double*)slot(c, fl, MONC_TVD_PROG), (const double*)slot(c, fl, MONC_TVD_TANGENT), n);
    HIP_TRY(hipGetLastError());
  }
  // flows that are not advected keep their (value, tangent) across steps
  for (int fl : {MONC_TVD_U, MONC_TVD_V, MONC_TVD_W})
    if (!(c->advect & MONC_TVD_ADVECT_FLOW))
      HIP_TRY(hipMemcpyAsync(c->dual[1][fl], c->dual[0][fl], c->n_elem * sizeof(Dual), hipMemcpyDeviceToDevice,
                             c->stream));
  int cur = 0;
  for (int s = 0; s < nsteps; s++) {
    LaunchArgs<Dual> L;
    L.d = dims_of(c);
    L.u = (const Dual*)c->dual[cur][MONC_TVD_U];
    L.v = (const Dual*)c->dual[cur][MONC_TVD_V];
    L.w = (const Dual*)c->dual[cur][MONC_TVD_W];
    L.mc = vmet_of(c, false);
    L.mw = vmet_of(c, true);
    L.dtm = Dual(dtm);
    L.cx = Dual(c->cx);
    L.cy = Dual(c->cy);
    for (size_t first = 0; first < fields.size(); first += MAX_FIELDS) {
      FieldSet<Dual> S{};
      S.nf = (int)std::min(fields.size() - first, (size_t)MAX_FIELDS);
      for (int m = 0; m < S.nf; m++) {
        const int fl = fields[first + m];
        S.kind[m] = kind_of(fl);
        S.F[m] = (const Dual*)c->dual[cur][fl];
        S.f[m] = S.F[m];
        S.out[m] = (Dual*)c->dual[1 - cur][fl];
        S.dt[m] = Dual(dtm);
      }
      dim3 block(64, 4, 1);
      dim3 grid((c->nz + 63) / 64, (c->ny + 3) / 4, c->nx);
      hipLaunchKernelGGL((cell_kernel<Dual, MODE_STEP>), grid, block, 0, c->stream, L, S);
      HIP_TRY(hipGetLastError());
    }
    cur = 1 - cur;  // flows that are not advected hold the same pair in both buffers
  }
  for (int fl : state) {
    hipLaunchKernelGGL(unpack_dual_kernel, dim3(nb), dim3(256), 0, c->stream, (double*)slot(c, fl, MONC_TVD_PROG),
                       (double*)slot(c, fl, MONC_TVD_TANGENT), (const Dual*)c->dual[cur][fl], n);
    HIP_TRY(hipGetLastError());
  }
  return MONC_TVD_OK;
}

// Weighted sum of one field through wsum_partial_kernel / wsum_final_kernel
// on `scratch` (wsum_scratch_len(nz) doubles: weights, partials, result);
// synchronises the stream.
size_t wsum_scratch_len(int nz) { return (size_t)nz + WSUM_BLOCKS + 1; }

int weighted_sum_on(monc_tvd_ctx* c, const void* field, const double* weights, double* scratch, double* out) {
  double* dw = scratch;
  double* dpart = scratch + c->nz;
  double* dres = dpart + WSUM_BLOCKS;
  const int64_t ngroups = (int64_t)((c->ny + 3) / 4) * c->nx;
  const int nb = (int)std::min<int64_t>(ngroups, WSUM_BLOCKS);
  HIP_TRY(hipMemcpyAsync(dw, weights, sizeof(double) * c->nz, hipMemcpyHostToDevice, c->stream));
  if (c->precision == MONC_TVD_FP64)
    hipLaunchKernelGGL((wsum_partial_kernel<double>), dim3(nb), dim3(64, 4, 1), 0, c->stream, (const double*)field,
                       dims_of(c), (const double*)dw, dpart);
  else
    hipLaunchKernelGGL((wsum_partial_kernel<float>), dim3(nb), dim3(64, 4, 1), 0, c->stream, (const float*)field,
                       dims_of(c), (const double*)dw, dpart);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL((wsum_final_kernel<double>), dim3(1), dim3(256), 0, c->stream, (const double*)dpart, nb, dres);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, dres, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MONC_TVD_OK;
}

// Scratch of monc_tvd_weighted_sum_noalloc: a file-scope device array (one
// copy per GPU, held until the process ends) of weights, partials and the
// result for columns of up to WSUM_NOALLOC_MAX_NZ levels.  Host code never
// passes the array's name to a kernel or copy: it is a host-side stand-in.
// Each device's own address is taken once, with that device current, by
// hipGetSymbolAddress and kept in wsum_noalloc_ptr.
constexpr int WSUM_NOALLOC_MAX_NZ = 4096;
constexpr int WSUM_NOALLOC_LEN = WSUM_NOALLOC_MAX_NZ + WSUM_BLOCKS + 1;
static __device__ __align__(256) double scratch[WSUM_NOALLOC_LEN];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[MAX_DEVICES] = {};

int wsum_noalloc_ptr(double** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(MONC_TVD_ERR_ARG, "device %d beyond the %d supported", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return MONC_TVD_OK;
}

}  // namespace

extern "C" {

const char* monc_tvd_last_error(void) { return g_err.c_str(); }

int monc_tvd_test_fail_allocations(int n) {
  return __atomic_exchange_n(&g_fail_allocs, n > 0 ? n : 0, __ATOMIC_RELAXED);
}

int monc_tvd_create(int nz, int ny, int nx, int halo, int n_q, int n_tracers, int advect_flags, double cx,
                    double cy, const double* rdz, const double* rdzn, const double* dz, const double* dzn,
                    const double* tzc1, const double* tzc2, const double* tzd1, const double* tzd2, int precision,
                    monc_tvd_ctx** out) {
  g_err.clear();
  if (!out) return fail(MONC_TVD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (halo != 2) return fail(MONC_TVD_ERR_ARG, "halo depth %d unsupported: the TVD stencil needs exactly 2", halo);
  if (nz < 4 || ny < 4 || nx < 4)
    return fail(MONC_TVD_ERR_ARG, "local grid %dx%dx%d too small (need nz, ny, nx >= 4)", nz, ny, nx);
  if (n_q < 0 || n_q > 64) return fail(MONC_TVD_ERR_ARG, "n_q %d out of range [0,64]", n_q);
  if (n_tracers < 0 || n_tracers > 64) return fail(MONC_TVD_ERR_ARG, "n_tracers %d out of range [0,64]", n_tracers);
  if (precision != MONC_TVD_FP64 && precision != MONC_TVD_FP32)
    return fail(MONC_TVD_ERR_ARG, "unknown precision %d", precision);
  if (advect_flags & ~15) return fail(MONC_TVD_ERR_ARG, "unknown advect flags 0x%x", advect_flags);
  if (!(cx > 0.0) || !(cy > 0.0) || !std::isfinite(cx) || !std::isfinite(cy))
    return fail(MONC_TVD_ERR_ARG, "cx, cy must be positive and finite");
  const double* mets[MET_N] = {rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2};
  for (int m = 0; m < MET_N; m++)
    if (!mets[m]) return fail(MONC_TVD_ERR_ARG, "metric vector %d is null", m);
  const size_t n_elem = (size_t)nz * (size_t)(ny + 4) * (size_t)(nx + 4);
  if (n_elem > ((size_t)1 << 40)) return fail(MONC_TVD_ERR_ARG, "field too large");

  monc_tvd_ctx* c = new monc_tvd_ctx();
  c->nz = nz;
  c->ny = ny;
  c->nx = nx;
  c->n_q = n_q;
  c->n_tracers = n_tracers;
  c->advect = advect_flags;
  c->precision = precision;
  c->cx = cx;
  c->cy = cy;
  c->n_elem = n_elem;
  c->elem_size = precision == MONC_TVD_FP64 ? 8 : 4;
  c->nfields = 4 + n_q + n_tracers;
  c->active.assign((size_t)c->nfields, 1);
  c->buf.assign((size_t)c->nfields * NVAR, nullptr);
  c->owned.assign((size_t)c->nfields * NVAR, 0);
  c->bound.assign((size_t)c->nfields, nullptr);
  c->bound_z.assign((size_t)c->nfields, nullptr);
  c->spare.assign((size_t)c->nfields, nullptr);
  c->spare_z.assign((size_t)c->nfields, nullptr);
  c->adj_tmp.assign((size_t)c->nfields, nullptr);
  c->rec.assign((size_t)c->nfields, nullptr);
  c->dual[0].assign((size_t)c->nfields, nullptr);
  c->dual[1].assign((size_t)c->nfields, nullptr);
  c->stream = nullptr;
  c->dmet = nullptr;
  c->kernel = MONC_TVD_KERNEL_MARCH;
  for (int q = 0; q < 4; q++) {
    c->rows[q] = c->strip[q] = 0;
    c->kind_kernel[q] = MONC_TVD_KERNEL_MARCH;
  }
  c->buffer_ok = n_elem * c->elem_size < ((size_t)1 << 32) - ((size_t)1 << 20);
  c->profiling = 0;
  c->diag_flags = 0;
  c->rprefrcp = nullptr;
  c->diag_before.assign((size_t)c->nfields, nullptr);
  c->diag_zf.assign((size_t)c->nfields, nullptr);
  for (int q = 0; q < MONC_TVD_NCOMPONENTS; q++) {
    c->diag_3d[q].assign((size_t)c->nfields, nullptr);
    c->diag_prof[q].assign((size_t)c->nfields, nullptr);
    c->diag_have[q].assign((size_t)c->nfields, 0);
    c->tabs_3d[q] = nullptr;
    c->tabs_prof[q] = nullptr;
    c->tabs_have[q] = 0;
  }
  c->diag_partial = nullptr;
  for (int k = 0; k < MONC_TVD_NKERNEL_KINDS; k++) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  std::vector<double> h((size_t)MET_N * nz);
  for (int m = 0; m < MET_N; m++) std::memcpy(&h[(size_t)m * nz], mets[m], sizeof(double) * nz);
  hipError_t e = dev_alloc(&c->dmet, h.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->dmet, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (c->dmet) (void)hipFree(c->dmet);
    delete c;
    return fail(MONC_TVD_ERR_HIP, "metric upload: %s", hipGetErrorString(e));
  }
  *out = c;
  return MONC_TVD_OK;
}

int monc_tvd_set_field_active(monc_tvd_ctx* c, int field, int active) {
  if (int rc = check_ctx(c)) return rc;
  if (field != MONC_TVD_TH && (field < MONC_TVD_Q(0) || field >= MONC_TVD_Q(c->n_q)))
    return fail(MONC_TVD_ERR_ARG, "field %d has no active flag: only th and the q fields are gated", field);
  c->active[field] = active ? 1 : 0;
  return MONC_TVD_OK;
}

int monc_tvd_destroy(monc_tvd_ctx* c) {
  if (!c) return MONC_TVD_OK;
  (void)hipStreamSynchronize(c->stream);
  for (auto& t : c->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (size_t n = 0; n < c->buf.size(); n++) {
    void* p = c->buf[n];
    if (!p) continue;
    const int field = (int)(n / NVAR), variant = (int)(n % NVAR);
    // after an odd number of steps a bound prognostic and its spare swapped
    // roles; free only library memory
    if (variant == MONC_TVD_PROG && c->bound[field]) {
      if (p != c->bound[field]) (void)hipFree(p);
      continue;
    }
    if (variant == MONC_TVD_ZFIELD && c->bound_z[field]) {
      if (p != c->bound_z[field]) (void)hipFree(p);
      continue;
    }
    if (c->owned[n]) (void)hipFree(p);
  }
  for (int f = 0; f < c->nfields; f++) {
    if (c->spare[f] && c->spare[f] != c->bound[f]) (void)hipFree(c->spare[f]);
    if (c->spare_z[f] && c->spare_z[f] != c->bound_z[f]) (void)hipFree(c->spare_z[f]);
    if (c->adj_tmp[f]) (void)hipFree(c->adj_tmp[f]);
    for (int b = 0; b < 2; b++)
      if (c->dual[b][f]) (void)hipFree(c->dual[b][f]);
  }
  for (int f = 0; f < c->nfields; f++) {
    for (void* p : {c->diag_before[f], c->diag_zf[f]})
      if (p) (void)hipFree(p);
    for (int q = 0; q < MONC_TVD_NCOMPONENTS; q++)
      for (void* p : {c->diag_3d[q][f], (void*)c->diag_prof[q][f]})
        if (p) (void)hipFree(p);
  }
  for (int q = 0; q < MONC_TVD_NCOMPONENTS; q++)
    for (void* p : {c->tabs_3d[q], (void*)c->tabs_prof[q]})
      if (p) (void)hipFree(p);
  for (void* p : {(void*)c->rprefrcp, (void*)c->diag_partial, (void*)c->thadv_tab,
                  (void*)c->wsum_scratch})
    if (p) (void)hipFree(p);
  if (c->dmet) (void)hipFree(c->dmet);
  delete c;
  return MONC_TVD_OK;
}

int monc_tvd_set_tiling(monc_tvd_ctx* c, int kernel, int rows, int strip) {
  if (int rc = check_ctx(c)) return rc;
  if (kernel != MONC_TVD_KERNEL_MARCH && kernel != MONC_TVD_KERNEL_CELL && kernel != MONC_TVD_KERNEL_STAGE)
    return fail(MONC_TVD_ERR_ARG, "kernel %d is not 0 (march), 1 (cell) or 2 (stage)", kernel);
  if (kernel == MONC_TVD_KERNEL_STAGE && rows != 0 && !stage_rows_ok(rows))
    return fail(MONC_TVD_ERR_ARG, "no stage shape with %d rows (4, 6, 8, 11 or 15)", rows);
  if (strip < 0 || strip > 4096) return fail(MONC_TVD_ERR_ARG, "strip %d out of range [0,4096]", strip);
  const int nch = (c->nz + 63) / 64;
  if (kernel == MONC_TVD_KERNEL_MARCH && rows != 0 && !march_rows_ok(nch, rows))
    return fail(MONC_TVD_ERR_ARG, "no march shape with %d rows for %d-level columns", rows, c->nz);
  if (rows < 0) return fail(MONC_TVD_ERR_ARG, "rows %d < 0", rows);
  c->kernel = kernel == MONC_TVD_KERNEL_CELL ? MONC_TVD_KERNEL_CELL : MONC_TVD_KERNEL_MARCH;
  for (int q = 0; q < 4; q++) {
    c->rows[q] = kernel == MONC_TVD_KERNEL_CELL ? 0 : rows;
    c->strip[q] = strip;
    c->kind_kernel[q] = kernel == MONC_TVD_KERNEL_STAGE ? MONC_TVD_KERNEL_STAGE : MONC_TVD_KERNEL_MARCH;
  }
  return MONC_TVD_OK;
}

int monc_tvd_get_tiling(monc_tvd_ctx* c, int* chosen) {
  if (int rc = check_ctx(c)) return rc;
  if (!chosen) return fail(MONC_TVD_ERR_ARG, "null output array");
  const int nch = (c->nz + 63) / 64;
  const bool march_fits = nch <= 4 && c->buffer_ok;
  for (int q = 0; q < 4; q++) {
    const int kind = q == 3 ? KIND_SCALAR : q;
    int kernel = c->kernel == MONC_TVD_KERNEL_CELL || !march_fits ? MONC_TVD_KERNEL_CELL : c->kind_kernel[q];
    int rows = c->rows[q], strip = c->strip[q];
    const int dstrip = stage_default_strip(c, kind);
    if (kernel == MONC_TVD_KERNEL_MARCH && !rows && !strip && c->buffer_ok && dstrip > 0) {
      kernel = MONC_TVD_KERNEL_STAGE;
      rows = stage_default_rows(c->precision, kind);
      strip = dstrip;
    } else if (kernel == MONC_TVD_KERNEL_STAGE) {
      rows = rows > 0 ? rows : 6;
      strip = strip > 0 ? strip : 64;
    } else if (kernel == MONC_TVD_KERNEL_MARCH) {
      rows = rows > 0 ? rows : default_rows(c->precision, nch, q);
      const bool wide = c->precision == MONC_TVD_FP64 && q == 3 && nch == 4 && rows <= 2;
      strip = strip > 0 ? strip : default_strip(c->ny, c->nx, rows, wide);
    } else {
      rows = strip = 0;
    }
    chosen[3 * q] = kernel;
    chosen[3 * q + 1] = rows;
    chosen[3 * q + 2] = strip;
  }
  return MONC_TVD_OK;
}

int monc_tvd_march_shapes(monc_tvd_ctx* c, int* rows, int n) {
  if (int rc = check_ctx(c)) return rc;
  if (n < 0 || (n > 0 && !rows)) return fail(MONC_TVD_ERR_ARG, "null output array");
  const int nch = (c->nz + 63) / 64;
  int m = 0;
  if (nch >= 1 && nch <= 4 && c->buffer_ok)
    for (int r : MARCH_ROWS[nch - 1])
      if (r) {
        if (m < n) rows[m] = r;
        m++;
      }
  return m;
}

int monc_tvd_weighted_sum(monc_tvd_ctx* c, int field, int variant, const double* weights, double* out) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_fv(c, field, variant)) return rc;
  if (!weights || !out) return fail(MONC_TVD_ERR_ARG, "null weights or output");
  if (int rc = need(c, field, variant)) return rc;
  // scratch of the context, allocated on the first call and kept
  if (!c->wsum_scratch) HIP_TRY(dev_alloc(&c->wsum_scratch, sizeof(double) * wsum_scratch_len(c->nz)));
  return weighted_sum_on(c, slot(c, field, variant), weights, c->wsum_scratch, out);
}

int monc_tvd_weighted_sum_noalloc(monc_tvd_ctx* c, int field, int variant, const double* weights, double* out) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_fv(c, field, variant)) return rc;
  if (!weights || !out) return fail(MONC_TVD_ERR_ARG, "null weights or output");
  if (c->nz > WSUM_NOALLOC_MAX_NZ)
    return fail(MONC_TVD_ERR_ARG, "nz = %d exceeds the %d levels of the fixed scratch of the no-allocation sum", c->nz,
                WSUM_NOALLOC_MAX_NZ);
  if (int rc = need(c, field, variant)) return rc;
  double* ptr = nullptr;
  if (int rc = wsum_noalloc_ptr(&ptr)) return rc;
  // every word of the scratch a call reads is written earlier in the same call
  // (weights by the copy, partials by the first kernel, the result by the
  // second): nothing needs clearing between calls
  return weighted_sum_on(c, slot(c, field, variant), weights, ptr, out);
}

int monc_tvd_set_profiling(monc_tvd_ctx* c, int on) {
  if (int rc = check_ctx(c)) return rc;
  c->profiling = on ? 1 : 0;
  return MONC_TVD_OK;
}

int monc_tvd_kernel_times(monc_tvd_ctx* c, double* ms, long long* launches, int n, int reset) {
  if (int rc = check_ctx(c)) return rc;
  if (n < 0 || (n > 0 && (!ms || !launches))) return fail(MONC_TVD_ERR_ARG, "null output arrays");
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto& t : c->pending) {
    float el = 0.f;
    hipError_t e = hipEventElapsedTime(&el, t.a, t.b);
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
    if (e != hipSuccess) continue;
    c->prof_ms[t.kind] += el;
    c->prof_n[t.kind] += 1;
  }
  c->pending.clear();
  for (int k = 0; k < n && k < MONC_TVD_NKERNEL_KINDS; k++) {
    ms[k] = c->prof_ms[k];
    launches[k] = c->prof_n[k];
  }
  if (reset)
    for (int k = 0; k < MONC_TVD_NKERNEL_KINDS; k++) {
      c->prof_ms[k] = 0.0;
      c->prof_n[k] = 0;
    }
  return MONC_TVD_OK;
}

int monc_tvd_set_stream(monc_tvd_ctx* c, void* s) {
  if (int rc = check_ctx(c)) return rc;
  c->stream = (hipStream_t)s;
  return MONC_TVD_OK;
}

int monc_tvd_synchronize(monc_tvd_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MONC_TVD_OK;
}

int monc_tvd_bind_field(monc_tvd_ctx* c, int field, int variant, void* dptr) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_fv(c, field, variant)) return rc;
  if (!dptr) return fail(MONC_TVD_ERR_ARG, "null device pointer");
  const size_t n = (size_t)field * NVAR + variant;
  if (c->buf[n] && c->owned[n]) HIP_TRY(hipFree(c->buf[n]));
  c->buf[n] = dptr;
  c->owned[n] = 0;
  if (variant == MONC_TVD_PROG) c->bound[field] = dptr;
  if (variant == MONC_TVD_ZFIELD) c->bound_z[field] = dptr;
  return MONC_TVD_OK;
}

int monc_tvd_field_device_ptr(monc_tvd_ctx* c, int field, int variant, void** dptr) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_fv(c, field, variant)) return rc;
  if (!dptr) return fail(MONC_TVD_ERR_ARG, "null output pointer");
  if (int rc = ensure(c, field, variant)) return rc;
  *dptr = slot(c, field, variant);
  return MONC_TVD_OK;
}

int monc_tvd_set_field(monc_tvd_ctx* c, int field, int variant, const double* host) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_fv(c, field, variant)) return rc;
  if (!host) return fail(MONC_TVD_ERR_ARG, "null host pointer");
  if (int rc = ensure(c, field, variant)) return rc;
  void* d = slot(c, field, variant);
  if (c->precision == MONC_TVD_FP64) {
    HIP_TRY(hipMemcpyAsync(d, host, c->n_elem * 8, hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> tmp(c->n_elem);
    for (size_t n = 0; n < c->n_elem; n++) tmp[n] = (float)host[n];
    HIP_TRY(hipMemcpyAsync(d, tmp.data(), c->n_elem * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MONC_TVD_OK;
}

int monc_tvd_get_field(monc_tvd_ctx* c, int field, int variant, double* host) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = check_fv(c, field, variant)) return rc;
  if (!host) return fail(MONC_TVD_ERR_ARG, "null host pointer");
  if (int rc = need(c, field, variant)) return rc;
  void* d = slot(c, field, variant);
  if (c->precision == MONC_TVD_FP64) {
    HIP_TRY(hipMemcpyAsync(host, d, c->n_elem * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(c->n_elem);
    HIP_TRY(hipMemcpyAsync(tmp.data(), d, c->n_elem * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t n = 0; n < c->n_elem; n++) host[n] = (double)tmp[n];
  }
  return MONC_TVD_OK;
}

int monc_tvd_halo_fill(monc_tvd_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  for (int f = 0; f < c->nfields; f++)
    for (int v : {MONC_TVD_PROG, MONC_TVD_ZFIELD, MONC_TVD_TANGENT}) {
      void* p = slot(c, f, v);
      if (!p) continue;
      int rc = c->precision == MONC_TVD_FP64 ? launch_halo<double>(c, (double*)p) : launch_halo<float>(c, (float*)p);
      if (rc) return rc;
    }
  return MONC_TVD_OK;
}

int monc_tvd_tendencies(monc_tvd_ctx* c, double dtm, int field_stepping, int momentum_stepping, int scalar_stepping) {
  if (int rc = check_ctx(c)) return rc;
  for (int s : {field_stepping, momentum_stepping, scalar_stepping})
    if (s != MONC_TVD_CENTRED_STEPPING && s != MONC_TVD_FORWARD_STEPPING)
      return fail(MONC_TVD_ERR_ARG, "stepping flag %d is neither CENTRED (0) nor FORWARD (1)", s);
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return c->precision == MONC_TVD_FP64
             ? tendencies_impl<double>(c, dtm, field_stepping, momentum_stepping, scalar_stepping)
             : tendencies_impl<float>(c, dtm, field_stepping, momentum_stepping, scalar_stepping);
}

int monc_tvd_adjoint_step(monc_tvd_ctx* c, double dtm) {
  if (int rc = check_ctx(c)) return rc;
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return c->precision == MONC_TVD_FP64 ? adjoint_step_impl<double>(c, dtm) : adjoint_step_impl<float>(c, dtm);
}

int monc_tvd_adjoint_fold(monc_tvd_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  return c->precision == MONC_TVD_FP64 ? fold_halo<double>(c) : fold_halo<float>(c);
}

int monc_thadv_tendencies(monc_tvd_ctx* c, int anelastic, const double* dthref, int baroclinicity,
                          const double* thref, double thref0, double fcoriol, double g, double rate_x, double rate_y,
                          double ugal, double vgal) {
  if (int rc = check_ctx(c)) return rc;
  if (!anelastic && !baroclinicity && !(c->diag_flags & MONC_TVD_DIAG_TENDENCY)) return MONC_TVD_OK;
  if (anelastic && !dthref) return fail(MONC_TVD_ERR_ARG, "anelastic th_advection needs dthref");
  if (baroclinicity && anelastic && !thref)
    return fail(MONC_TVD_ERR_ARG, "anelastic advection of the mean baroclinicity needs thref");
  if (baroclinicity && !(g != 0.0)) return fail(MONC_TVD_ERR_ARG, "g must be non-zero");
  if (int rc = need(c, MONC_TVD_TH, MONC_TVD_SOURCE)) return rc;
  if (anelastic)
    if (int rc = need(c, MONC_TVD_W, MONC_TVD_PROG)) return rc;
  if (baroclinicity) {
    if (int rc = need(c, MONC_TVD_U, MONC_TVD_PROG)) return rc;
    if (int rc = need(c, MONC_TVD_V, MONC_TVD_PROG)) return rc;
  }
  return c->precision == MONC_TVD_FP64
             ? thadv_impl<double>(c, anelastic, dthref, baroclinicity, thref, thref0, fcoriol, g, rate_x, rate_y,
                                  ugal, vgal)
             : thadv_impl<float>(c, anelastic, dthref, baroclinicity, thref, thref0, fcoriol, g, rate_x, rate_y,
                                 ugal, vgal);
}

int monc_pw_tendencies(monc_tvd_ctx* c, int advect_flow, int advect_th, int advect_q, double tcx, double tcy) {
  if (int rc = check_ctx(c)) return rc;
  if (!advect_flow && !advect_th && !advect_q) return MONC_TVD_OK;
  if (4 + c->n_q > MAX_FIELDS) return fail(MONC_TVD_ERR_ARG, "%d fields above the kernel's %d", 4 + c->n_q, MAX_FIELDS);
  if (!(tcx == tcx) || !(tcy == tcy)) return fail(MONC_TVD_ERR_ARG, "tcx, tcy must not be NaN");
  for (int f = MONC_TVD_U; f <= MONC_TVD_W; f++)
    if (int rc = need(c, f, MONC_TVD_PROG)) return rc;
  if (advect_flow)
    for (int f = MONC_TVD_U; f <= MONC_TVD_W; f++)
      if (int rc = need(c, f, MONC_TVD_SOURCE)) return rc;
  if (advect_th && c->active[MONC_TVD_TH])
    for (int v : {MONC_TVD_PROG, MONC_TVD_SOURCE})
      if (int rc = need(c, MONC_TVD_TH, v)) return rc;
  if (advect_q)
    for (int n = 0; n < c->n_q; n++)
      for (int v : {MONC_TVD_PROG, MONC_TVD_SOURCE})
        if (int rc = need(c, MONC_TVD_Q(n), v)) return rc;
  return c->precision == MONC_TVD_FP64 ? pw_impl<double>(c, advect_flow, advect_th, advect_q, tcx, tcy)
                                       : pw_impl<float>(c, advect_flow, advect_th, advect_q, tcx, tcy);
}

int monc_tvd_tangent_steps(monc_tvd_ctx* c, double dtm, int nsteps) {
  if (int rc = check_ctx(c)) return rc;
  if (c->precision != MONC_TVD_FP64) return fail(MONC_TVD_ERR_ARG, "the tangent-linear model runs in FP64 only");
  if (nsteps < 0) return fail(MONC_TVD_ERR_ARG, "nsteps %d < 0", nsteps);
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return tangent_steps_impl(c, dtm, nsteps);
}

int monc_tvd_forward_step_into(monc_tvd_ctx* c, double dtm, void* const* in, void* const* out) {
  if (int rc = check_ctx(c)) return rc;
  if (!in || !out) return fail(MONC_TVD_ERR_ARG, "null buffer array");
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return c->precision == MONC_TVD_FP64 ? forward_into_impl<double>(c, dtm, in, out)
                                       : forward_into_impl<float>(c, dtm, in, out);
}

int monc_tvd_forward_step_record(monc_tvd_ctx* c, double dtm, void* const* in, void* const* out,
                                 unsigned char* const* rec) {
  if (int rc = check_ctx(c)) return rc;
  if (!in || !out || !rec) return fail(MONC_TVD_ERR_ARG, "null buffer array");
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return c->precision == MONC_TVD_FP64 ? forward_record_impl<double>(c, dtm, in, out, rec)
                                       : forward_record_impl<float>(c, dtm, in, out, rec);
}

int monc_tvd_bind_record(monc_tvd_ctx* c, int field, const unsigned char* rec) {
  if (int rc = check_ctx(c)) return rc;
  if (field < 0 || field >= c->nfields) return fail(MONC_TVD_ERR_ARG, "field %d out of range", field);
  c->rec[field] = rec;
  return MONC_TVD_OK;
}

int monc_tvd_leapfrog_steps(monc_tvd_ctx* c, double dtm, double tsmth, double ugal, double vgal, int nsteps,
                            int first_forward) {
  if (int rc = check_ctx(c)) return rc;
  if (nsteps < 0) return fail(MONC_TVD_ERR_ARG, "nsteps %d < 0", nsteps);
  if (!std::isfinite(dtm) || !std::isfinite(tsmth) || !std::isfinite(ugal) || !std::isfinite(vgal))
    return fail(MONC_TVD_ERR_ARG, "dtm, tsmth, ugal and vgal must be finite");
  return c->precision == MONC_TVD_FP64 ? leapfrog_impl<double>(c, dtm, tsmth, ugal, vgal, nsteps, first_forward)
                                       : leapfrog_impl<float>(c, dtm, tsmth, ugal, vgal, nsteps, first_forward);
}

int monc_tvd_forward_steps(monc_tvd_ctx* c, double dtm, int nsteps) {
  if (int rc = check_ctx(c)) return rc;
  if (nsteps < 0) return fail(MONC_TVD_ERR_ARG, "nsteps %d < 0", nsteps);
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return c->precision == MONC_TVD_FP64 ? forward_steps_impl<double>(c, dtm, nsteps)
                                       : forward_steps_impl<float>(c, dtm, nsteps);
}

int monc_tvd_set_diagnostics(monc_tvd_ctx* c, int flags, const double* rprefrcp) {
  if (int rc = check_ctx(c)) return rc;
  if (flags & ~(MONC_TVD_DIAG_TENDENCY | MONC_TVD_DIAG_ZFACE))
    return fail(MONC_TVD_ERR_ARG, "unknown diagnostics flags 0x%x", flags);
  if (rprefrcp) {
    if (!c->rprefrcp) HIP_TRY(dev_alloc(&c->rprefrcp, sizeof(double) * c->nz));
    HIP_TRY(hipMemcpyAsync(c->rprefrcp, rprefrcp, sizeof(double) * c->nz, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  c->diag_flags = flags;
  return MONC_TVD_OK;
}

int monc_tvd_get_diagnostic(monc_tvd_ctx* c, int field, int which, double* host) {
  return monc_tvd_get_component_diagnostic(c, MONC_TVD_COMPONENT_TVD, field, which, host);
}

int monc_tvd_get_component_diagnostic(monc_tvd_ctx* c, int comp, int field, int which, double* host) {
  if (int rc = check_ctx(c)) return rc;
  if (!host) return fail(MONC_TVD_ERR_ARG, "null host pointer");
  if (comp < 0 || comp >= MONC_TVD_NCOMPONENTS) return fail(MONC_TVD_ERR_ARG, "unknown component %d", comp);
  if (which == MONC_TVD_DIAG_ZFACES && comp != MONC_TVD_COMPONENT_TVD)
    return fail(MONC_TVD_ERR_ARG, "z-face diagnostics belong to the TVD component only");
  const bool tabs = which == MONC_TVD_DIAG_TABS_3D || which == MONC_TVD_DIAG_TABS_PROFILE;
  if (!tabs && (field < 0 || field >= c->nfields))
    return fail(MONC_TVD_ERR_ARG, "field id %d out of range [0,%d)", field, c->nfields);
  const void* src = nullptr;
  size_t n = 0;
  bool fp64 = false;
  switch (which) {
    case MONC_TVD_DIAG_3D:
      if (!c->diag_have[comp][field]) break;
      src = c->diag_3d[comp][field];
      n = (size_t)c->nz * c->ny * c->nx;
      break;
    case MONC_TVD_DIAG_PROFILE:
      if (!c->diag_have[comp][field]) break;
      src = c->diag_prof[comp][field];
      n = c->nz;
      fp64 = true;
      break;
    case MONC_TVD_DIAG_TABS_3D:
      if (!c->tabs_have[comp]) break;
      src = c->tabs_3d[comp];
      n = (size_t)c->nz * c->ny * c->nx;
      break;
    case MONC_TVD_DIAG_TABS_PROFILE:
      if (!c->tabs_have[comp]) break;
      src = c->tabs_prof[comp];
      n = c->nz;
      fp64 = true;
      break;
    case MONC_TVD_DIAG_ZFACES:
      src = c->diag_zf[field];
      n = c->n_elem;
      break;
    default:
      return fail(MONC_TVD_ERR_ARG, "unknown diagnostic %d", which);
  }
  if (!src)
    return fail(MONC_TVD_ERR_STATE, "diagnostic %d of field %d (component %d) was never recorded", which, field, comp);
  if (fp64 || c->precision == MONC_TVD_FP64) {
    HIP_TRY(hipMemcpyAsync(host, src, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < n; q++) host[q] = (double)tmp[q];
  }
  return MONC_TVD_OK;
}

int monc_tvd_autotune(monc_tvd_ctx* c, double dtm, int reps, int* chosen) {
  if (int rc = check_ctx(c)) return rc;
  if (reps < 1 || reps > 100) return fail(MONC_TVD_ERR_ARG, "reps %d out of range [1,100]", reps);
  if (!std::isfinite(dtm)) return fail(MONC_TVD_ERR_ARG, "dtm must be finite");
  return c->precision == MONC_TVD_FP64 ? autotune_impl<double>(c, dtm, reps, chosen)
                                       : autotune_impl<float>(c, dtm, reps, chosen);
}

}  // extern "C"
