// Stage kernel: the step of one field kind (u, v, w, or the scalars advected
// by the same flows) per launch, with the raw fields staged in LDS.
//
// A block owns a band of 64 levels (lane = level) of R cell rows and marches
// along an x strip.  u, v, w and the field(s) the kind advects are staged in
// LDS as x planes: each plane holds rows j0-2 .. j0+R+1 and levels
// k0-2 .. k0+65 of one column, levels outside [1, nz] replicating the nearest
// level (the clamping of ultflx's stencils), so every operand of every face
// is an LDS read.  A ring of four planes holds columns i-1 .. i+2 while
// column i is computed; the plane of column i+3 is fetched into registers
// during the faces of column i and written into the ring slot of column i-1
// once the block has finished reading it.  (v and w of the scalar kinds are
// read at columns i and i+1 only: two planes, column i+2 written into the
// slot of column i, whose values the divergence takes before the barrier.)
// Each staged value thus crosses HBM once per launch (plus the 4/R row halo
// and the 4/64 level halo).
//
// Each of the R row waves computes the right x face, the north y face and the
// top z face of its cells (the face arithmetic of tvd_faces.hpp /
// tvd_march.hpp, ultimateflux.F90:81-614, with the same operand order); one
// more wave, the edge wave, computes the faces the row waves do not own: the
// south y face of the first row and the z face below the band (the top face
// of level k0-1) of every row.  Faces cross between waves through LDS,
// between lanes by a wave shift.  After one barrier every row wave forms the
// flux divergence (tvdadvection.F90:911-960) and writes the step, the
// tendency or the leapfrog update of its cells (the modes below).
#pragma once
#include "tvd_march.hpp"

namespace tvd {

constexpr int ST_KB = 68;    // staged levels: k0-2 .. k0+65
constexpr int ST_NMET = 65;  // metric levels: k0-1 .. k0+63

// Staged arrays: the flows, then the scalars (MODE_STEP advects each flow
// field by itself, so u, v and w are also the advected fields of their kinds).
enum { SA_U = 0, SA_V = 1, SA_W = 2, SA_S = 3 };

// Ring depth of staged array a (NA arrays): four x planes (columns i-1 .. i+2
// while column i is computed, column c in slot c & 3), except, with V2 (the
// scalar kinds), v and w, whose faces and divergence then read only columns i
// and i+1: two planes (slot c & 1).  The planes of all arrays lie one after
// the other, array by array.
template <int NA, bool V2>
__host__ __device__ constexpr int st_depth(int a) { return V2 && (a == SA_V || a == SA_W) ? 2 : 4; }
template <int NA, bool V2>
__host__ __device__ constexpr int st_plane0(int a) {
  return a <= SA_V ? 4 * a
                   : (a == SA_W ? 4 + st_depth<NA, V2>(SA_V)
                                : 4 + st_depth<NA, V2>(SA_V) + st_depth<NA, V2>(SA_W) + 4 * (a - SA_S));
}
template <int NA, bool V2>
__host__ __device__ constexpr int st_planes() { return st_plane0<NA, V2>(NA); }
// The column array a stages when the arrays of depth 4 stage column c4 and
// those of depth 2 column c2, and the ring plane that column goes to.
template <int NA, bool V2>
__device__ __forceinline__ int st_col(int a, int c4, int c2) {
  if constexpr (V2) return st_depth<NA, V2>(a) == 2 ? c2 : c4;
  else return c4;
}
template <int NA, bool V2>
__device__ __forceinline__ int st_plane(int a, int c) {
  return st_plane0<NA, V2>(a) + (c & (st_depth<NA, V2>(a) - 1));
}

// LDS of one block for a field kind with NSF fields.  The metric table comes
// first: the views of the edge wave start a row or a level before the ring
// (their stencils reach only later elements).
template <typename T, int R, int NA, int NSF, bool V2>
struct StageLds {
  T met[M_COUNT][ST_NMET];  // the kind's metric roles (w: swapped), levels k0-1 .. k0+63
  T ring[st_planes<NA, V2>()][R + 4][ST_KB];
  T yn[NSF][R + 1][64];     // north y face of rows j0-1 .. j0+R-1
  T zb[NSF][R];             // top face of level k0-1 of each row
};

// Read access to the ring: array a at level k, row j, column i (i uniform,
// within the four staged columns).  Addresses are formed from the lane's own
// position (k_at, j_at) so that the constant level and row shifts of the
// stencils fold into the instruction's offset; only the column's ring slot is
// added per access.
template <typename T, int R, int NA, bool V2>
struct SView {
  static constexpr int PLANE = (R + 4) * ST_KB;  // elements per x plane
  const T* base;    // element (row j_at-2, level k_at-2) of plane 0
  int k_at, j_at;   // the lane's own level and row
  int j0;           // first row of the block
  __device__ __forceinline__ SView(const T* ring, int k0, int j0_, int k, int j)
      : base(ring + (j - j0_) * ST_KB + (k - k0)), k_at(k), j_at(j), j0(j0_) {}
  // every stencil offset is >= -2 in rows and levels, so the constant part
  // of an index is non-negative and folds into the LDS instruction's offset
  __device__ __forceinline__ T operator()(int a, int k, int j, int i) const {
    return base[(st_plane0<NA, V2>(a) + (i & (st_depth<NA, V2>(a) - 1))) * PLANE + (j - j_at + 2) * ST_KB +
                (k - k_at + 2)];
  }
};

// A flow as the field kind sees it (stencil.F90:147-169): raw for scalars,
// the dual-grid interpolation 0.5*(f + f shifted along u's x, v's y or w's z)
// for momentum.  At k = nz the level above replicates nz, so w's
// interpolation there is 0.5*(f + f) = f, the reference's raw copy.
template <typename T, int KIND, int R, int NA, bool V2>
__device__ __forceinline__ T sflow(const SView<T, R, NA, V2>& s, int a, int k, int j, int i) {
  const T x = s(a, k, j, i);
  if constexpr (KIND == KIND_U) return T(0.5) * (x + s(a, k, j, i + 1));
  else if constexpr (KIND == KIND_V) return T(0.5) * (x + s(a, k, j + 1, i));
  else if constexpr (KIND == KIND_W) return T(0.5) * (x + s(a, k + 1, j, i));
  else return x;
}

// The metric factors of one level of the field kind (tabulated per block).
template <typename T>
struct SMet {
  T rdz_k, rdz_kd, zs_up, zs_dn, rdc_up, rdc_dn, rdu_up, rdu_dn, rdd, dd2, t1, t2;
};

template <typename T, class LDS>
__device__ __forceinline__ SMet<T> smet(const LDS& lds, int li) {
  SMet<T> m;
  m.rdz_k = lds.met[M_RDZ][li];
  m.rdz_kd = lds.met[M_RDZKD][li];
  m.zs_up = lds.met[M_ZS_UP][li];
  m.zs_dn = lds.met[M_ZS_DN][li];
  m.rdc_up = lds.met[M_RDC_UP][li];
  m.rdc_dn = lds.met[M_RDC_DN][li];
  m.rdu_up = lds.met[M_RDU_UP][li];
  m.rdu_dn = lds.met[M_RDU_DN][li];
  m.rdd = lds.met[M_RDD][li];
  m.dd2 = lds.met[M_DD2][li];
  m.t1 = lds.met[M_T1][li];
  m.t2 = lds.met[M_T2][li];
  return m;
}

// Right x face of column (j, i) at level k (ultimateflux.F90:102-165) for the
// NS fields a0 .. a0+NS-1 of kind KIND; kq = (min(k+1, nz)) - 1.
template <typename T, int KIND, int NS, int R, int NA, bool V2>
__device__ __forceinline__ void sx_face(const SView<T, R, NA, V2>& s, int a0, int k, int kq, int j, int i, T rdz_k,
                                        T rdz_kd, T dt, T cx, T cy, T out[NS], int cls[NS]) {
  const T u_m = sflow<T, KIND>(s, SA_U, k, j, i - 1), u_0 = sflow<T, KIND>(s, SA_U, k, j, i);
  const T u_p = sflow<T, KIND>(s, SA_U, k, j, i + 1);
  const T v_0 = sflow<T, KIND>(s, SA_V, k, j, i), v_s = sflow<T, KIND>(s, SA_V, k, j - 1, i);
  const T v_0p = sflow<T, KIND>(s, SA_V, k, j, i + 1), v_sp = sflow<T, KIND>(s, SA_V, k, j - 1, i + 1);
  const T w_0 = sflow<T, KIND>(s, SA_W, k, j, i), w_p = sflow<T, KIND>(s, SA_W, k, j, i + 1);
  const T wm_0 = sflow<T, KIND>(s, SA_W, k - 1, j, i), wm_p = sflow<T, KIND>(s, SA_W, k - 1, j, i + 1);
  const bool upos = u_0 > T(0.0);
  T sm = rdz_k * (pos(upos ? w_0 : w_p) + negabs(upos ? wm_0 : wm_p));
  sm = sm + cy * (pos(upos ? v_0 : v_0p) + negabs(upos ? v_s : v_sp));
  sm = sm + cx * (upos ? u_0 + negabs(u_m) : pos(u_p) - u_0);
  sm = sm * dt;
  const T sz = T(0.125) * (w_0 + wm_0 + w_p + wm_p);
  const bool pz = sign_pos(sz);
  const T cz = fabs(sz) * rdz_kd * dt;
  const T sy = T(0.125) * (v_0p + v_0 + v_sp + v_s);
  const int js = sign_pos(sy) ? j - 1 : j + 1;
  const T cyt = fabs(sy) * cy * dt;
  const HLim<T> lim(u_0, cx, dt, sm);
  const bool ap = sign_pos(u_0);
  const int ksl = pz ? k : kq, ksu = pz ? k - 1 : k + 1;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const int a = a0 + n;
    const T fm = s(a, k, j, i - 1), fa = s(a, k, j, i), fb = s(a, k, j, i + 1), f2 = s(a, k, j, i + 2);
    const T fu = ap ? fm : f2;
    const T fc = ap ? fa : fb;
    const T fd = ap ? fb : fa;
    const T d0 = s(a, ksl, j, i) - s(a, ksu, j, i);
    const T d1 = s(a, ksl, j, i + 1) - s(a, ksu, j, i + 1);
    const T g1 = cz * (ap ? d0 : d1);
    const T t = ap ? s(a, k, js, i) : s(a, k, js, i + 1);
    const T g2 = cyt * (fc - t);
    out[n] = lim.face_cls(fu, fc, fd, g1, g2, cls[n]);
  }
}

// North y face of row j at column i, level k (ultimateflux.F90:189-233).
template <typename T, int KIND, int NS, int R, int NA, bool V2>
__device__ __forceinline__ void sy_face(const SView<T, R, NA, V2>& s, int a0, int k, int kq, int j, int i, T rdz_k,
                                        T rdz_kd, T dt, T cx, T cy, T out[NS], int cls[NS]) {
  const T vk = sflow<T, KIND>(s, SA_V, k, j, i), vn = sflow<T, KIND>(s, SA_V, k, j + 1, i);
  const T vs = sflow<T, KIND>(s, SA_V, k, j - 1, i);
  const T w0 = sflow<T, KIND>(s, SA_W, k, j, i), w0m = sflow<T, KIND>(s, SA_W, k - 1, j, i);
  const T w1 = sflow<T, KIND>(s, SA_W, k, j + 1, i), w1m = sflow<T, KIND>(s, SA_W, k - 1, j + 1, i);
  const T u0 = sflow<T, KIND>(s, SA_U, k, j, i), u0m = sflow<T, KIND>(s, SA_U, k, j, i - 1);
  const T u1 = sflow<T, KIND>(s, SA_U, k, j + 1, i), u1m = sflow<T, KIND>(s, SA_U, k, j + 1, i - 1);
  const bool vpos = sign_pos(vk);
  T sm = rdz_k * (pos(vpos ? w0 : w1) + negabs(vpos ? w0m : w1m));
  sm = sm + cy * (pos(vpos ? vk : vn) + negabs(vpos ? vs : vk));
  sm = sm + cx * (pos(vpos ? u0 : u1) + negabs(vpos ? u0m : u1m));
  sm = sm * dt;
  const T sz = T(0.125) * (w0 + w0m + w1 + w1m);
  const bool pz = sign_pos(sz);
  const int jj = vpos ? j : j + 1;  // j + jofset
  const int ksl = pz ? k : kq, ksu = pz ? k - 1 : k + 1;
  const T cz = fabs(sz) * rdz_kd * dt;
  const T sx = T(0.125) * (u0m + u0 + u1m + u1);
  const int it = sign_pos(sx) ? i - 1 : i + 1;
  const T cxt = fabs(sx) * cx * dt;
  const HLim<T> lim(vk, cy, dt, sm);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const int a = a0 + n;
    const T fa = s(a, k, j, i), fb = s(a, k, j + 1, i);
    const T fs = s(a, k, j - 1, i), f2 = s(a, k, j + 2, i);
    const T fu = vpos ? fs : f2;
    const T fc = vpos ? fa : fb;
    const T fd = vpos ? fb : fa;
    const T g1 = cz * (s(a, ksl, jj, i) - s(a, ksu, jj, i));
    const T g2 = cxt * (fc - s(a, k, jj, it));
    out[n] = lim.face_cls(fu, fc, fd, g1, g2, cls[n]);
  }
}

// Top z face of level k (flux_z(k+1)), ultimateflux.F90:236-523, the three
// branches in one as in mz_face: `edge` marks k = 1 and k = nz-1, `top`
// k = nz-1, `bot` k = 1.
template <typename T, int KIND, int NS, int R, int NA, bool V2>
__device__ __forceinline__ void sz_face(const SView<T, R, NA, V2>& s, int a0, int k, int nz, int j, int i,
                                        const SMet<T>& m, T dt, T cx, T cy, T out[NS], int cls[NS]) {
  const bool edge = k == 1 || k == nz - 1, top = k == nz - 1, bot = k == 1;
  const T wk = sflow<T, KIND>(s, SA_W, k, j, i);
  const T wm_0 = sflow<T, KIND>(s, SA_W, k - 1, j, i), wk_0 = sflow<T, KIND>(s, SA_W, k + 1, j, i);
  const T u_m = sflow<T, KIND>(s, SA_U, k, j, i - 1), u_0 = sflow<T, KIND>(s, SA_U, k, j, i);
  const T v_0 = sflow<T, KIND>(s, SA_V, k, j, i), v_s = sflow<T, KIND>(s, SA_V, k, j - 1, i);
  const T uk_m = sflow<T, KIND>(s, SA_U, k + 1, j, i - 1), uk_0 = sflow<T, KIND>(s, SA_U, k + 1, j, i);
  const T vk_s = sflow<T, KIND>(s, SA_V, k + 1, j - 1, i), vk_0 = sflow<T, KIND>(s, SA_V, k + 1, j, i);
  const bool up = edge ? wk >= T(0.0) : sign_pos(wk);
  const T wa = up ? pos(wk) : (top ? T(0.0) : pos(wk_0));
  const T wb = up ? (bot ? T(0.0) : negabs(wm_0)) : negabs(wk);
  T sm = (up ? m.zs_up : m.zs_dn) * (wa + wb);
  sm = sm + cy * (pos(up ? v_0 : vk_0) + negabs(up ? v_s : vk_s));
  sm = sm + cx * (pos(up ? u_0 : uk_0) + negabs(up ? u_m : uk_m));
  sm = sm * dt;
  const T sy = T(0.125) * (v_s + v_0 + vk_s + vk_0);
  const T cyt = fabs(sy) * cy * dt;
  const T sx = T(0.125) * (u_m + u_0 + uk_m + uk_0);
  const T cxt = fabs(sx) * cx * dt;
  const int kk = up ? k : k + 1;
  const int js = sign_pos(sy) ? j - 1 : j + 1;
  const int ix = sign_pos(sx) ? i - 1 : i + 1;
  const T rdc = up ? m.rdc_up : m.rdc_dn;
  const T rdu = up ? m.rdu_up : m.rdu_dn;
  const T rdd = m.rdd;
  const T r6 = T(1.0 / 6.0);
  const T cflmod = fabs(wk * rdd * dt);
  const T hrcfl = recip(sm + T(1.e-30));
  const T hcfl = T(0.5) * cflmod;
  const T cflcoef = r6 * (T(1.0) - cflmod * cflmod) * (m.dd2 * rdc);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const int a = a0 + n;
    const T fk = s(a, k, j, i), fk1 = s(a, k + 1, j, i), fkm = s(a, k - 1, j, i), fk2 = s(a, k + 2, j, i);
    const T fdn = top ? T(2.0) * fk1 - fk : fk2;
    const T fu = up ? fkm : fdn;
    const T fc = up ? fk : fk1;
    const T fd = up ? fk1 : fk;
    const T g1 = cyt * (fc - s(a, kk, js, i));
    const T g2 = cxt * (fc - s(a, kk, j, ix));
    const T fcurvs = fd - T(2.0) * fc + fu;
    const T fdels = fd - fu;
    const T rfdels = recip(guard(fdels));
    const T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * ((fd - fc) * rdd - (fc - fu) * rdu) - g1 - g2;
    const T fnfs = (ffs - fu) * rfdels;
    const T fncs = (fc - fu) * rfdels;
    const T fnrefs = fncs * hrcfl;
    const T ftemp = tmax(fncs, tmin(tmin(fnfs, fnrefs), T(1.0)));
    const T lim = ftemp * fdels + fu;
    out[n] = fabs(fcurvs) >= fabs(fdels) ? fc : lim;
    cls[n] = lim_class(fcurvs, fdels, fnfs, fncs, fnrefs);
  }
}

// The staged array a kind advects: its own flow, or the fourth array (the
// scalars, or a flow kind's separate advected field).
template <int KIND, int NA>
constexpr int skind_first() { return KIND == KIND_SCALAR || NA > 3 ? SA_S : KIND; }

// Faces of one field kind for a row wave at (k, j), column i: right x, north
// y and top z face of each field of the kind; the z face only where the kind
// has one (scalars, u, v: 2 <= k <= nz-1; w: 1 <= k <= nz-1).  code[n]: the
// limiter arms the three faces took, x | y << 2 | z << 4 (the recording
// step's byte; unused otherwise).
template <typename T, int KIND, int NSF, int R, int NA, bool V2>
__device__ __forceinline__ void skind_faces(const SView<T, R, NA, V2>& s, const StageLds<T, R, NA, NSF, V2>& lds, int k, int kq,
                                            int nz, int j, int i, int li, T dt, T cx, T cy, T fxr[NSF], T fyn[NSF],
                                            T fzt[NSF], int code[NSF]) {
  constexpr int A0 = skind_first<KIND, NA>();
  const SMet<T> m = smet<T>(lds, li);
  int ax[NSF], ay[NSF], az[NSF];
  sx_face<T, KIND, NSF>(s, A0, k, kq, j, i, m.rdz_k, m.rdz_kd, dt, cx, cy, fxr, ax);
  sy_face<T, KIND, NSF>(s, A0, k, kq, j, i, m.rdz_k, m.rdz_kd, dt, cx, cy, fyn, ay);
  T zf[NSF];
  sz_face<T, KIND, NSF>(s, A0, k, nz, j, i, m, dt, cx, cy, zf, az);
#pragma unroll
  for (int n = 0; n < NSF; n++) code[n] = ax[n] | (ay[n] << 2) | (az[n] << 4);
  const bool has_z = KIND == KIND_W ? (k >= 1 && k <= nz - 1) : (k >= 2 && k <= nz - 1);
#pragma unroll
  for (int n = 0; n < NSF; n++) fzt[n] = has_z ? zf[n] : T(0);
}

// Divergence of the fields of one kind at (k, j, i) (tvdadvection.F90:
// 925-959).  MODE_STEP and MODE_REC, the forward step (stepfields.F90:
// 597-627): writes f + dtm*s at the cell and its periodic halo images; levels the reference
// leaves untouched (k = 1; k = nz for w) keep f.  MODE_ACCUM, the drop-in
// call: adds s to the source field `out` at levels 2..nz (w: 2..nz-1),
// overwriting it at k = nz for u, v and the scalars (:942-959).
// The flows the divergence of cell (k, j, i) reads at its own column (ur:
// the right x face; vs, vn: south and north y faces; wb, wk: bottom and top z
// faces).
template <typename T>
struct SDivFlows {
  T ur, vs, vn, wb, wk;
};

template <typename T, int KIND, int R, int NA, bool V2>
__device__ __forceinline__ SDivFlows<T> sdiv_flows(const SView<T, R, NA, V2>& s, int k, int j, int i) {
  return {sflow<T, KIND>(s, SA_U, k, j, i), sflow<T, KIND>(s, SA_V, k, j - 1, i), sflow<T, KIND>(s, SA_V, k, j, i),
          sflow<T, KIND>(s, SA_W, k - 1, j, i), sflow<T, KIND>(s, SA_W, k, j, i)};
}

// MODE_LEAP: the current fields (f), their z-fields (z) and the buffers of
// the next z-fields (o2), with each field's galilean shift.
template <typename T, int NSF>
struct SLeap {
  __amdgpu_buffer_rsrc_t f[NSF], z[NSF], o2[NSF];
  T gal[NSF];
};

template <typename T, int KIND, int NSF, int R, int NA, bool V2, int MODE>
__device__ __forceinline__ void skind_step(const SView<T, R, NA, V2>& s, const StageLds<T, R, NA, NSF, V2>& lds,
                                           const __amdgpu_buffer_rsrc_t (&out)[NSF], const Geo& g,
                                           const LaunchArgs<T>& L, const SLeap<T, NSF>& lp, int k, int j, int i,
                                           int li, int lane, int r, T dtm, T cx, T cy, T ul, const SDivFlows<T>& q,
                                           const T fxl[NSF], const T fxr[NSF], const T fyn[NSF],
                                           const T fzt[NSF]) {
  const Dims& d = L.d;
  constexpr int A0 = skind_first<KIND, NA>();
  constexpr bool IS_W = KIND == KIND_W;
  const int nz = d.nz, ny = d.ny, nx = d.nx;
  const T ur = q.ur, vs = q.vs, vn = q.vn, wb = q.wb, wk = q.wk;
  const T t1 = lds.met[M_T1][li], t2 = lds.met[M_T2][li];
  const T wt = k < nz ? wk : T(0);
  const int kmax = IS_W ? nz - 1 : nz;
  const bool cell = k >= 2 && k <= kmax;
  const bool top = !IS_W && k == nz;
  const unsigned vo = (unsigned)(k - 1) * (unsigned)sizeof(T);
  const unsigned so = g.so(j, i);
  if constexpr (MODE == MODE_ACCUM) {
    // the source read is clamped to the column (lanes past nz store nothing)
    const unsigned vc = (unsigned)((k < nz ? k : nz) - 1) * (unsigned)sizeof(T);
#pragma unroll
    for (int n = 0; n < NSF; n++) {
      const T fys = lds.yn[n][r][lane];
      T fzb = __shfl_up(fzt[n], 1);
      if (lane == 0) fzb = lds.zb[n][r];
      const T s0 = ldb<T>(out[n], vc, so);
      T sv;
      if (top) {
        sv = (vs * fys - vn * fyn[n]) * cy;
        sv = sv + T(4.0) * t1 * wb * fzb;
        sv = sv + (ul * fxl[n] - ur * fxr[n]) * cx;
      } else {
        sv = s0 + (vs * fys - vn * fyn[n]) * cy;
        sv = sv + T(4.0) * (wb * fzb * t1 - wt * fzt[n] * t2);
        sv = sv + (ul * fxl[n] - ur * fxr[n]) * cx;
      }
      if (cell) stb<T>(out[n], sv, vo, so);
    }
    return;
  }
  const int jh = (j <= 4) ? j + ny : (j >= ny + 1 ? j - ny : 0);
  const int ih = (i <= 4) ? i + nx : (i >= nx + 1 ? i - nx : 0);
  const unsigned soj = g.so(jh ? jh : j, i), soi = g.so(j, ih ? ih : i), soji = g.so(jh ? jh : j, ih ? ih : i);
#pragma unroll
  for (int n = 0; n < NSF; n++) {
    const int a = A0 + n;
    const T fys = lds.yn[n][r][lane];
    T fzb = __shfl_up(fzt[n], 1);
    if (lane == 0) fzb = lds.zb[n][r];
    T sv;
    if (top) {
      sv = (vs * fys - vn * fyn[n]) * cy;
      sv = sv + T(4.0) * t1 * wb * fzb;
      sv = sv + (ul * fxl[n] - ur * fxr[n]) * cx;
    } else {
      sv = (vs * fys - vn * fyn[n]) * cy;
      sv = sv + T(4.0) * (wb * fzb * t1 - wt * fzt[n] * t2);
      sv = sv + (ul * fxl[n] - ur * fxr[n]) * cx;
    }
    if constexpr (MODE == MODE_LEAP) {
      // the fused MONC step (leap_update): the current field and its
      // z-field from memory (the staged field is the advected one, the
      // z-field in a centred step), reads clamped to the column
      const unsigned vc = (unsigned)((k < nz ? k : nz) - 1) * (unsigned)sizeof(T);
      const T fc = ldb<T>(lp.f[n], vc, so);
      const T zc = L.forward ? T(0) : ldb<T>(lp.z[n], vc, so);
      T fn, zn;
      leap_update<T, KIND>(L, lp.gal[n], fc, zc, cell ? sv : T(0), fn, zn);
      if (k <= nz) {
        stb<T>(out[n], fn, vo, so);
        stb<T>(lp.o2[n], zn, vo, so);
        if (jh) {
          stb<T>(out[n], fn, vo, soj);
          stb<T>(lp.o2[n], zn, vo, soj);
        }
        if (ih) {
          stb<T>(out[n], fn, vo, soi);
          stb<T>(lp.o2[n], zn, vo, soi);
        }
        if (jh && ih) {
          stb<T>(out[n], fn, vo, soji);
          stb<T>(lp.o2[n], zn, vo, soji);
        }
      }
    } else {
      const T fcur = s(a, k, j, i);
      const T val = cell ? fcur + dtm * sv : fcur;
      if (k <= nz) {
        stb<T>(out[n], val, vo, so);
        if (jh) stb<T>(out[n], val, vo, soj);
        if (ih) stb<T>(out[n], val, vo, soi);
        if (jh && ih) stb<T>(out[n], val, vo, soji);
      }
    }
  }
}

// Staging of one x plane (column c): pairs (array, row) are dealt to the
// waves round-robin; a pair is one 64-lane load of levels k0 .. k0+63 and
// four edge levels (k0-2, k0-1, k0+64, k0+65) gathered by the wave's low
// lanes.  Levels are clamped to [1, nz]; rows past ny+4 read as 0.
template <int R, int NA>
struct StagePlan {
  static constexpr int PAIRS = NA * (R + 4);
  static constexpr int WAVES = R + 1;
  static constexpr int PER = (PAIRS + WAVES - 1) / WAVES;  // pairs per wave (upper bound)
};

template <typename T, int R, int NA>
struct StageFetch {
  T main[StagePlan<R, NA>::PER];
  T edge;
};

// The loads of a plane are all in flight together; only the store into the
// ring waits for them.  Pairs past the last are clamped to it (their loads are
// redundant and not stored), rows past ny+4 to row ny+4 (no active cell's
// stencil reaches them).  The four edge levels of the wave's q-th pair go to
// lanes 4q .. 4q+3 of one register.  The arrays of depth 4 stage column c4,
// those of depth 2 column c2.
template <typename T, int R, int NA, bool V2>
__device__ __forceinline__ void stage_fetch(StageFetch<T, R, NA>& f, const __amdgpu_buffer_rsrc_t (&src)[NA],
                                            const Geo& g, int wv, int lane, int k0, int j0, int nz, int ny, int c4,
                                            int c2) {
  using P = StagePlan<R, NA>;
  const unsigned es = sizeof(T);
  const unsigned lev = (unsigned)(min(k0 + lane, nz) - 1) * es;
  const int e = lane & 3, kk = e < 2 ? k0 - 2 + e : k0 + 62 + e;
  const unsigned elev = (unsigned)((kk < 1 ? 1 : (kk > nz ? nz : kk)) - 1) * es;
  f.edge = T(0);
#pragma unroll
  for (int q = 0; q < P::PER; q++) {
    // the pair is uniform over the wave: its array selects a buffer resource
    const int p = min((int)__builtin_amdgcn_readfirstlane(wv + q * P::WAVES), P::PAIRS - 1);
    const int a = p / (R + 4), row = p % (R + 4);
    const unsigned sc = (unsigned)(st_col<NA, V2>(a, c4, c2) - 1) * g.sib;
    const unsigned jo = (unsigned)(min(j0 - 2 + row, ny + 4) - 1) * g.sjb;
    __amdgpu_buffer_rsrc_t r = src[0];
#pragma unroll
    for (int x = 1; x < NA; x++) r = a == x ? src[x] : r;
    f.main[q] = ldb<T>(r, lev + jo, sc);
    if ((lane >> 2) == q) f.edge = ldb<T>(r, elev + jo, sc);
  }
}

template <typename T, int R, int NA, bool V2, class LDS>
__device__ __forceinline__ void stage_store(const StageFetch<T, R, NA>& f, LDS& lds, int wv,
                                            int lane, int c4, int c2) {
  using P = StagePlan<R, NA>;
#pragma unroll
  for (int q = 0; q < P::PER; q++) {
    const int p = wv + q * P::WAVES;
    if (p < P::PAIRS) {
      const int a = p / (R + 4);
      lds.ring[st_plane<NA, V2>(a, st_col<NA, V2>(a, c4, c2))][p % (R + 4)][2 + lane] = f.main[q];
    }
  }
  const int q = lane >> 2, e = lane & 3;
  const int p = wv + q * P::WAVES;
  if (q < P::PER && p < P::PAIRS) {
    const int a = p / (R + 4);
    lds.ring[st_plane<NA, V2>(a, st_col<NA, V2>(a, c4, c2))][p % (R + 4)][e < 2 ? e : 64 + e] = f.edge;
  }
}

// The edge wave's faces of one field kind: the south y face of the block's
// first row (the north face of row j0-1) for the 64 levels, and the z face
// below the band (top face of level k0-1) of each row (lane r = row j0+r).
// The recording step (MODE_REC) writes the arm of the south face of row 3 as
// the y arm of owner (k, 2, i), the one face row 2 owns; the faces below the
// band belong to the band below, which records them.
template <typename T, int KIND, int NSF, int R, int NA, bool V2, int MODE>
__device__ __forceinline__ void skind_edge(const SView<T, R, NA, V2>& s, StageLds<T, R, NA, NSF, V2>& lds,
                                           const FieldSet<T>& G, const Dims& d, int lane, int k0, int nz, int ny,
                                           int i, T dt, T cx, T cy) {
  constexpr int A0 = skind_first<KIND, NA>();
  const int k = k0 + lane, kq = k < nz ? k : nz - 1;
  const int j = s.j0 - 1;
  const SView<T, R, NA, V2> se(&lds.ring[0][0][0], k0, s.j0, k, j);
  {
    const T rdz_k = lds.met[M_RDZ][lane + 1], rdz_kd = lds.met[M_RDZKD][lane + 1];
    T f[NSF];
    int ay[NSF];
    sy_face<T, KIND, NSF>(se, A0, k, kq, j, i, rdz_k, rdz_kd, dt, cx, cy, f, ay);
#pragma unroll
    for (int n = 0; n < NSF; n++) lds.yn[n][0][lane] = f[n];
    if constexpr (MODE == MODE_REC) {
      if (j == 2 && k <= nz) {
#pragma unroll
        for (int n = 0; n < NSF; n++) G.rec[n][at(d, k, 2, i)] = (unsigned char)(ay[n] << 2);
      }
    }
  }
  if (k0 > 1 && lane < R && s.j0 + lane <= ny + 2) {
    const SMet<T> m = smet<T>(lds, 0);
    const SView<T, R, NA, V2> sz(&lds.ring[0][0][0], k0, s.j0, k0 - 1, s.j0 + lane);
    T f[NSF];
    int az[NSF];
    sz_face<T, KIND, NSF>(sz, A0, k0 - 1, nz, s.j0 + lane, i, m, dt, cx, cy, f, az);
#pragma unroll
    for (int n = 0; n < NSF; n++) lds.zb[n][lane] = f[n];
  }
}

// One launch of the step of one field kind (u, v, w, or NS scalars advected
// by the same flows): a 1-D grid of (row tile, band, strip) tiles, row tiles
// fastest so that consecutive tiles of one XCD share halo rows in its L2
// (blocks are dealt to the 8 XCDs round-robin).  MODE_STEP with every advected
// field its own prognostic, MODE_REC (the same step, also writing each
// owner's limiter arms to G.rec), MODE_ACCUM (the drop-in tendency call)
// with the flow kinds advecting their own prognostic, or MODE_LEAP (the
// fused MONC timestep: the advected fields F, the current fields f and
// z-fields z, the next fields out and z-fields out2); G holds the kind's
// fields in order.
template <typename T, int KIND, int NS, int R, int WPE, int MODE>
__global__ void __launch_bounds__(64 * (R + 1)) __attribute__((amdgpu_waves_per_eu(WPE)))
stage_kernel(LaunchArgs<T> L, FieldSet<T> G, int strip, int nbands, int tiles_y) {
  constexpr bool SC = KIND == KIND_SCALAR;
  // a flow kind with NS = 1 advects a field other than its staged flow (a
  // z-field): staged as a fourth array
  constexpr int NA = 3 + NS;
  constexpr bool V2 = SC;
  constexpr int NSF = SC ? NS : 1;
  constexpr int A0 = skind_first<KIND, NA>();
  constexpr int W = KIND == KIND_W ? 1 : 0;
  __shared__ StageLds<T, R, NA, NSF, V2> lds;
  const int nb = (int)gridDim.x, bid = (int)blockIdx.x;
  const int full = nb - nb % 8;
  const int t = bid < full ? (bid % 8) * (full / 8) + bid / 8 : bid;
  const int ty = t % tiles_y, band = (t / tiles_y) % nbands, tx = t / (tiles_y * nbands);
  const int nz = L.d.nz, ny = L.d.ny, nx = L.d.nx;
  const int lane = threadIdx.x;
  const int wv = (int)__builtin_amdgcn_readfirstlane(threadIdx.y);
  const int k0 = 1 + 64 * band, j0 = 3 + ty * R;
  const int i_begin = 3 + tx * strip, i_end = min(i_begin + strip, nx + 3);
  const T dt = G.dt[0], dtm = L.dtm, cx = L.cx, cy = L.cy;
  const unsigned es = sizeof(T);
  Geo g;
  g.sjb = (unsigned)L.d.sj * es;
  g.sib = (unsigned)L.d.si * es;
  __amdgpu_buffer_rsrc_t src[NA], out[NSF];
  src[SA_U] = rsrc(L.u);
  src[SA_V] = rsrc(L.v);
  src[SA_W] = rsrc(L.w);
#pragma unroll
  for (int n = 0; n < NA - 3; n++) src[SA_S + n] = rsrc(G.F[n]);
#pragma unroll
  for (int n = 0; n < NSF; n++) out[n] = rsrc(G.out[n]);
  SLeap<T, NSF> lp{};
  if constexpr (MODE == MODE_LEAP) {
#pragma unroll
    for (int n = 0; n < NSF; n++) {
      lp.f[n] = rsrc(G.f[n]);
      lp.z[n] = rsrc(G.z[n]);
      lp.o2[n] = rsrc(G.out2[n]);
      lp.gal[n] = G.gal[n];
    }
  }
  // metric table of the kind, levels k0-1 .. k0+63
  for (int e = lane + 64 * wv; e < M_COUNT * ST_NMET; e += 64 * (R + 1)) {
    const int mm = e / ST_NMET, kk = k0 - 1 + e % ST_NMET;
    lds.met[mm][e % ST_NMET] = met_value<T>(mm, kk, nz, W, W ? L.mw : L.mc);
  }
  StageFetch<T, R, NA> fetch;
  // the planes the faces of column i_begin-1 read: i_begin-2 .. i_begin+1
  // (arrays of depth 2: i_begin-1 and i_begin, the last staged twice)
  for (int c = i_begin - 2; c <= i_begin + 1; c++) {
    stage_fetch<T, R, NA, V2>(fetch, src, g, wv, lane, k0, j0, nz, ny, c, min(c, i_begin));
    stage_store<T, R, NA, V2>(fetch, lds, wv, lane, c, min(c, i_begin));
  }
  __syncthreads();
  const bool row_wave = wv < R;
  const int r = wv, j = j0 + r;
  const bool active = row_wave && j <= ny + 2;  // wave-uniform
  const int k = k0 + lane, kq = k < nz ? k : nz - 1, li = lane + 1;
  // the row waves' view; the edge wave forms its own for its faces
  const SView<T, R, NA, V2> sv(&lds.ring[0][0][0], k0, j0, k, row_wave ? j : j0 - 1);
  T fxl[NSF], fxr[NSF], fyn[NSF], fzt[NSF], ul = T(0);
  SDivFlows<T> dq{};
#pragma unroll
  for (int n = 0; n < NSF; n++) fxl[n] = fxr[n] = fyn[n] = fzt[n] = T(0);
  if (active) {  // the left x faces of column i_begin
    int ax[NSF];
    sx_face<T, KIND, NSF>(sv, A0, k, kq, j, i_begin - 1, lds.met[M_RDZ][li], lds.met[M_RDZKD][li], dt, cx, cy,
                          fxl, ax);
    if constexpr (MODE == MODE_REC) {  // the right x face of column 2 (owner (k, j, 2))
      if (i_begin == 3 && k <= nz) {
#pragma unroll
        for (int n = 0; n < NSF; n++) G.rec[n][at(L.d, k, j, 2)] = (unsigned char)ax[n];
      }
    }
  }
  __syncthreads();  // planes i_begin-2 (depth 2: i_begin-1) are no longer read
  stage_fetch<T, R, NA, V2>(fetch, src, g, wv, lane, k0, j0, nz, ny, i_begin + 2, i_begin + 1);
  stage_store<T, R, NA, V2>(fetch, lds, wv, lane, i_begin + 2, i_begin + 1);
  __syncthreads();
  for (int i = i_begin; i < i_end; i++) {
    // planes up to i_end+1 are read (depth 2: up to i_end)
    const bool more = i + 3 <= i_end + 1;
    if (more) stage_fetch<T, R, NA, V2>(fetch, src, g, wv, lane, k0, j0, nz, ny, i + 3, i + 2);
    if (row_wave) {
      if (active) {
        int code[NSF];
        skind_faces<T, KIND, NSF>(sv, lds, k, kq, nz, j, i, li, dt, cx, cy, fxr, fyn, fzt, code);
        ul = sflow<T, KIND>(sv, SA_U, k, j, i - 1);  // its plane is overwritten before the step
        // so are the planes of v and w at column i where they are two deep
        if constexpr (V2) dq = sdiv_flows<T, KIND>(sv, k, j, i);
        if constexpr (MODE == MODE_REC) {  // the arms of the faces cell (k, j, i) owns
          if (k <= nz) {
#pragma unroll
            for (int n = 0; n < NSF; n++) G.rec[n][at(L.d, k, j, i)] = (unsigned char)code[n];
          }
        }
      }
#pragma unroll
      for (int n = 0; n < NSF; n++) lds.yn[n][r + 1][lane] = fyn[n];
    } else {
      skind_edge<T, KIND, NSF, R, NA, V2, MODE>(sv, lds, G, L.d, lane, k0, nz, ny, i, dt, cx, cy);
    }
    __syncthreads();
    if (active) {
      if constexpr (!V2) dq = sdiv_flows<T, KIND>(sv, k, j, i);
      skind_step<T, KIND, NSF, R, NA, V2, MODE>(sv, lds, out, g, L, lp, k, j, i, li, lane, r, dtm, cx, cy, ul, dq, fxl,
                                            fxr, fyn, fzt);
    }
    if (more) stage_store<T, R, NA, V2>(fetch, lds, wv, lane, i + 3, i + 2);
#pragma unroll
    for (int n = 0; n < NSF; n++) fxl[n] = fxr[n];
    __syncthreads();
  }
}

}  // namespace tvd
