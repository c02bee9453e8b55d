// March kernel: the default x-marching kernel of the TVD step on gfx950.
//
// A block owns R cell rows (j0 .. j0+R-1) over all levels: NCH waves of 64
// levels per row (lanes along z), R*NCH waves, and marches along an x strip.
// Per column i every cell (k, j, i) computes each face it owns exactly once:
//   * the right x face, carried in registers to the next column as its left face;
//   * the north y face, handed to the row above through LDS (the first row of
//     the block also computes its south face);
//   * the top z face, handed to the level above by a wave shuffle and across
//     64-level sub-columns through LDS.
// One barrier per column orders the LDS exchange (double-buffered by column
// parity).
//
// The face arithmetic is that of tvd_faces.hpp (ultimateflux.F90:81-614) with
// three changes of form that keep every value identical:
//   * the reference's upwind weights ap = 1/2 + sign(1/2, vel), an = 1 - ap are
//     exactly 0 or 1, so ap*x + an*y is the selection ap ? x : y;
//   * the three vertical branches of ultflx (bottom k = 1 for w, middle, top
//     k = nz-1; :236-523) are one computation whose few differing operands are
//     chosen per lane: no wave runs two branches;
//   * the per-level metric factors of each branch are tabulated once per block
//     in LDS in the working precision.
// Loads use buffer instructions with the wave-uniform row and column in the
// scalar offset and the per-lane level in the vector offset (levels clamped to
// [1, nz], so every load stays inside its own column).
#pragma once
#include "tvd_kernels.hpp"

namespace tvd {

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ void stb(__amdgpu_buffer_rsrc_t r, T v, unsigned vo, unsigned so) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vo, so, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 0);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)0xffffffff, BUFFER_RSRC_WORD3);
}

// Byte strides shared by every field of a launch.
struct Geo {
  unsigned sjb, sib;
  __device__ __forceinline__ unsigned so(int j, int i) const {
    return (unsigned)(j - 1) * sjb + (unsigned)(i - 1) * sib;
  }
};

// A flow as a field kind sees it (stencil.F90:147-169): raw for scalars, the
// dual-grid interpolation 0.5*(f + f shifted one step along u's x, v's y or
// w's z) for momentum.  A level is given by its byte offset o and that of the
// level above it, o1, clamped to nz: at k = nz the interpolation of w reads
// 0.5*(f + f) = f, the reference's raw copy (:160-163).
template <typename T, int KIND>
struct MFlow {
  static constexpr int DK = KIND == KIND_W ? 1 : 0;
  static constexpr int DJ = KIND == KIND_V ? 1 : 0;
  static constexpr int DI = KIND == KIND_U ? 1 : 0;
  __amdgpu_buffer_rsrc_t r;
  Geo g;
  __device__ __forceinline__ T operator()(unsigned o, unsigned o1, int j, int i) const {
    const unsigned s = g.so(j, i);
    if constexpr (DK) return T(0.5) * (ldb<T>(r, o, s) + ldb<T>(r, o1, s));
    else if constexpr (DJ) return T(0.5) * (ldb<T>(r, o, s) + ldb<T>(r, o, s + g.sjb));
    else if constexpr (DI) return T(0.5) * (ldb<T>(r, o, s) + ldb<T>(r, o, s + g.sib));
    else return ldb<T>(r, o, s);
  }
};

// Per-lane byte offsets of the levels a lane reads, each clamped to [1, nz].
struct Levels {
  unsigned k, km1, kp1, kp0, kp2;  // k, k-1, k+1, (k+1 clamped) - 1, k+2
  unsigned edge;                   // the wave's edge levels: k0-1 (lane 0), k0+64 (lane 1), k0+65 (lane 2)
};

// The advected field around the lane's level k, clamped to [1, nz] as the
// loads clamp: m1 = F(k-1), p1 = F(k+1), p2 = F(k+2), p0 = F(min(k+1, nz)-1),
// taken from the lane's own F(k) by wave shifts instead of loads (lane L of a
// wave holds level k0+L); the levels beyond the wave (k0-1, k0+64, k0+65)
// come from one load per column.  Lanes past nz hold F(nz) and compute faces
// nothing uses.
template <typename T>
struct VEdge {
  T lo, hi1, hi2;
};
template <typename T>
struct VNb {
  T m1, p1, p2, p0;
};

template <typename T>
__device__ __forceinline__ VEdge<T> vedge(__amdgpu_buffer_rsrc_t r, const Levels& lv, unsigned so) {
  const T v = ldb<T>(r, lv.edge, so);
  return {lane_value(v, 0), lane_value(v, 1), lane_value(v, 2)};
}

template <typename T>
__device__ __forceinline__ VNb<T> vnb(T x, const VEdge<T>& e, int lane, int k, int nz) {
  VNb<T> v;
  const T a = from_above(x), b = from_below(x);
  v.p1 = k >= nz ? x : (lane == 63 ? e.hi1 : a);
  v.m1 = k == 1 ? x : (lane == 0 ? e.lo : b);
  const T a2 = from_above(v.p1);
  v.p2 = lane == 63 ? e.hi2 : a2;
  v.p0 = k >= nz ? v.m1 : x;
  return v;
}

// Metric factors per level, tabulated in LDS for the field kind (for w the
// reference passes (rdzn, rdz, dz) in the roles of (rdz, rdzn, dzn), kdof = 1;
// tvdadvection.F90:881-894).  Indices clamped to [1, nz].
enum {
  M_RDZ,     // rdz(k): x/y outflow sums
  M_RDZKD,   // rdz(k+kdof): x/y vertical gradient terms
  M_ZS_UP,   // z outflow-sum factor, W >= 0: rdz(k); rdz(1+kdof) at k = 1, rdz(k+kdof) at k = nz-1
  M_ZS_DN,   // W < 0: rdz(k+1); rdz(2+kdof) at k = 1
  M_RDC_UP,  // rdz(k+kdof)
  M_RDC_DN,  // rdz(k+1+kdof)
  M_RDU_UP,  // rdzn(k); rdzn(2) at k = 1
  M_RDU_DN,  // rdzn(k+2) (rdzn(nz) at k = nz-1)
  M_RDD,     // rdzn(k+1)
  M_DD2,     // dzn(k+1)^2
  M_T1,      // tzc1 / tzd1 (k)
  M_T2,      // tzc2 / tzd2 (k); 0 at k = nz
  M_COUNT
};

template <typename T>
__device__ __forceinline__ T met_value(int m, int k, int nz, int kdof, const VMet& v) {
  auto c = [nz](int x) { return x < 1 ? 1 : (x > nz ? nz : x); };
  switch (m) {
    case M_RDZ: return vm<T>(v.rdz, c(k));
    case M_RDZKD: return vm<T>(v.rdz, c(k + kdof));
    case M_ZS_UP: return vm<T>(v.rdz, c(k == 1 ? 1 + kdof : (k == nz - 1 ? k + kdof : k)));
    case M_ZS_DN: return vm<T>(v.rdz, c(k == 1 ? 2 + kdof : k + 1));
    case M_RDC_UP: return vm<T>(v.rdz, c(k + kdof));
    case M_RDC_DN: return vm<T>(v.rdz, c(k + 1 + kdof));
    case M_RDU_UP: return vm<T>(v.rdzn, c(k == 1 ? 2 : k));
    case M_RDU_DN: return vm<T>(v.rdzn, c(k + 2));
    case M_RDD: return vm<T>(v.rdzn, c(k + 1));
    case M_DD2: {
      const T dd = vm<T>(v.dzn, c(k + 1));
      return dd * dd;
    }
    case M_T1: return vm<T>(v.t1, c(k));
    default: return k < nz ? vm<T>(v.t2, c(k)) : T(0);
  }
}

// Flows at the level of the lane (and k-1, k+1) around column (j, i) that the
// three faces of cell (k, j, i) and its divergence read.
template <typename T>
struct FlowsAt {
  T u_m, u_0, u_p, u1_m, u1_0;      // U(k) at (j,i-1) (j,i) (j,i+1) (j+1,i-1) (j+1,i)
  T v_0, v_s, v_0p, v_sp, v_n;      // V(k) at (j,i) (j-1,i) (j,i+1) (j-1,i+1) (j+1,i)
  T w_0, w_p, w_n;                  // W(k) at (j,i) (j,i+1) (j+1,i)
  T wm_0, wm_p, wm_n;               // W(k-1) at the same
  T uk_m, uk_0, vk_s, vk_0, wk_0;   // U(k+1) at (j,i-1) (j,i); V(k+1) at (j-1,i) (j,i); W(k+1) at (j,i)
};

// Right x face of column (j, i) at the lane's level (ultimateflux.F90:102-165);
// xw[n] holds F(k, j, i-1 .. i+2) of field n (the x window of the march); with
// VD the vertical neighbours at columns i and i+1 come from c0 and c1, else
// they are loaded.
template <typename T, int NS, bool VD>
__device__ __forceinline__ void mx_face(const __amdgpu_buffer_rsrc_t (&F)[NS], const Geo& g, const Levels& lv, int j,
                                        int i, const T (&xw)[NS][4], const VNb<T> (&c0)[NS], const VNb<T> (&c1)[NS],
                                        T u_m, T u_0, T u_p, T v_0, T v_s, T v_0p, T v_sp, T w_0, T w_p, T wm_0,
                                        T wm_p, T rdz_k, T rdz_kd, T dt, T cx, T cy, T out[NS], int cls[NS]) {
  const bool upos = u_0 > T(0.0);
  T s = rdz_k * (pos(upos ? w_0 : w_p) + negabs(upos ? wm_0 : wm_p));
  s = s + cy * (pos(upos ? v_0 : v_0p) + negabs(upos ? v_s : v_sp));
  s = s + cx * (upos ? u_0 + negabs(u_m) : pos(u_p) - u_0);
  s = s * dt;
  const T sz = T(0.125) * (w_0 + wm_0 + w_p + wm_p);
  const bool pz = sign_pos(sz);  // ksl = k, ksu = k-1; else ksl = (k+1 clamped) - 1, ksu = k+1
  const T cz = fabs(sz) * rdz_kd * dt;
  const T sy = T(0.125) * (v_0p + v_0 + v_sp + v_s);
  const unsigned ojs = lv.k + (sign_pos(sy) ? 0u : 2u * g.sjb);  // row j-1 or j+1, from row j-1
  const T cyt = fabs(sy) * cy * dt;
  const HLim<T> lim(u_0, cx, dt, s);
  const bool ap = sign_pos(u_0);
  const unsigned sj0 = g.so(j - 1, i), sj1 = g.so(j - 1, i + 1);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    // every load is unconditional and only values are selected: a load whose
    // scalar offset were chosen per lane would have to be serialised
    const T fm = xw[n][0], fa = xw[n][1], fb = xw[n][2], f2 = xw[n][3];
    const T fu = ap ? fm : f2;
    const T fc = ap ? fa : fb;
    const T fd = ap ? fb : fa;
    T d0, d1;
    if constexpr (VD) {
      d0 = pz ? fa - c0[n].m1 : c0[n].p0 - c0[n].p1;
      d1 = pz ? fb - c1[n].m1 : c1[n].p0 - c1[n].p1;
    } else {
      const unsigned osl = pz ? lv.k : lv.kp0, osu = pz ? lv.km1 : lv.kp1;
      const unsigned s_0 = g.so(j, i), s_1 = g.so(j, i + 1);
      d0 = ldb<T>(F[n], osl, s_0) - ldb<T>(F[n], osu, s_0);
      d1 = ldb<T>(F[n], osl, s_1) - ldb<T>(F[n], osu, s_1);
    }
    const T g1 = cz * (ap ? d0 : d1);
    const T t0 = ldb<T>(F[n], ojs, sj0), t1 = ldb<T>(F[n], ojs, sj1);
    const T t = ap ? t0 : t1;
    const T g2 = cyt * (fc - t);
    out[n] = lim.face_cls(fu, fc, fd, g1, g2, cls[n]);
  }
}

// North y face of row j at column i (ultimateflux.F90:189-233); vk, vn, vs =
// V at rows j, j+1, j-1; w0/w0m, w1/w1m = W(k), W(k-1) at rows j, j+1; u0/u0m,
// u1/u1m = U at (row j, i), (row j, i-1), (row j+1, i), (row j+1, i-1).
template <typename T, int NS>
__device__ __forceinline__ void my_face(const __amdgpu_buffer_rsrc_t (&F)[NS], const Geo& g, const Levels& lv, int j,
                                        int i, const T (&fc0)[NS], T vk, T vn, T vs, T w0, T w0m, T w1, T w1m, T u0,
                                        T u0m, T u1, T u1m, T rdz_k, T rdz_kd, T dt, T cx, T cy, T out[NS],
                                        int cls[NS]) {
  const bool vpos = sign_pos(vk);
  T s = rdz_k * (pos(vpos ? w0 : w1) + negabs(vpos ? w0m : w1m));
  s = s + cy * (pos(vpos ? vk : vn) + negabs(vpos ? vs : vk));
  s = s + cx * (pos(vpos ? u0 : u1) + negabs(vpos ? u0m : u1m));
  s = s * dt;
  const T sz = T(0.125) * (w0 + w0m + w1 + w1m);
  const bool pz = sign_pos(sz);
  const unsigned orow = vpos ? 0u : g.sjb;  // row jj = j + jofset
  const unsigned osl = (pz ? lv.k : lv.kp0) + orow;
  const unsigned osu = (pz ? lv.km1 : lv.kp1) + orow;
  const T cz = fabs(sz) * rdz_kd * dt;
  const T sx = T(0.125) * (u0m + u0 + u1m + u1);
  const bool xup = sign_pos(sx);
  const unsigned ot = lv.k + orow + (xup ? 0u : 2u * g.sib);  // (jj, i-1) or (jj, i+1), from column i-1
  const T cxt = fabs(sx) * cx * dt;
  const HLim<T> lim(vk, cy, dt, s);
  const unsigned s_s = g.so(j - 1, i), s_0 = g.so(j, i), s_1 = g.so(j + 1, i), s_2 = g.so(j + 2, i);
  const unsigned s_im = g.so(j, i - 1);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fa = fc0[n], fb = ldb<T>(F[n], lv.k, s_1);
    const T fs = ldb<T>(F[n], lv.k, s_s), f2 = ldb<T>(F[n], lv.k, s_2);
    const T fu = vpos ? fs : f2;
    const T fc = vpos ? fa : fb;
    const T fd = vpos ? fb : fa;
    const T g1 = cz * (ldb<T>(F[n], osl, s_0) - ldb<T>(F[n], osu, s_0));
    const T g2 = cxt * (fc - ldb<T>(F[n], ot, s_im));
    out[n] = lim.face_cls(fu, fc, fd, g1, g2, cls[n]);
  }
}

// Top z face of the lane's level k (the reference's flux_z(k+1)),
// ultimateflux.F90:236-523, all three branches in one: `edge` marks k = 1 and
// k = nz-1, whose upwind test is W >= 0 (the middle one takes the sign, so they
// differ for W = -0), `top` marks k = nz-1 (W < 0 extrapolates linearly above
// the top and has no inflow from above), `bot` marks k = 1 (no level below).
// With VD the levels k-1, k+1, k+2 come from c0, else they are loaded.
template <typename T, int NS, bool VD>
__device__ __forceinline__ void mz_face(const __amdgpu_buffer_rsrc_t (&F)[NS], const Geo& g, const Levels& lv, int j,
                                        int i, const T (&fc0)[NS], const VNb<T> (&c0)[NS], const FlowsAt<T>& q,
                                        bool edge, bool top, bool bot,
                                        T zs_up, T zs_dn,
                                        T rdc_up, T rdc_dn, T rdu_up, T rdu_dn, T rdd, T dd2, T dt, T cx, T cy,
                                        T out[NS], int cls[NS]) {
  const T wk = q.w_0;
  const bool up = edge ? wk >= T(0.0) : sign_pos(wk);
  // outflow sum of the upwind cell kk = k (up) or k+1
  const T wa = up ? pos(wk) : (top ? T(0.0) : pos(q.wk_0));
  const T wb = up ? (bot ? T(0.0) : negabs(q.wm_0)) : negabs(wk);
  T s = (up ? zs_up : zs_dn) * (wa + wb);
  s = s + cy * (pos(up ? q.v_0 : q.vk_0) + negabs(up ? q.v_s : q.vk_s));
  s = s + cx * (pos(up ? q.u_0 : q.uk_0) + negabs(up ? q.u_m : q.uk_m));
  s = s * dt;
  // transverse gradient weights at level k (:697-761)
  const T sy = T(0.125) * (q.v_s + q.v_0 + q.vk_s + q.vk_0);
  const T cyt = fabs(sy) * cy * dt;
  const T sx = T(0.125) * (q.u_m + q.u_0 + q.uk_m + q.uk_0);
  const bool xup = sign_pos(sx);
  const T cxt = fabs(sx) * cx * dt;
  const unsigned okk = up ? lv.k : lv.kp1;
  const unsigned ojs = okk + (sign_pos(sy) ? 0u : 2u * g.sjb);  // (kk, j-1 or j+1), from row j-1
  const unsigned oxs = okk + (xup ? 0u : 2u * g.sib);           // (kk, i-1 or i+1), from column i-1
  // limiter, velocity and metric part (calculate_flux_in_w, :525-554)
  const T rdc = up ? rdc_up : rdc_dn;
  const T rdu = up ? rdu_up : rdu_dn;
  const T r6 = T(1.0 / 6.0);
  const T cflmod = fabs(wk * rdd * dt);
  const T hrcfl = recip(s + T(1.e-30));
  const T hcfl = T(0.5) * cflmod;
  const T cflcoef = r6 * (T(1.0) - cflmod * cflmod) * (dd2 * rdc);
  const unsigned s_jm = g.so(j - 1, i), s_im = g.so(j, i - 1);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fk = fc0[n];
    T fk1, fkm, fk2;
    if constexpr (VD) {
      fk1 = c0[n].p1;
      fkm = c0[n].m1;
      fk2 = c0[n].p2;
    } else {
      const unsigned s_0 = g.so(j, i);
      fk1 = ldb<T>(F[n], lv.kp1, s_0);
      fkm = ldb<T>(F[n], lv.km1, s_0);
      fk2 = ldb<T>(F[n], lv.kp2, s_0);
    }
    const T fdn = top ? T(2.0) * fk1 - fk : fk2;
    const T fu = up ? fkm : fdn;
    const T fc = up ? fk : fk1;
    const T fd = up ? fk1 : fk;
    const T g1 = cyt * (fc - ldb<T>(F[n], ojs, s_jm));
    const T g2 = cxt * (fc - ldb<T>(F[n], oxs, s_im));
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

// LDS of one block: the metric table of the field kind (NKMAX levels), the
// north y faces of every wave (double-buffered by column parity) and the top z
// face of the last level of every wave's sub-column.
template <typename T, int NSL, int WAVES, int NKMAX>
struct MarchLds {
  T met[M_COUNT][NKMAX];
  T yn[2][WAVES][NSL][64];
  T zt[2][WAVES][NSL];
};

// One tile: cell rows j0 = 3 + ty*R .. j0+R-1 (R = WAVES / nch), all levels
// (nch sub-columns of 64), x columns 3 + tx*strip onwards.  Wave wv takes
// sub-column wv % nch of row wv / nch; waves past R*nch only take the
// barriers.  Every wave of the block calls this with the same tile.  WIN:
// carry the x windows from column to column (below) or reload them; VD: take
// the advected field's vertical neighbours by wave shifts (VNb) or load them.
template <typename T, int KIND, int NS, int MODE, int WAVES, int NKMAX, int NSL, bool WIN, bool VD>
__device__ __forceinline__ void march_tile(const LaunchArgs<T>& L, const FieldSet<T>& S, int nch, int ty, int tx,
                                           int strip, MarchLds<T, NSL, WAVES, NKMAX>& lds) {
  constexpr bool IS_W = KIND == KIND_W;
  constexpr int kdof = IS_W ? 1 : 0;
  const int R = WAVES / nch;
  const int NK = 64 * nch;
  const int nz = L.d.nz, ny = L.d.ny, nx = L.d.nx;
  const VMet m = IS_W ? L.mw : L.mc;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  for (int e = tid; e < M_COUNT * NK; e += 64 * WAVES) {
    const int mm = e / NK, kk = 1 + e % NK;
    lds.met[mm][kk - 1] = met_value<T>(mm, kk, nz, kdof, m);
  }
  const int lane = threadIdx.x;
  const int wv = (int)__builtin_amdgcn_readfirstlane(threadIdx.y);
  const int qc = wv % nch, r = wv / nch;
  const int j = 3 + ty * R + r;
  const bool active = r < R && j <= ny + 2;  // wave-uniform
  const int k = 1 + 64 * qc + lane;
  const int kmax = IS_W ? nz - 1 : nz;
  const bool cell = k >= 2 && k <= kmax;
  const bool has_z = IS_W ? (k >= 1 && k <= nz - 1) : (k >= 2 && k <= nz - 1);
  const bool edge = k == 1 || k == nz - 1;
  const bool ztop = k == nz - 1, zbot = k == 1;
  const bool top = !IS_W && k == nz;
  const int i_begin = 3 + tx * strip;
  const int i_end = min(i_begin + strip, nx + 3);
  const T dt = S.dt[0], cx = L.cx, cy = L.cy;

  const unsigned es = sizeof(T);
  auto off = [&](int kk) { return (unsigned)((kk < 1 ? 1 : (kk > nz ? nz : kk)) - 1) * es; };
  Levels lv;
  lv.k = off(k);
  lv.km1 = off(k - 1);
  lv.kp1 = off(k + 1);
  lv.kp0 = off((k + 1 > nz ? nz : k + 1) - 1);
  lv.kp2 = off(k + 2);
  lv.edge = off(lane == 1 ? k - lane + 64 : (lane == 2 ? k - lane + 65 : k - lane - 1));
  Geo g;
  g.sjb = (unsigned)L.d.sj * es;
  g.sib = (unsigned)L.d.si * es;
  MFlow<T, KIND> U{rsrc(L.u), g}, V{rsrc(L.v), g}, W{rsrc(L.w), g};
  __amdgpu_buffer_rsrc_t F[NS], Fo[NS], Ff[NS], Fz[NS], Fo2[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) {
    F[n] = rsrc(S.F[n]);
    Fo[n] = rsrc(S.out[n]);
    Ff[n] = rsrc(S.f[n]);
    if constexpr (MODE == MODE_LEAP) {
      Fz[n] = rsrc(S.z[n]);
      Fo2[n] = rsrc(S.out2[n]);
    }
  }
  __syncthreads();
  const int kl = (k <= NK ? k : NK) - 1;
  const T rdz_k = lds.met[M_RDZ][kl], rdz_kd = lds.met[M_RDZKD][kl];

  // The march carries, from column i to column i+1, the x window of the
  // advected fields at (k, j): xw[n] = F(i-1 .. i+2), one new load per column
  // (the centre value also serves the y and z faces and the step's base), and
  // the flows the two columns share (U at i-1, i; U(j+1) at i-1; V at i and
  // (j-1, i); W(k), W(k-1) at i; U(k+1) at i-1): each value is loaded once.
  // The kernels are bound by the vector-memory address path, so loads saved
  // are time saved.  With WIN = false every column reloads them.
  T fxl[NS];
  T xw[NS][4];
  VEdge<T> e0[NS], e1[NS];  // the wave's edge levels of F at (j, i) and (j, i+1)
  T c_um = T(0), c_u0 = T(0), c_u1m = T(0), c_v0 = T(0), c_vs = T(0), c_w0 = T(0), c_wm0 = T(0), c_ukm = T(0);
#pragma unroll
  for (int n = 0; n < NS; n++) fxl[n] = xw[n][0] = xw[n][1] = xw[n][2] = xw[n][3] = T(0);
  if (active) {
    const int i = i_begin - 1;
#pragma unroll
    for (int n = 0; n < NS; n++)
#pragma unroll
      for (int d = 0; d < 4; d++) xw[n][d] = ldb<T>(F[n], lv.k, g.so(j, i - 1 + d));
    const T um = U(lv.k, lv.kp1, j, i - 1);
    c_um = U(lv.k, lv.kp1, j, i);
    c_u0 = U(lv.k, lv.kp1, j, i + 1);
    c_u1m = U(lv.k, lv.kp1, j + 1, i);
    const T v0 = V(lv.k, lv.kp1, j, i), vs = V(lv.k, lv.kp1, j - 1, i);
    c_v0 = V(lv.k, lv.kp1, j, i + 1);
    c_vs = V(lv.k, lv.kp1, j - 1, i + 1);
    const T w0 = W(lv.k, lv.kp1, j, i), wm0 = W(lv.km1, lv.k, j, i);
    c_w0 = W(lv.k, lv.kp1, j, i + 1);
    c_wm0 = W(lv.km1, lv.k, j, i + 1);
    c_ukm = U(lv.kp1, lv.kp2, j, i);
    VNb<T> c0[NS], c1[NS];
    if constexpr (VD) {
#pragma unroll
      for (int n = 0; n < NS; n++) {
        e0[n] = vedge<T>(F[n], lv, g.so(j, i));
        e1[n] = vedge<T>(F[n], lv, g.so(j, i + 1));
        c0[n] = vnb(xw[n][1], e0[n], lane, k, nz);
        c1[n] = vnb(xw[n][2], e1[n], lane, k, nz);
        e0[n] = e1[n];
      }
    }
    int cpro[NS];
    mx_face<T, NS, VD>(F, g, lv, j, i, xw, c0, c1, um, c_um, c_u0, v0, vs, c_v0, c_vs, w0, c_w0, wm0, c_wm0, rdz_k,
                   rdz_kd, dt, cx, cy, fxl, cpro);
    if constexpr (MODE == MODE_REC) {  // the right x face of column 2 (owner (k, j, 2))
      if (i == 2 && k <= nz) {
#pragma unroll
        for (int n = 0; n < NS; n++) S.rec[n][at(L.d, k, j, 2)] = (unsigned char)cpro[n];
      }
    }
#pragma unroll
    for (int n = 0; n < NS; n++) {  // window of column i_begin: F(i_begin-1 .. i_begin+1), F(i_begin+2) below
      xw[n][0] = xw[n][1];
      xw[n][1] = xw[n][2];
      xw[n][2] = xw[n][3];
    }
  }
  int par = 0;
  for (int i = i_begin; i < i_end; i++) {
    T fxr[NS], fyn[NS], fys[NS], fzt[NS], fzb[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) fxr[n] = fyn[n] = fys[n] = fzt[n] = T(0);
    FlowsAt<T> q;
    T fc0[NS];
    if (active) {
      if constexpr (!WIN) {  // reload what the window would carry
#pragma unroll
        for (int n = 0; n < NS; n++)
#pragma unroll
          for (int d = 0; d < 3; d++) xw[n][d] = ldb<T>(F[n], lv.k, g.so(j, i - 1 + d));
        c_um = U(lv.k, lv.kp1, j, i - 1);
        c_u0 = U(lv.k, lv.kp1, j, i);
        c_u1m = U(lv.k, lv.kp1, j + 1, i - 1);
        c_v0 = V(lv.k, lv.kp1, j, i);
        c_vs = V(lv.k, lv.kp1, j - 1, i);
        c_w0 = W(lv.k, lv.kp1, j, i);
        c_wm0 = W(lv.km1, lv.k, j, i);
        c_ukm = U(lv.kp1, lv.kp2, j, i - 1);
      }
      VNb<T> c0[NS], c1[NS];
#pragma unroll
      for (int n = 0; n < NS; n++) {
        xw[n][3] = ldb<T>(F[n], lv.k, g.so(j, i + 2));
        fc0[n] = xw[n][1];
        if constexpr (VD) {
          if constexpr (!WIN) e0[n] = vedge<T>(F[n], lv, g.so(j, i));
          e1[n] = vedge<T>(F[n], lv, g.so(j, i + 1));
          c0[n] = vnb(xw[n][1], e0[n], lane, k, nz);
          c1[n] = vnb(xw[n][2], e1[n], lane, k, nz);
        }
      }
      q.u_m = c_um;
      q.u_0 = c_u0;
      q.u_p = U(lv.k, lv.kp1, j, i + 1);
      q.u1_m = c_u1m;
      q.u1_0 = U(lv.k, lv.kp1, j + 1, i);
      q.v_0 = c_v0;
      q.v_s = c_vs;
      q.v_0p = V(lv.k, lv.kp1, j, i + 1);
      q.v_sp = V(lv.k, lv.kp1, j - 1, i + 1);
      q.v_n = V(lv.k, lv.kp1, j + 1, i);
      q.w_0 = c_w0;
      q.w_p = W(lv.k, lv.kp1, j, i + 1);
      q.w_n = W(lv.k, lv.kp1, j + 1, i);
      q.wm_0 = c_wm0;
      q.wm_p = W(lv.km1, lv.k, j, i + 1);
      q.wm_n = W(lv.km1, lv.k, j + 1, i);
      q.uk_m = c_ukm;
      q.uk_0 = U(lv.kp1, lv.kp2, j, i);
      q.vk_s = V(lv.kp1, lv.kp2, j - 1, i);
      q.vk_0 = V(lv.kp1, lv.kp2, j, i);
      q.wk_0 = W(lv.kp1, lv.kp2, j, i);
      int cyn[NS], cys[NS] = {}, cxr[NS], czf[NS];
      my_face<T, NS>(F, g, lv, j, i, fc0, q.v_0, q.v_n, q.v_s, q.w_0, q.wm_0, q.w_n, q.wm_n, q.u_0, q.u_m, q.u1_0,
                     q.u1_m, rdz_k, rdz_kd, dt, cx, cy, fyn, cyn);
      if (r == 0) {  // south face of the block's first row
        T fs0[NS];
#pragma unroll
        for (int n = 0; n < NS; n++) fs0[n] = ldb<T>(F[n], lv.k, g.so(j - 1, i));
        my_face<T, NS>(F, g, lv, j - 1, i, fs0, q.v_s, q.v_0, V(lv.k, lv.kp1, j - 2, i), W(lv.k, lv.kp1, j - 1, i),
                       W(lv.km1, lv.k, j - 1, i), q.w_0, q.wm_0, U(lv.k, lv.kp1, j - 1, i),
                       U(lv.k, lv.kp1, j - 1, i - 1), q.u_0, q.u_m, rdz_k, rdz_kd, dt, cx, cy, fys, cys);
      }
      mx_face<T, NS, VD>(F, g, lv, j, i, xw, c0, c1, q.u_m, q.u_0, q.u_p, q.v_0, q.v_s, q.v_0p, q.v_sp, q.w_0, q.w_p,
                     q.wm_0, q.wm_p, rdz_k, rdz_kd, dt, cx, cy, fxr, cxr);
      T zf[NS];
      mz_face<T, NS, VD>(F, g, lv, j, i, fc0, c0, q, edge, ztop, zbot, lds.met[M_ZS_UP][kl], lds.met[M_ZS_DN][kl],
                     lds.met[M_RDC_UP][kl], lds.met[M_RDC_DN][kl], lds.met[M_RDU_UP][kl], lds.met[M_RDU_DN][kl],
                     lds.met[M_RDD][kl], lds.met[M_DD2][kl], dt, cx, cy, zf, czf);
#pragma unroll
      for (int n = 0; n < NS; n++) fzt[n] = has_z ? zf[n] : T(0);
      if constexpr (MODE == MODE_REC) {  // the arms of the faces cell (k, j, i) owns
        if (k <= nz) {
#pragma unroll
          for (int n = 0; n < NS; n++) {
            S.rec[n][at(L.d, k, j, i)] = (unsigned char)(cxr[n] | (cyn[n] << 2) | (czf[n] << 4));
            // the south face of row 3 is the north face of owner (k, 2, i)
            if (j == 3) S.rec[n][at(L.d, k, 2, i)] = (unsigned char)(cys[n] << 2);
          }
        }
      }
    }
#pragma unroll
    for (int n = 0; n < NS; n++) {
      lds.yn[par][wv][n][lane] = fyn[n];
      fzb[n] = __shfl_up(fzt[n], 1);
    }
    if (nch > 1 && lane == 63) {
#pragma unroll
      for (int n = 0; n < NS; n++) lds.zt[par][wv][n] = fzt[n];
    }
    __syncthreads();
    if (r > 0) {  // the south face is the north face of the row below: wave wv - nch
#pragma unroll
      for (int n = 0; n < NS; n++) fys[n] = lds.yn[par][wv - nch][n][lane];
    }
    if (nch > 1 && lane == 0 && qc > 0) {
#pragma unroll
      for (int n = 0; n < NS; n++) fzb[n] = lds.zt[par][wv - 1][n];
    }
    par ^= 1;
    if (active) {
      // scalar offsets of the cell and of its periodic images, computed outside
      // the per-lane branches
      const unsigned so = g.so(j, i);
      const int jh = (j <= 4) ? j + ny : (j >= ny + 1 ? j - ny : 0);
      const int ih = (i <= 4) ? i + nx : (i >= nx + 1 ? i - nx : 0);
      const unsigned soj = g.so(jh ? jh : j, i), soi = g.so(j, ih ? ih : i), soji = g.so(jh ? jh : j, ih ? ih : i);
      const T t1 = lds.met[M_T1][kl], t2 = lds.met[M_T2][kl];
      const T wt = k < nz ? q.w_0 : T(0);
#pragma unroll
      for (int n = 0; n < NS; n++) {
        T sv;
        const T s0 = MODE == MODE_ACCUM ? ldb<T>(Fo[n], lv.k, so) : T(0);
        T fcur = T(0);
        if constexpr (MODE == MODE_STEP || MODE == MODE_REC) fcur = S.f[n] == S.F[n] ? fc0[n] : ldb<T>(Ff[n], lv.k, so);
        else if constexpr (MODE == MODE_LEAP) fcur = ldb<T>(Ff[n], lv.k, so);
        if (top) {
          sv = (q.v_s * fys[n] - q.v_0 * fyn[n]) * cy;
          sv = sv + T(4.0) * t1 * q.wm_0 * fzb[n];
          sv = sv + (q.u_m * fxl[n] - q.u_0 * fxr[n]) * cx;
        } else {
          sv = s0 + (q.v_s * fys[n] - q.v_0 * fyn[n]) * cy;
          sv = sv + T(4.0) * (q.wm_0 * fzb[n] * t1 - wt * fzt[n] * t2);
          sv = sv + (q.u_m * fxl[n] - q.u_0 * fxr[n]) * cx;
        }
        if constexpr (MODE == MODE_ACCUM) {
          if (cell) stb<T>(Fo[n], sv, lv.k, so);
          if (S.zf[n] && k <= nz) stb<T>(rsrc(S.zf[n]), k == 1 ? T(0) : fzb[n], lv.k, so);
        } else if constexpr (MODE == MODE_LEAP) {
          // cleared sources: 0 where the reference leaves s untouched
          const T zcur = L.forward ? T(0) : ldb<T>(Fz[n], lv.k, so);
          T fn, zn;
          leap_update<T, KIND>(L, S.gal[n], fcur, zcur, cell ? sv : T(0), fn, zn);
          if (k <= nz) {
            stb<T>(Fo[n], fn, lv.k, so);
            stb<T>(Fo2[n], zn, lv.k, so);
            if (jh) {
              stb<T>(Fo[n], fn, lv.k, soj);
              stb<T>(Fo2[n], zn, lv.k, soj);
            }
            if (ih) {
              stb<T>(Fo[n], fn, lv.k, soi);
              stb<T>(Fo2[n], zn, lv.k, soi);
            }
            if (jh && ih) {
              stb<T>(Fo[n], fn, lv.k, soji);
              stb<T>(Fo2[n], zn, lv.k, soji);
            }
          }
        } else {
          // levels the reference leaves untouched (k = 1; k = nz for w) keep f
          const T val = cell ? fcur + L.dtm * sv : fcur;
          if (k <= nz) {
            stb<T>(Fo[n], val, lv.k, so);
            if (jh) stb<T>(Fo[n], val, lv.k, soj);
            if (ih) stb<T>(Fo[n], val, lv.k, soi);
            if (jh && ih) stb<T>(Fo[n], val, lv.k, soji);
          }
        }
      }
    }
#pragma unroll
    for (int n = 0; n < NS; n++) {
      fxl[n] = fxr[n];
      xw[n][0] = xw[n][1];
      xw[n][1] = xw[n][2];
      xw[n][2] = xw[n][3];
      if constexpr (VD) e0[n] = e1[n];
    }
    if (active) {
      c_um = q.u_0;
      c_u0 = q.u_p;
      c_u1m = q.u1_0;
      c_v0 = q.v_0p;
      c_vs = q.v_sp;
      c_w0 = q.w_p;
      c_wm0 = q.wm_p;
      c_ukm = q.uk_0;
    }
  }
  __syncthreads();  // the LDS is reused by the block's next tile
}

// The FP64 scalar pair in tiles of at most 2 rows of 256 levels runs at 2
// waves per SIMD, with registers to carry its x windows (3.03 against 3.18 ms
// per launch at 512x512x256 in 2 x 64 tiles).
template <typename T, int NS, int NCH, int R>
constexpr bool pair_wide() {
  return sizeof(T) == 8 && NS == 2 && NCH == 4 && R <= 2;
}

// Single-domain launch: a 1-D grid of tiles_y * tiles_x blocks of NCH*R waves.
// Consecutive blocks of one XCD (blocks are dealt to the 8 XCDs round-robin)
// take consecutive tiles, so that rows shared by neighbouring tiles meet in
// one L2.
template <typename T, int KIND, int NS, int MODE, int NCH, int R, int WPE>
__global__ void __launch_bounds__(64 * NCH * R) __attribute__((amdgpu_waves_per_eu(WPE)))
march_kernel(LaunchArgs<T> L, FieldSet<T> S, int strip, int tiles_y) {
  __shared__ MarchLds<T, NS, NCH * R, 64 * NCH> lds;
  const int nb = (int)gridDim.x, bid = (int)blockIdx.x;
  const int full = nb - nb % 8;
  const int t = bid < full ? (bid % 8) * (full / 8) + bid / 8 : bid;
  // The scalar pair reloads its x windows: two windows leave FP64 no
  // registers at 3 waves per SIMD and cost FP32 a wave per SIMD, which
  // measured slower than the loads they save (pair_wide shapes excepted).
  // With max and min as single instructions the FP32 pair fits its windows
  // in 76-77 registers at 6 waves per SIMD, but measured no faster (1.424
  // against 1.429 ms per launch at 512x512x256), so it still reloads.  Vertical neighbours by wave
  // shifts pay for the w kernel (3% in FP64, 2% in FP32); the FP64 u, v and
  // scalar kernels, at 3 waves per SIMD and two shifts per value, run faster
  // loading them, and in FP32 the gain depends on the tile shape.
  constexpr bool VD = KIND == KIND_W && NS == 1;
  constexpr bool WIN = NS != 2 || pair_wide<T, NS, NCH, R>();
  march_tile<T, KIND, NS, MODE, NCH * R, 64 * NCH, NS, WIN, VD>(L, S, NCH, t % tiles_y, t / tiles_y, strip, lds);
}

}  // namespace tvd
