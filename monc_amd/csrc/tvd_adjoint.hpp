// Hand-written adjoint of one forward-Euler TVD step.
//
// The forward step is X' = X + dtm * T(X) for every advected field (u, v, w,
// th, q), T the tendency of tvdadvection.F90:528-960 with forward stepping
// (ultflx dt = dtm, advected field = the prognostic).  Given the adjoint of the
// new state, lambda' (interior), the adjoint of the old state is
//   lambda = lambda' + dtm * J(X)^T lambda'
// assembled face by face: each face owner recomputes its face's operands from
// the state (flow signs and the upwind picks of the transverse terms are
// re-derived from the same inputs, so they are the ones the forward pass
// took), takes the limiter arm the forward step recorded (MODE_REC) or
// re-derives it from the same arithmetic, receives the face's adjoint from
// the two cells it bounds, and scatters the reverse-mode partials to the field
// and flow adjoints: in registers, then in per-block LDS rings, then to the
// global adjoints with one device-scope atomic per value and ring column.  A
// fold then returns the adjoint that landed on halo points to their periodic
// sources (the adjoint of the periodic fill).
//
// Derivative conventions at kinks follow the tangent-linear code obtained by
// instantiating the forward templates with Dual: max/min differentiate the
// chosen operand, d|x| = sign-bit(x) ? -dx : dx.
#pragma once
#include "tvd_kernels.hpp"

namespace tvd {

template <typename T>
__device__ __forceinline__ T sgn_abs(T x) { return signbit(x) ? T(-1.0) : T(1.0); }
// derivatives of pos(x) = max(0,x) and negabs(x) = |min(0,x)| as coded
template <typename T>
__device__ __forceinline__ T dpos(T x) { return T(0.0) > x ? T(0.0) : T(1.0); }
template <typename T>
__device__ __forceinline__ T dnegabs(T x) { return T(0.0) < x ? T(0.0) : sgn_abs(x); }

// Adjoint of one target field accumulated in LDS: a ring of 4 x columns
// (slot = i mod 4) x `rows` rows (from j0 - 1) x ADJ_KB levels (from k0 - 1).
// Every contribution of the face owners at (k, j, i) lands
// within levels k-1..k+2, rows j-1..j+2 and columns i-1..i+2 (the stencils of
// the three faces and the dual-grid interpolation of the flows).
constexpr int ADJ_KB = 67;
template <typename T>
struct LdsAcc {
  T* s;
  int k0, j0, rows;
  __device__ __forceinline__ T* ptr(int k, int j, int i) const {
    return &s[((i & 3) * rows + (j - j0 + 1)) * ADJ_KB + (k - k0 + 1)];
  }
};

// One face owner's scatter gathered in registers before it reaches the LDS
// rings.  LDS floating-point atomics issue at a small fraction of the rate of
// plain LDS traffic (measured on MI355X: ~190 CU cycles per 64-lane ds_add_f32
// against ~12 for a read and a write), and an owner adds ~110 times.  Slot
// (t, dk, dj, di) holds the owner's total for target t at (k+dk, j+dj, i+di),
// dk, dj, di in -1..2.  The offsets are compile-time constants at the call
// sites of add() (an offset an upwind pick chooses at run time becomes adds
// to each candidate slot, the value selected per lane), so the slots no call
// reaches stay constant zero and vanish.  After each face the level offsets
// are folded onto each lane's own level with wave shifts (lane L of a wave
// holds level k0+L), leaving one LDS read-add-write per (t, dj, di) and the
// band's edge levels (flush_slots).  No LDS atomic remains.
template <typename T, int NT>
struct Slots {
  T v[NT][4][4][4];  // [target][dk+1][dj+1][di+1]
};

// Target t's slot for (kk, jj, ii), a fixed offset from the owner (k, j, i).
template <typename T, int NT>
__device__ __forceinline__ void slot_add(Slots<T, NT>* s, const LdsAcc<T>* ring, int t, int k, int j, int i, int kk,
                                         int jj, int ii, T v) {
  const int dk = kk - k, dj = jj - j, di = ii - i;
  (void)ring;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (dk == a - 1 && dj == b - 1 && di == c - 1) s->v[t][a][b][c] += v;
}

// A flow's accumulator: target TGT.
template <typename T, int TGT, int NT>
struct SlotAcc {
  Slots<T, NT>* s;
  const LdsAcc<T>* ring;
  int k, j, i;
  __device__ __forceinline__ void add(int kk, int jj, int ii, T v) const { slot_add(s, ring, TGT, k, j, i, kk, jj, ii, v); }
};

// The advected fields' accumulators: field n is target n (n a constant once
// the loops over the fields are unrolled).
template <typename T, int NT>
struct SlotAccF {
  Slots<T, NT>* s;
  const LdsAcc<T>* ring;
  int k, j, i;
  __device__ __forceinline__ void add(int n, int kk, int jj, int ii, T v) const {
    slot_add(s, ring, n, k, j, i, kk, jj, ii, v);
  }
};

// A slot that some add() reaches (not the constant zero of one no call site
// reaches); decided at compile time.
template <typename T>
__device__ __forceinline__ bool live(T x) { return !(__builtin_constant_p(x) && x == T(0)); }

// Adds the slots of every lane of the wave (lane L = level k0+L of row j,
// column i) to the rings, without atomics.  Called by every wave of the block
// together (it holds barriers), after each face.  The slots of row offset dj
// go in phase dj+1: in one phase the wave of owner row r writes ring row
// r+dj only, so no two waves write one row, and within a wave each lane
// writes its own level (lane 0 and 63 alone the band's edges); a barrier
// closes each phase (the last one also orders the phases against the next
// face's flush and the ring's column flush).  Within a
// phase every ring value is read before any is written (the locations are
// distinct), so the reads go out together instead of one LDS round trip per
// read-add-write.
template <typename T, int NT>
__device__ __forceinline__ void flush_slots(const Slots<T, NT>& s, const LdsAcc<T> (&ring)[NT], int k, int j, int i,
                                            int lane) {
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int jj = j + b - 1;
    // level k0+L collects dk = 0 of lane L, +1 of lane L-1, +2 of lane L-2 and
    // -1 of lane L+1; `up` at lane L is what lanes L and L-1 send to level
    // k0+L+1
    T tot[NT][4], up[NT][4];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const T vm = s.v[t][0][b][c], v0 = s.v[t][1][b][c], v1 = s.v[t][2][b][c], v2 = s.v[t][3][b][c];
        up[t][c] = v1;
        if (live(v2)) up[t][c] += from_below(v2);
        tot[t][c] = v0;
        if (live(v1) || live(v2)) tot[t][c] += from_below(up[t][c]);
        if (live(vm)) tot[t][c] += from_above(vm);
      }
    T* p[NT][4];
    T old[NT][4];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (live(tot[t][c])) {
          p[t][c] = ring[t].ptr(k, jj, i + c - 1);
          old[t][c] = *p[t][c];
        }
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (live(tot[t][c])) *p[t][c] = old[t][c] + tot[t][c];
    // the band's edges: level k0-1 from lane 0, k0+64 and k0+65 from lane 63
    if (lane == 0) {
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (live(s.v[t][0][b][c])) {
            p[t][c] = ring[t].ptr(k - 1, jj, i + c - 1);
            old[t][c] = *p[t][c];
          }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (live(s.v[t][0][b][c])) *p[t][c] = old[t][c] + s.v[t][0][b][c];
    }
    if (lane == 63) {
      T* q[NT][4];
      T old2[NT][4];
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (live(up[t][c])) {
            p[t][c] = ring[t].ptr(k + 1, jj, i + c - 1);
            old[t][c] = *p[t][c];
          }
          if (live(s.v[t][3][b][c])) {
            q[t][c] = ring[t].ptr(k + 2, jj, i + c - 1);
            old2[t][c] = *q[t][c];
          }
        }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (live(up[t][c])) *p[t][c] = old[t][c] + up[t][c];
          if (live(s.v[t][3][b][c])) *q[t][c] = old2[t][c] + s.v[t][3][b][c];
        }
    }
    __syncthreads();
  }
}

// Adjoint of Flow<T,DK,DJ,DI>: the dual-grid interpolation scatters to both points.
template <typename T, int DK, int DJ, int DI, class ACC>
struct AccFlow {
  ACC a;
  int nz;
  __device__ __forceinline__ void add(int k, int j, int i, T v) const {
    auto to = [&](int kk, int jj, int ii, T x) { a.add(kk, jj, ii, x); };
    if constexpr (DK == 0 && DJ == 0 && DI == 0) {
      to(k, j, i, v);
    } else if constexpr (DK == 1) {
      if (k < nz) {
        to(k, j, i, T(0.5) * v);
        to(k + 1, j, i, T(0.5) * v);
      } else {
        to(k, j, i, v);
      }
    } else {
      to(k, j, i, T(0.5) * v);
      to(k, j + DJ, i + DI, T(0.5) * v);
    }
  }
};

// Adjoints of the limiter inputs.
template <typename T>
struct LimAdj {
  T fu, fc, fd, g1, g2, hrcfl, hcfl, cflcoef;
};

// Adjoint of cflmod = |vel*ch*dt| through hcfl = cflmod/2 and
// cflcoef = r6*(1 - cflmod^2)*geom: returns d/dvel.
template <typename T>
__device__ __forceinline__ T courant_adj(T vel, T ch, T dt, T geom, T lhcfl, T lcflcoef) {
  const T r6 = T(1.0 / 6.0);
  const T x = vel * ch * dt;
  const T cflmod = fabs(x);
  const T lcfl = T(0.5) * lhcfl - r6 * T(2.0) * cflmod * geom * lcflcoef;
  return lcfl * sgn_abs(x) * ch * dt;
}

// Reverse of a face whose limiter arm is `arm` (LIM_*): with the arm fixed
// the face is linear in the advected field (fc, fd, the QUICKEST value ffs,
// or fu + (fc - fu)*hrcfl), so the adjoints of its operands follow from the
// arm alone.  Returns the face value (from the same arm).  The arm's value
// equals the forward's ftemp*fdels + fu up to the rounding of rf*fdels = 1
// (rf = 1/(fdels + 1e-30)), so the derivative is that of the forward arm up
// to a relative 1e-30/|fdels|.  ZFORM: `rdd`, `rdu` give the z form's curvature
// (fd-fc)*rdd - (fc-fu)*rdu instead of fd - 2fc + fu.
template <typename T, bool ZFORM>
__device__ __forceinline__ T arm_adjoint(int arm, T lam, T fu, T fc, T fd, T g1, T g2, T hrcfl, T hcfl, T cflcoef,
                                         T rdd, T rdu, LimAdj<T>& r) {
  const T curv = ZFORM ? (fd - fc) * rdd - (fc - fu) * rdu : fd - T(2.0) * fc + fu;
  const T ffs = T(0.5) * (fd + fc) - hcfl * (fd - fc) - cflcoef * curv - g1 - g2;
  const T fr = fu + (fc - fu) * hrcfl;
  const bool ac = arm == LIM_C, ad = arm == LIM_D, aq = arm == LIM_Q, ar = arm == LIM_R;
  const T lq = aq ? lam : T(0.0), lr = ar ? lam : T(0.0);
  // ffs = 0.5(fd+fc) - hcfl(fd-fc) - cflcoef*curv - g1 - g2
  const T lcurv = -(lq * cflcoef);
  r.fd = (ad ? lam : T(0.0)) + lq * (T(0.5) - hcfl) + (ZFORM ? lcurv * rdd : lcurv);
  r.fc = (ac ? lam : T(0.0)) + lq * (T(0.5) + hcfl) - (ZFORM ? lcurv * (rdd + rdu) : T(2.0) * lcurv) + lr * hrcfl;
  r.fu = (ZFORM ? lcurv * rdu : lcurv) + lr * (T(1.0) - hrcfl);
  r.hcfl = -(lq * (fd - fc));
  r.cflcoef = -(lq * curv);
  r.g1 = -lq;
  r.g2 = -lq;
  r.hrcfl = lr * (fc - fu);
  return ac ? fc : (ad ? fd : (aq ? ffs : fr));
}

// The adds whose offset an upwind pick chooses at run time go to fixed-offset
// slots with the value selected per lane (0 to the slots not chosen).  This
// one: +x at level ksl and -x at level ksu of (jj, ii), the z-gradient pair of
// the x and y faces, where (ksl, ksu) = (k, k-1) when pz, else (k, k+1) below
// level nz and (k-1, k) at k = nz (knz).
template <class AF, typename T>
__device__ __forceinline__ void add_zpair(const AF& a, int n, bool pz, bool knz, int k, int jj, int ii, T x) {
  a.add(n, k, jj, ii, pz || !knz ? x : -x);
  a.add(n, k - 1, jj, ii, pz ? -x : (knz ? x : T(0.0)));
  a.add(n, k + 1, jj, ii, pz || knz ? T(0.0) : -x);
}

// Right x face of column (j,i) at level k for the NS advected fields: values
// (face[n], for the divergence adjoint of the flow coefficient) and reverse.
// lam[n] is the adjoint of field n's face value.  The velocity-only
// quantities are recomputed once and their adjoints summed over the fields
// before they reach the flows.
template <typename T, int NS, class AF, class FU, class FV, class FW, class AU, class AV, class AW>
__device__ __forceinline__ void face_x_adj(int k, int j, int i, const Adv<T, NS>& F, const AF& aF, const FU& U,
                                           const FV& V, const FW& W, const AU& aU, const AV& aV, const AW& aW,
                                           const VMet& m, int kdof, T dt, T cx, T cy, int nz, const T (&lam)[NS],
                                           T (&face)[NS], bool rec, const int (&arm)[NS]) {
  const int kp1 = k + 1 < nz ? k + 1 : nz;
  const int kp0 = kp1 - 1;
  const T uk = U(k, j, i);
  const T sm = half_sign(uk);
  const T ap = T(0.5) + sm, an = T(0.5) - sm;
  const bool upos = uk > T(0.0);
  const T w0 = W(k, j, i), w0m = W(k - 1, j, i), w1 = W(k, j, i + 1), w1m = W(k - 1, j, i + 1);
  const T v0 = V(k, j, i), v0m = V(k, j - 1, i), v1 = V(k, j, i + 1), v1m = V(k, j - 1, i + 1);
  const T um = U(k, j, i - 1), up1 = U(k, j, i + 1);
  const T wa = upos ? w0 : w1, wb = upos ? w0m : w1m;
  const T va = upos ? v0 : v1, vb = upos ? v0m : v1m;
  const T rz0 = vm<T>(m.rdz, k);
  T s = rz0 * (pos(wa) + negabs(wb));
  s = s + cy * (pos(va) + negabs(vb));
  s = s + cx * (upos ? uk + negabs(um) : pos(up1) - uk);
  s = s * dt;
  const T sz = T(0.125) * (w0 + w0m + w1 + w1m);
  const bool pz = sign_pos(sz);
  const int ksu = pz ? k - 1 : kp1;
  const int ksl = pz ? k : kp0;
  const T rzk = vm<T>(m.rdz, k + kdof);
  const T cz = fabs(sz) * rzk * dt;
  const T sy = T(0.125) * (v1 + v0 + v1m + v0m);
  const bool syp = sign_pos(sy);
  const int js = syp ? j - 1 : j + 1;
  const T cyt = fabs(sy) * cy * dt;
  const bool knz = k >= nz;  // (ksl, ksu) = (k-1, k) for W < 0 at the top level
  const HLim<T> lim(uk, cx, dt, s);
  // adjoints of cz, cyt, hrcfl, hcfl and cflcoef, summed over the fields
  T lcz = T(0.0), lcyt = T(0.0), lhr = T(0.0), lhc = T(0.0), lcc = T(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fa = F(n, k, j, i), fb = F(n, k, j, i + 1);
    const T fm = F(n, k, j, i - 1), fp = F(n, k, j, i + 2);
    const T fu = fm * ap + fp * an, fc = fa * ap + fb * an, fd = fb * ap + fa * an;
    const T l0 = F(n, ksl, j, i), u0 = F(n, ksu, j, i), l1 = F(n, ksl, j, i + 1), u1 = F(n, ksu, j, i + 1);
    const T tmp = (l0 - u0) * ap + (l1 - u1) * an;
    const T g1 = cz * tmp;
    const T s0 = F.lane(n, k, js, i), s1 = F.lane(n, k, js, i + 1);
    const T t = s0 * ap + s1 * an;
    const T g2 = cyt * (fc - t);
    LimAdj<T> r;
    face[n] = arm_adjoint<T, false>(rec ? arm[n] : lim.cls_of(fu, fc, fd, g1, g2), lam[n], fu, fc, fd, g1, g2,
                                    lim.hrcfl, lim.hcfl, lim.cflcoef, T(0), T(0), r);
    // transverse terms
    lcyt += r.g2 * (fc - t);
    const T lfc = r.fc + r.g2 * cyt;
    const T lt = -(r.g2 * cyt);
    lcz += r.g1 * tmp;
    const T ltmp = r.g1 * cz;
    lhr += r.hrcfl;
    lhc += r.hcfl;
    lcc += r.cflcoef;
    // advected field
    aF.add(n, k, j, i, lfc * ap + r.fd * an);
    aF.add(n, k, j, i + 1, lfc * an + r.fd * ap);
    aF.add(n, k, j, i - 1, r.fu * ap);
    aF.add(n, k, j, i + 2, r.fu * an);
    add_zpair(aF, n, pz, knz, k, j, i, ltmp * ap);
    add_zpair(aF, n, pz, knz, k, j, i + 1, ltmp * an);
    const T lta = lt * ap, ltb = lt * an;  // at row js = j-1 (syp) or j+1
    aF.add(n, k, j - 1, i, syp ? lta : T(0.0));
    aF.add(n, k, j + 1, i, syp ? T(0.0) : lta);
    aF.add(n, k, j - 1, i + 1, syp ? ltb : T(0.0));
    aF.add(n, k, j + 1, i + 1, syp ? T(0.0) : ltb);
  }
  // flows (hrcfl multiplies its adjoint first: hrcfl^2 may overflow where
  // the outflow sum is zero, and no face there has a non-zero adjoint)
  const T lS = -(lim.hrcfl * (lim.hrcfl * lhr));
  T luk = courant_adj(uk, cx, dt, T(1.0), lhc, lcc);
  const T lsz = T(0.125) * (lcz * sgn_abs(sz) * rzk * dt);
  const T lsy = T(0.125) * (lcyt * sgn_abs(sy) * cy * dt);
  const T lS3 = lS * dt;
  const T lA = lS3 * rz0, lB = lS3 * cy, lC = lS3 * cx;
  const T lwa = lA * dpos(wa), lwb = lA * dnegabs(wb);
  const T lva = lB * dpos(va), lvb = lB * dnegabs(vb);
  T lum = T(0.0), lup1 = T(0.0);
  if (upos) {
    luk = luk + lC;
    lum = lC * dnegabs(um);
  } else {
    luk = luk - lC;
    lup1 = lC * dpos(up1);
  }
  aW.add(k, j, i, lsz + (upos ? lwa : T(0.0)));
  aW.add(k - 1, j, i, lsz + (upos ? lwb : T(0.0)));
  aW.add(k, j, i + 1, lsz + (upos ? T(0.0) : lwa));
  aW.add(k - 1, j, i + 1, lsz + (upos ? T(0.0) : lwb));
  aV.add(k, j, i, lsy + (upos ? lva : T(0.0)));
  aV.add(k, j - 1, i, lsy + (upos ? lvb : T(0.0)));
  aV.add(k, j, i + 1, lsy + (upos ? T(0.0) : lva));
  aV.add(k, j - 1, i + 1, lsy + (upos ? T(0.0) : lvb));
  aU.add(k, j, i, luk);
  aU.add(k, j, i - 1, lum);
  aU.add(k, j, i + 1, lup1);
}

// North y face of row j at column i, level k: values and reverse.
template <typename T, int NS, class AF, class FU, class FV, class FW, class AU, class AV, class AW>
__device__ __forceinline__ void face_y_adj(int k, int j, int i, const Adv<T, NS>& F, const AF& aF, const FU& U,
                                           const FV& V, const FW& W, const AU& aU, const AV& aV, const AW& aW,
                                           const VMet& m, int kdof, T dt, T cx, T cy, int nz, const T (&lam)[NS],
                                           T (&face)[NS], bool rec, const int (&arm)[NS]) {
  const int kp1 = k + 1 < nz ? k + 1 : nz;
  const int kp0 = kp1 - 1;
  const T vk = V(k, j, i);
  const T sm = half_sign(vk);
  const T ap = T(0.5) + sm, an = T(0.5) - sm;
  const bool vpos = sign_pos(vk);
  const int jj = vpos ? j : j + 1;
  const T w0 = W(k, j, i), w0m = W(k - 1, j, i), w1 = W(k, j + 1, i), w1m = W(k - 1, j + 1, i);
  const T vn = V(k, j + 1, i), vs = V(k, j - 1, i);
  const T u0 = U(k, j, i), u0m = U(k, j, i - 1), u1 = U(k, j + 1, i), u1m = U(k, j + 1, i - 1);
  const T wa = vpos ? w0 : w1, wb = vpos ? w0m : w1m;
  const T va = vpos ? vk : vn, vb = vpos ? vs : vk;
  const T ua = vpos ? u0 : u1, ub = vpos ? u0m : u1m;
  const T rz0 = vm<T>(m.rdz, k);
  T s = rz0 * (pos(wa) + negabs(wb));
  s = s + cy * (pos(va) + negabs(vb));
  s = s + cx * (pos(ua) + negabs(ub));
  s = s * dt;
  const T sz = T(0.125) * (w0 + w0m + w1 + w1m);
  const bool pz = sign_pos(sz);
  const int ksu = pz ? k - 1 : kp1;
  const int ksl = pz ? k : kp0;
  const T rzk = vm<T>(m.rdz, k + kdof);
  const T cz = fabs(sz) * rzk * dt;
  const T sx = T(0.125) * (u0m + u0 + u1m + u1);
  const T posuf = T(0.5) + half_sign(sx);
  const T cxt = fabs(sx) * cx * dt;
  const bool knz = k >= nz;  // (ksl, ksu) = (k-1, k) for W < 0 at the top level
  const HLim<T> lim(vk, cy, dt, s);
  T lcz = T(0.0), lcxt = T(0.0), lhr = T(0.0), lhc = T(0.0), lcc = T(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const T fa = F(n, k, j, i), fb = F(n, k, j + 1, i);
    const T fm = F(n, k, j - 1, i), fp = F(n, k, j + 2, i);
    const T fu = fm * ap + fp * an, fc = fa * ap + fb * an, fd = fb * ap + fa * an;
    const T fl = F.lane(n, ksl, jj, i), fs = F.lane(n, ksu, jj, i);
    const T g1 = cz * (fl - fs);
    const T fe = F.lane(n, k, jj, i + 1), fw = F.lane(n, k, jj, i - 1);
    const T tx = fe * (T(1.0) - posuf) + fw * posuf;
    const T g2 = cxt * (fc - tx);
    LimAdj<T> r;
    face[n] = arm_adjoint<T, false>(rec ? arm[n] : lim.cls_of(fu, fc, fd, g1, g2), lam[n], fu, fc, fd, g1, g2,
                                    lim.hrcfl, lim.hcfl, lim.cflcoef, T(0), T(0), r);
    lcxt += r.g2 * (fc - tx);
    const T lfc = r.fc + r.g2 * cxt;
    const T ltx = -(r.g2 * cxt);
    lcz += r.g1 * (fl - fs);
    const T ldiff = r.g1 * cz;
    lhr += r.hrcfl;
    lhc += r.hcfl;
    lcc += r.cflcoef;
    aF.add(n, k, j, i, lfc * ap + r.fd * an);
    aF.add(n, k, j + 1, i, lfc * an + r.fd * ap);
    aF.add(n, k, j - 1, i, r.fu * ap);
    aF.add(n, k, j + 2, i, r.fu * an);
    // row jj = j (vpos) or j+1
    add_zpair(aF, n, pz, knz, k, j, i, vpos ? ldiff : T(0.0));
    add_zpair(aF, n, pz, knz, k, j + 1, i, vpos ? T(0.0) : ldiff);
    const T lte = ltx * (T(1.0) - posuf), ltw = ltx * posuf;
    aF.add(n, k, j, i + 1, vpos ? lte : T(0.0));
    aF.add(n, k, j + 1, i + 1, vpos ? T(0.0) : lte);
    aF.add(n, k, j, i - 1, vpos ? ltw : T(0.0));
    aF.add(n, k, j + 1, i - 1, vpos ? T(0.0) : ltw);
  }
  const T lS = -(lim.hrcfl * (lim.hrcfl * lhr));
  T lvk = courant_adj(vk, cy, dt, T(1.0), lhc, lcc);
  const T lsz = T(0.125) * (lcz * sgn_abs(sz) * rzk * dt);
  const T lsx = T(0.125) * (lcxt * sgn_abs(sx) * cx * dt);
  const T lS3 = lS * dt;
  const T lA = lS3 * rz0, lB = lS3 * cy, lC = lS3 * cx;
  const T lwa = lA * dpos(wa), lwb = lA * dnegabs(wb);
  const T lva = lB * dpos(va), lvb = lB * dnegabs(vb);
  const T lua = lC * dpos(ua), lub = lC * dnegabs(ub);
  aW.add(k, j, i, lsz + (vpos ? lwa : T(0.0)));
  aW.add(k - 1, j, i, lsz + (vpos ? lwb : T(0.0)));
  aW.add(k, j + 1, i, lsz + (vpos ? T(0.0) : lwa));
  aW.add(k - 1, j + 1, i, lsz + (vpos ? T(0.0) : lwb));
  // va = vk or vn, vb = vs or vk
  if (vpos) {
    lvk = lvk + lva;
    aV.add(k, j - 1, i, lvb);
  } else {
    aV.add(k, j + 1, i, lva);
    lvk = lvk + lvb;
  }
  aV.add(k, j, i, lvk);
  aU.add(k, j, i, lsx + (vpos ? lua : T(0.0)));
  aU.add(k, j, i - 1, lsx + (vpos ? lub : T(0.0)));
  aU.add(k, j + 1, i, lsx + (vpos ? T(0.0) : lua));
  aU.add(k, j + 1, i - 1, lsx + (vpos ? T(0.0) : lub));
}

// Top face of level k (flux_z(k+1)), 1 <= k <= nz-1: values and reverse.
// The transverse terms (ZGrad) are reversed per field, their weights' adjoints
// summed over the fields.
template <typename T, int NS, class AF, class FU, class FV, class FW, class AU, class AV, class AW>
__device__ __forceinline__ void face_z_adj(int k, int j, int i, const Adv<T, NS>& F, const AF& aF, const FU& U,
                                           const FV& V, const FW& W, const AU& aU, const AV& aV, const AW& aW,
                                           const VMet& m, int kdof, int kmin, T dt, T cx, T cy, int nz,
                                           const T (&lam)[NS], T (&face)[NS], bool rec, const int (&arm)[NS]) {
  if (k == 1 && kmin > 1) {
#pragma unroll
    for (int n = 0; n < NS; n++) face[n] = T(0.0);
    return;
  }
  const T wk = W(k, j, i);
  T s, rdu, rdc;
  int kk;
  // S = ((a0 + cy*B) + cx*C) * dt with a0 the W part; record its inputs
  int kw0 = k, kw1 = k - 1, kv = k;  // W(kw0), W(kw1), V(kv, j/j-1), U(kv, i/i-1)
  int mode;                          // 0 middle, 1 top/bottom up, 2 top down, 3 bottom down
  T rzw;
  const bool bottom = k == 1, top = k == nz - 1;
  T ap = T(0.0), an = T(0.0);
  if (bottom) {
    if (wk >= T(0.0)) {
      mode = 1;
      kv = 1;
      rzw = vm<T>(m.rdz, 1 + kdof);
      s = wk * rzw;
      rdu = vm<T>(m.rdzn, 2);
      rdc = vm<T>(m.rdz, 1 + kdof);
      kk = 1;
    } else {
      mode = 3;
      kv = 2;
      kw0 = 2;
      rzw = vm<T>(m.rdz, 2 + kdof);
      s = rzw * (pos(W(2, j, i)) - wk);
      rdu = vm<T>(m.rdzn, 3);
      rdc = vm<T>(m.rdz, 2 + kdof);
      kk = 2;
    }
  } else if (top) {
    if (wk >= T(0.0)) {
      mode = 1;
      kv = k;
      rzw = vm<T>(m.rdz, k + kdof);
      s = rzw * (wk + negabs(W(k - 1, j, i)));
      rdu = vm<T>(m.rdzn, k);
      rdc = vm<T>(m.rdz, k + kdof);
      kk = k;
    } else {
      mode = 2;
      kv = k + 1;
      const int kt = k + 1 + kdof < nz ? k + 1 + kdof : nz;
      rzw = vm<T>(m.rdz, kt);
      s = -(wk * rzw);
      rdu = vm<T>(m.rdzn, nz);
      rdc = rzw;
      kk = k + 1;
    }
  } else {
    mode = 0;
    const T sm = half_sign(wk);
    ap = T(0.5) + sm, an = T(0.5) - sm;
    kk = sign_pos(wk) ? k : k + 1;
    kv = kk;
    kw0 = kk;
    kw1 = kk - 1;
    rzw = vm<T>(m.rdz, kk);
    s = rzw * (pos(W(kk, j, i)) + negabs(W(kk - 1, j, i)));
    rdu = vm<T>(m.rdzn, k) * ap + vm<T>(m.rdzn, k + 2) * an;
    rdc = vm<T>(m.rdz, k + kdof) * ap + vm<T>(m.rdz, k + 1 + kdof) * an;
  }
  const T vb0 = V(kv, j, i), vb1 = V(kv, j - 1, i), ub0 = U(kv, j, i), ub1 = U(kv, j, i - 1);
  s = s + cy * (pos(vb0) + negabs(vb1));
  s = s + cx * (pos(ub0) + negabs(ub1));
  s = s * dt;
  // transverse weights (ZGrad), kept with their sums for the reverse
  const T v00 = V(k, j - 1, i), v01 = V(k, j, i), v10 = V(k + 1, j - 1, i), v11 = V(k + 1, j, i);
  const T sy = T(0.125) * (v00 + v01 + v10 + v11);
  const bool syp = sign_pos(sy);
  const int js = syp ? j - 1 : j + 1;
  const T cyt = fabs(sy) * cy * dt;
  const bool kup = kk == k;  // the upwind level kk is k (else k+1)
  const T u00 = U(k, j, i - 1), u01 = U(k, j, i), u10 = U(k + 1, j, i - 1), u11 = U(k + 1, j, i);
  const T sx = T(0.125) * (u00 + u01 + u10 + u11);
  const T posuf = T(0.5) + half_sign(sx);
  const T cxt = fabs(sx) * cx * dt;
  const WLim<T> lim(k, wk, dt, s, rdc, rdu, m);
  T ly = T(0.0), lx = T(0.0), lhr = T(0.0), lhc = T(0.0), lcc = T(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    T fu, fc, fd;
    if (bottom) {
      const T f1 = F(n, 1, j, i), f2 = F(n, 2, j, i);
      const bool up = mode == 1;
      fu = up ? f1 : F(n, 3, j, i);
      fc = up ? f1 : f2;
      fd = up ? f2 : f1;
    } else if (top) {
      const T fk = F(n, k, j, i), fk1 = F(n, k + 1, j, i);
      const bool up = mode == 1;
      // W < 0: linear extrapolation to a level equidistant above the top
      fu = up ? F(n, k - 1, j, i) : T(2.0) * fk1 - fk;
      fc = up ? fk : fk1;
      fd = up ? fk1 : fk;
    } else {
      fu = F(n, k - 1, j, i) * ap + F(n, k + 2, j, i) * an;
      fc = F(n, kk, j, i);
      fd = F(n, k + 1, j, i) * ap + F(n, k, j, i) * an;
    }
    const T fsj = F.lane(n, kk, js, i);
    const T fe = F(n, kk, j, i + 1), fw = F(n, kk, j, i - 1);
    const T tx = fe * (T(1.0) - posuf) + fw * posuf;
    const T g1 = cyt * (fc - fsj), g2 = cxt * (fc - tx);
    LimAdj<T> r;
    face[n] = arm_adjoint<T, true>(rec ? arm[n] : lim.cls_of(fu, fc, fd, g1, g2), lam[n], fu, fc, fd, g1, g2,
                                   lim.hrcfl, lim.hcfl, lim.cflcoef, lim.rdd, lim.rdu, r);
    lhr += r.hrcfl;
    lhc += r.hcfl;
    lcc += r.cflcoef;
    // g1 = cyt*(fc - fsj), g2 = cxt*(fc - tx)
    const T lfc = r.fc + r.g1 * cyt + r.g2 * cxt;
    // level kk = k (kup) or k+1, row js = j-1 (syp) or j+1
    const T ly1 = -(r.g1 * cyt);
    aF.add(n, k, j - 1, i, kup && syp ? ly1 : T(0.0));
    aF.add(n, k, j + 1, i, kup && !syp ? ly1 : T(0.0));
    aF.add(n, k + 1, j - 1, i, !kup && syp ? ly1 : T(0.0));
    aF.add(n, k + 1, j + 1, i, !kup && !syp ? ly1 : T(0.0));
    const T lxe = -(r.g2 * cxt) * (T(1.0) - posuf), lxw = -(r.g2 * cxt) * posuf;
    aF.add(n, k, j, i + 1, kup ? lxe : T(0.0));
    aF.add(n, k + 1, j, i + 1, kup ? T(0.0) : lxe);
    aF.add(n, k, j, i - 1, kup ? lxw : T(0.0));
    aF.add(n, k + 1, j, i - 1, kup ? T(0.0) : lxw);
    ly += r.g1 * (fc - fsj);
    lx += r.g2 * (fc - tx);
    // stencil values
    if (mode == 0) {
      aF.add(n, k - 1, j, i, r.fu * ap);
      aF.add(n, k + 2, j, i, r.fu * an);
      aF.add(n, k, j, i, kup ? lfc : T(0.0));
      aF.add(n, k + 1, j, i, kup ? T(0.0) : lfc);
      aF.add(n, k + 1, j, i, r.fd * ap);
      aF.add(n, k, j, i, r.fd * an);
    } else if (bottom) {
      // k = 1 here
      if (mode == 1) {  // fu = fc = F(1), fd = F(2)
        aF.add(n, k, j, i, r.fu + lfc);
        aF.add(n, k + 1, j, i, r.fd);
      } else {  // fu = F(3), fc = F(2), fd = F(1)
        aF.add(n, k + 2, j, i, r.fu);
        aF.add(n, k + 1, j, i, lfc);
        aF.add(n, k, j, i, r.fd);
      }
    } else {
      if (mode == 1) {  // fu = F(k-1), fc = F(k), fd = F(k+1)
        aF.add(n, k - 1, j, i, r.fu);
        aF.add(n, k, j, i, lfc);
        aF.add(n, k + 1, j, i, r.fd);
      } else {  // fu = 2F(k+1) - F(k), fc = F(k+1), fd = F(k)
        aF.add(n, k + 1, j, i, T(2.0) * r.fu + lfc);
        aF.add(n, k, j, i, r.fd - r.fu);
      }
    }
  }
  const T lsy = T(0.125) * (ly * sgn_abs(sy) * cy * dt);
  const T lsx = T(0.125) * (lx * sgn_abs(sx) * cx * dt);
  aV.add(k, j - 1, i, lsy);
  aV.add(k, j, i, lsy);
  aV.add(k + 1, j - 1, i, lsy);
  aV.add(k + 1, j, i, lsy);
  aU.add(k, j, i - 1, lsx);
  aU.add(k, j, i, lsx);
  aU.add(k + 1, j, i - 1, lsx);
  aU.add(k + 1, j, i, lsx);
  // flows: courant number, outflow sum
  const T rdd = lim.rdd;
  const T dd = vm<T>(m.dzn, k + 1);
  T lwk = courant_adj(wk, rdd, dt, dd * dd * rdc, lhc, lcc);
  const T lS = -(lim.hrcfl * (lim.hrcfl * lhr));
  const T lS3 = lS * dt;
  const T lB = lS3 * cy, lC = lS3 * cx;
  // level kv = k or k+1
  const bool kvk = kv == k;
  const T lv0 = lB * dpos(vb0), lv1 = lB * dnegabs(vb1), lu0 = lC * dpos(ub0), lu1 = lC * dnegabs(ub1);
  aV.add(k, j, i, kvk ? lv0 : T(0.0));
  aV.add(k + 1, j, i, kvk ? T(0.0) : lv0);
  aV.add(k, j - 1, i, kvk ? lv1 : T(0.0));
  aV.add(k + 1, j - 1, i, kvk ? T(0.0) : lv1);
  aU.add(k, j, i, kvk ? lu0 : T(0.0));
  aU.add(k + 1, j, i, kvk ? T(0.0) : lu0);
  aU.add(k, j, i - 1, kvk ? lu1 : T(0.0));
  aU.add(k + 1, j, i - 1, kvk ? T(0.0) : lu1);
  const T lW = lS3;  // adjoint of the W part a0 of S
  if (mode == 0) {
    const T wa = W(kw0, j, i), wb = W(kw1, j, i);
    // kw0 = kk, kw1 = kk - 1
    const T l0 = lW * rzw * dpos(wa), l1 = lW * rzw * dnegabs(wb);
    aW.add(k, j, i, kup ? l0 : T(0.0));
    aW.add(k + 1, j, i, kup ? T(0.0) : l0);
    aW.add(k - 1, j, i, kup ? l1 : T(0.0));
    aW.add(k, j, i, kup ? T(0.0) : l1);
  } else if (bottom && mode == 1) {  // a0 = wk * rzw
    lwk = lwk + lW * rzw;
  } else if (bottom) {  // a0 = rzw * (pos(W(2)) - wk)
    aW.add(k + 1, j, i, lW * rzw * dpos(W(2, j, i)));  // k = 1
    lwk = lwk - lW * rzw;
  } else if (mode == 1) {  // a0 = rzw * (wk + negabs(W(k-1)))
    lwk = lwk + lW * rzw;
    aW.add(k - 1, j, i, lW * rzw * dnegabs(W(k - 1, j, i)));
  } else {  // a0 = -(wk * rzw)
    lwk = lwk - lW * rzw;
  }
  aW.add(k, j, i, lwk);
}

// One kernel kind's adjoint: NS advected fields (a momentum field alone, or
// up to two scalars that share their flows).
template <typename T, int NS>
struct AdjArgs {
  const T* F[NS];    // advected fields (forward stepping: the prognostics)
  const T* lin[NS];  // adjoint of the new state of each field (interior)
  T* lf[NS];         // adjoint accumulators of the fields
  const unsigned char* rec[NS];  // limiter arms the forward step recorded (MODE_REC), or null
  T* lu;             // adjoint accumulators of the flows u, v, w
  T* lv;
  T* lw;
  T dtm;
};

// The face owner (k, j, i), 1 <= k <= nz, 2 <= j <= ny+2, 2 <= i <= nx+2: the
// right x face, north y face and top z face of that point for the NS fields,
// their adjoints scattered into the accumulators aF (advected fields, target
// n for field n) and aU, aV, aW (flows, through the dual-grid interpolation of
// the field kind).  `after_face()` runs after each face in control flow that
// every lane of the wave reaches (lanes past nz take no face); register
// accumulators are emptied there.
template <typename T, int KIND, int NS, bool REC, class AF, class AU, class AV, class AW, class FL>
__device__ __forceinline__ void adjoint_point(const LaunchArgs<T>& L, const AdjArgs<T, NS>& A, int k, int j, int i,
                                              const AF& aF, const AU& aU, const AV& aV, const AW& aW,
                                              const FL& after_face) {
  const int nz = L.d.nz, ny = L.d.ny, nx = L.d.nx;
  // rows past ny+2 own no face, but their waves still reach every
  // after_face() (it may hold barriers); uniform over the wave
  const bool own = j <= ny + 2;
  const Flows<T, KIND> fl(L);
  Adv<T, NS> F{};
#pragma unroll
  for (int n = 0; n < NS; n++) F.a[n] = Arr<T>(A.F[n], L.d);
  const VMet m = KIND == KIND_W ? L.mw : L.mc;
  const int kdof = KIND == KIND_W ? 1 : 0;
  const int kmin = KIND == KIND_W ? 1 : 2;
  const int kmax = KIND == KIND_W ? nz - 1 : nz;
  const T dt = A.dtm;
  // adjoint lambda'(cell) of the new state, zero where no tendency is formed.
  // Differences of lambda' are taken before scaling by dtm: lambda' of a
  // smooth loss (e.g. 2 th) is large against its differences, and scaling
  // first would round each term before the cancellation.
  auto lraw = [&](int n, int kk, int jj, int ii) -> T {
    const bool cell = jj >= 3 && jj <= ny + 2 && ii >= 3 && ii <= nx + 2 && kk >= 2 && kk <= kmax;
    return cell ? A.lin[n][at(L.d, kk, jj, ii)] : T(0.0);
  };
  const bool row_in = own && j >= 3, col_in = own && i >= 3;
  T lself[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) lself[n] = lraw(n, k, j, i);
  // the limiter arms of the owner's faces, as the forward step recorded them
  // (REC), else re-derived from the state inside the face reverses
  int code[NS] = {};
  if constexpr (REC) {
    if (own && k <= nz) {
#pragma unroll
      for (int n = 0; n < NS; n++) code[n] = A.rec[n][at(L.d, k, j, i)];
    }
  }
  int ax[NS], ay[NS], az[NS];
#pragma unroll
  for (int n = 0; n < NS; n++) {
    ax[n] = code[n] & 3;
    ay[n] = (code[n] >> 2) & 3;
    az[n] = (code[n] >> 4) & 3;
  }
  // x face (k, j, i+1/2): s(i) -= U fx cx, s(i+1) += U fx cx
  if (row_in && k >= 2 && k <= kmax) {
    const T uf = fl.U(k, j, i);
    T d[NS], lam[NS], fx[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
      d[n] = A.dtm * (lraw(n, k, j, i + 1) - lself[n]);
      lam[n] = uf * L.cx * d[n];
    }
    face_x_adj(k, j, i, F, aF, fl.U, fl.V, fl.W, aU, aV, aW, m, kdof, dt, L.cx, L.cy, nz, lam, fx, REC, ax);
    T ld = T(0.0);
#pragma unroll
    for (int n = 0; n < NS; n++) ld += fx[n] * L.cx * d[n];
    aU.add(k, j, i, ld);
  }
  after_face();
  // y face (k, j+1/2, i)
  if (col_in && k >= 2 && k <= kmax) {
    const T vf = fl.V(k, j, i);
    T d[NS], lam[NS], fy[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
      d[n] = A.dtm * (lraw(n, k, j + 1, i) - lself[n]);
      lam[n] = vf * L.cy * d[n];
    }
    face_y_adj(k, j, i, F, aF, fl.U, fl.V, fl.W, aU, aV, aW, m, kdof, dt, L.cx, L.cy, nz, lam, fy, REC, ay);
    T ld = T(0.0);
#pragma unroll
    for (int n = 0; n < NS; n++) ld += fy[n] * L.cy * d[n];
    aV.add(k, j, i, ld);
  }
  after_face();
  // z face at level k (flux_z(k+1)): s(k) -= 4 W Fz t2(k), s(k+1) += 4 W Fz t1(k+1);
  // t1(k+1) l(k+1) - t2(k) l(k) = t1(k+1) (l(k+1) - l(k)) + (t1(k+1) - t2(k)) l(k),
  // the metric difference formed in FP64
  if (row_in && col_in && k <= nz - 1 && !(k == 1 && kmin > 1)) {
    const T dmet = (T)(m.t1[k] - m.t2[k - 1]);
    const T wf = fl.W(k, j, i);
    T e[NS], lam[NS], fz[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
      e[n] = T(4.0) * A.dtm * (vm<T>(m.t1, k + 1) * (lraw(n, k + 1, j, i) - lself[n]) + dmet * lself[n]);
      lam[n] = wf * e[n];
    }
    face_z_adj(k, j, i, F, aF, fl.U, fl.V, fl.W, aU, aV, aW, m, kdof, kmin, dt, L.cx, L.cy, nz, lam, fz, REC, az);
    T ld = T(0.0);
#pragma unroll
    for (int n = 0; n < NS; n++) ld += fz[n] * e[n];
    aW.add(k, j, i, ld);
  }
  after_face();
}

// LDS of the march adjoint: the rings of the NT target fields (the advected
// fields, then the u, v, w flows) for R owner rows.
template <typename T, int R, int NT = 4>
struct AdjLds {
  T ring[NT][4][R + 3][ADJ_KB];
};

// March adjoint tile: 64 levels (band `band`) x R owner rows (from row 2 +
// ty*R) x the x strip of owner columns from 2 + tx*strip, for NS fields of
// kind KIND.  Each column of owners adds into the LDS rings (register slots
// folded by wave shifts, in row phases); a column leaves the ring (x-1 after
// owners x) once no later owner of the tile can reach it, and is added to the
// global adjoints with one atomic per non-zero value.  Device-scope atomics
// are carried out beyond the L2, so the ring cuts the atomics that reach
// memory from ~70 per face owner to ~(strip+3)(R+3)/(strip R) per target.
// Every thread of the block calls this with the same tile.
template <typename T, int KIND, int R, int NS = 1, bool REC = false>
__device__ __forceinline__ void adjoint_tile(const LaunchArgs<T>& L0, const AdjArgs<T, NS>& A0, int band0, int ty0,
                                             int tx0, int strip0, AdjLds<T, R, NS + 3>& lds) {
  // the tile's arguments are uniform over the block (the ensemble reads them
  // from its member table)
  LaunchArgs<T> L = L0;
  L.d = uniform(L0.d);
  L.u = uniform_ptr(L0.u), L.v = uniform_ptr(L0.v), L.w = uniform_ptr(L0.w);
  for (VMet* m : {&L.mc, &L.mw}) {
    m->rdz = uniform_ptr(m->rdz), m->rdzn = uniform_ptr(m->rdzn), m->dzn = uniform_ptr(m->dzn);
    m->t1 = uniform_ptr(m->t1), m->t2 = uniform_ptr(m->t2);
  }
  L.dtm = uniform(L0.dtm), L.cx = uniform(L0.cx), L.cy = uniform(L0.cy);
  AdjArgs<T, NS> A;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    A.F[n] = uniform_ptr(A0.F[n]);
    A.lin[n] = uniform_ptr(A0.lin[n]);
    A.lf[n] = uniform_ptr(A0.lf[n]);
    A.rec[n] = REC ? uniform_ptr(A0.rec[n]) : nullptr;
  }
  A.lu = uniform_ptr(A0.lu), A.lv = uniform_ptr(A0.lv), A.lw = uniform_ptr(A0.lw);
  A.dtm = uniform(A0.dtm);
  const int band = uniform(band0), ty = uniform(ty0), tx = uniform(tx0), strip = uniform(strip0);
  constexpr int DK = Flows<T, KIND>::DK, DJ = Flows<T, KIND>::DJ, DI = Flows<T, KIND>::DI;
  constexpr int NT = NS + 3, ROWS = R + 3, PER = ROWS * ADJ_KB, NTH = 64 * R;
  const int nz = L.d.nz, ny = L.d.ny, nx = L.d.nx;
  const int lane = threadIdx.x, r = (int)__builtin_amdgcn_readfirstlane(threadIdx.y);
  const int tid = lane + 64 * r;
  const int k0 = __builtin_amdgcn_readfirstlane(1 + 64 * band), j0 = __builtin_amdgcn_readfirstlane(2 + ty * R);
  const int i_begin = 2 + tx * strip, i_end = min(i_begin + strip, nx + 3);
  // An owner's face adjoints are proportional to the incoming adjoint at
  // (k,j,i), (k,j,i+1), (k,j+1,i) and (k+1,j,i): a tile that reads only zeros
  // there adds nothing (the adjoint of a localised loss is zero far from it).
  // The check covers the box k..k+1, j..j+1, i_begin..i_end around the
  // tile's owners (a superset): each lane reads its own (k, j) over the
  // columns, the last row in range also row j+1 and lane 63 level k+1, all
  // loads independent of each other.
  {
    const int k = k0 + lane, j = j0 + r;
    int any = 0;
    if (j <= ny + 2 && k <= nz) {
      const bool row_up = r == R - 1 || j == ny + 2, lev_up = lane == 63 && k < nz;
      for (int i = i_begin; i <= i_end; i++)
#pragma unroll
        for (int n = 0; n < NS; n++) {
          any |= A.lin[n][at(L.d, k, j, i)] != T(0);
          if (row_up) any |= A.lin[n][at(L.d, k, j + 1, i)] != T(0);
          if (lev_up) any |= A.lin[n][at(L.d, k + 1, j, i)] != T(0);
        }
    }
    if (!__syncthreads_or(any)) return;
  }
  T* flat = &lds.ring[0][0][0][0];
  for (int e = tid; e < NT * 4 * PER; e += NTH) flat[e] = T(0);
  __syncthreads();
  LdsAcc<T> ring[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) ring[t] = LdsAcc<T>{&lds.ring[t][0][0][0], k0, j0, ROWS};
  T* dst[NT];
#pragma unroll
  for (int n = 0; n < NS; n++) dst[n] = A.lf[n];
  dst[NS] = A.lu;
  dst[NS + 1] = A.lv;
  dst[NS + 2] = A.lw;
  // add ring column c (all targets) to the global adjoints and clear it
  // (target by target, so that each target's pointer is known to be global)
  auto flush = [&](int c) {
    if (c < 1 || c > nx + 4) return;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      T* const g = dst[t];
      for (int e = tid; e < PER; e += NTH) {
        const int rr = e / ADJ_KB, lev = e - rr * ADJ_KB;
        T* p = &lds.ring[t][c & 3][rr][lev];
        const T v = *p;
        const int kk = k0 - 1 + lev, jj = j0 - 1 + rr;
        if (v != T(0) && kk >= 1 && kk <= nz && jj >= 1 && jj <= ny + 4) unsafeAtomicAdd(g + at(L.d, kk, jj, c), v);
        *p = T(0);
      }
    }
  };
  for (int ic = i_begin; ic < i_end; ic++) {
    // the column is uniform over the block; saying so keeps it in a scalar
    // register (a column offset in a vector register makes every buffer
    // load a loop over the distinct offsets of the wave)
    const int i = __builtin_amdgcn_readfirstlane(ic);
    {
      const int k = k0 + lane, j = j0 + r;
      Slots<T, NT> sl{};
      const SlotAccF<T, NT> aF{&sl, ring, k, j, i};
      const AccFlow<T, DK, DJ, DI, SlotAcc<T, NS, NT>> aU{{&sl, ring, k, j, i}, nz};
      const AccFlow<T, DK, DJ, DI, SlotAcc<T, NS + 1, NT>> aV{{&sl, ring, k, j, i}, nz};
      const AccFlow<T, DK, DJ, DI, SlotAcc<T, NS + 2, NT>> aW{{&sl, ring, k, j, i}, nz};
      adjoint_point<T, KIND, NS, REC>(L, A, k, j, i, aF, aU, aV, aW, [&] {
        flush_slots(sl, ring, k, j, i, lane);
        sl = Slots<T, NT>{};
      });
    }
    __syncthreads();
    flush(i - 1);
    __syncthreads();
  }
  // the last three columns the tile reached (distinct ring slots)
#pragma nounroll
  for (int c = i_end - 1; c <= i_end + 1; c++) flush(c);
  __syncthreads();
}

// March adjoint: grid (64-level bands, owner-row tiles, x strips). The FP32
// scalar-pair tile reading recorded arms is held to 128 registers per lane (4
// waves per SIMD): left free it takes 131 (3 waves) and runs 5.4 against
// 4.9 ms at 512x512x256. The hint slows the other FP32 tiles, which fit in 128
// without it, so they go without.
constexpr int ADJ_ROWS = 4;
template <typename T, int NS, bool REC>
constexpr int adj_waves_per_eu() { return sizeof(T) == 4 && NS == 2 && REC ? 4 : 1; }
template <typename T, int KIND, int NS = 1, bool REC = false>
__global__ void __launch_bounds__(64 * ADJ_ROWS) __attribute__((amdgpu_waves_per_eu(adj_waves_per_eu<T, NS, REC>())))
adjoint_march_kernel(LaunchArgs<T> L, AdjArgs<T, NS> A, int strip) {
  __shared__ AdjLds<T, ADJ_ROWS, NS + 3> lds;
  adjoint_tile<T, KIND, ADJ_ROWS, NS, REC>(L, A, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, strip, lds);
}

// Adjoint of the periodic halo fill, in two passes so every target is
// written by one thread: rows first (over all columns, corners included),
// then columns over the interior rows.  Halo entries are zeroed.
template <typename T>
__device__ __forceinline__ void fold_rows_block(T* f, const Dims& d, int bx, int by) {
  const int k = 1 + bx * 64 + (int)threadIdx.x;
  const int h = by * 4 + (int)threadIdx.y;  // 0..4*(nx+4)-1
  if (k > d.nz || h >= 4 * (d.nx + 4)) return;
  const int r = h / (d.nx + 4), i = 1 + h % (d.nx + 4);
  const int j = r < 2 ? 1 + r : d.ny + 1 + r;
  const int64_t src = at(d, k, j, i), dst = at(d, k, wrap(j, d.ny), i);
  f[dst] = f[dst] + f[src];
  f[src] = T(0.0);
}

template <typename T>
__device__ __forceinline__ void fold_cols_block(T* f, const Dims& d, int bx, int by) {
  const int k = 1 + bx * 64 + (int)threadIdx.x;
  const int h = by * 4 + (int)threadIdx.y;  // 0..4*ny-1
  if (k > d.nz || h >= 4 * d.ny) return;
  const int c = h / d.ny, j = 3 + h % d.ny;
  const int i = c < 2 ? 1 + c : d.nx + 1 + c;
  const int64_t src = at(d, k, j, i), dst = at(d, k, j, wrap(i, d.nx));
  f[dst] = f[dst] + f[src];
  f[src] = T(0.0);
}

// Up to MAX_FIELDS fields of one domain, one launch for all (blockIdx.z or
// blockIdx.y selects the field).
template <typename T>
struct FieldPtrs {
  T* f[MAX_FIELDS];
  const T* src[MAX_FIELDS];
};

template <typename T>
__global__ void __launch_bounds__(256) fold_rows_kernel(FieldPtrs<T> P, Dims d) {
  fold_rows_block<T>(P.f[blockIdx.z], d, (int)blockIdx.x, (int)blockIdx.y);
}

template <typename T>
__global__ void __launch_bounds__(256) fold_cols_kernel(FieldPtrs<T> P, Dims d) {
  fold_cols_block<T>(P.f[blockIdx.z], d, (int)blockIdx.x, (int)blockIdx.y);
}

// f[n] = src[n] for every field n (blockIdx.y), `words` 32-bit words each, in
// 16-byte pieces (both buffers 16-byte aligned) and a word tail.
template <typename T>
__global__ void __launch_bounds__(256) copy_fields_kernel(FieldPtrs<T> P, int64_t words) {
  const uint4* s4 = (const uint4*)P.src[blockIdx.y];
  uint4* d4 = (uint4*)P.f[blockIdx.y];
  const int64_t n4 = words / 4;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n4; t += stride) d4[t] = s4[t];
  const unsigned* s1 = (const unsigned*)P.src[blockIdx.y];
  unsigned* d1 = (unsigned*)P.f[blockIdx.y];
  const int64_t t = 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < words) d1[t] = s1[t];
}


}  // namespace tvd
