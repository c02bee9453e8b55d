/*
 * CPU restatement of MONC's TVD advection path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle the HIP kernels are checked against.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load it,
 * and only as the checker: the product library never links or calls it.
 *
 * It restates, in plain C and in the reference's own column-by-column order,
 *   components/tvdadvection/src/ultimateflux.F90   (whole file)
 *   components/tvdadvection/src/tvdadvection.F90   :528-599, :603-960, :962-1055
 *   model_core/src/grid/stencil.F90                :87-203 (dual interpolation)
 *   model_core/src/communication/halocommunication.F90 :1058-1168 (periodic fill)
 *   components/gridmanager/src/gridmanager.F90     :2048-2067, :2449-2456 (metrics)
 *   model_core/src/components/timestepper.F90      :65-122 (column loop, flags)
 *   components/stepfields/src/stepfields.F90       :597-627 (forward step)
 *
 * Conventions.  Every array is a Fortran column-major block data(k,j,i) =
 * data(z,y,x), 1-based, z fastest, with a depth-2 halo in x and y and none in z:
 * k = 1..nz, j = 1..ny+4, i = 1..nx+4, interior j,i = 3..n+2.  The helpers below
 * take those 1-based indices so each line can be read against the Fortran.
 * Vertical vectors (rdz, rdzn, dz, dzn, tzc*, tzd*) are length nz, 1-based.
 *
 * Parity status of each part is listed in the module header of
 * tests/test_oracle.py together with the check that pins it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* PLANT selects one deliberate restatement mistake (0: none).  Only the
 * pinning test of tests/test_oracle.py builds variants with PLANT > 0, to show
 * that each mistake fails at least one of its checks; the oracle proper is
 * always built with PLANT = 0. */
#ifndef PLANT
#define PLANT 0
#endif

#define OR_CENTRED 0 /* state.F90:15 CENTRED_STEPPING */
#define OR_FORWARD 1 /* state.F90:15 FORWARD_STEPPING */

typedef struct {
  int nz, ny, nx;
  double cx, cy;
  const double *rdz, *rdzn, *dz, *dzn, *tzc1, *tzc2, *tzd1, *tzd2;
} or_grid;

/* A flow field as seen by ultflx: either the model array itself (scalars) or
 * its dual-grid interpolation (momentum), stencil.F90:147-169. */
typedef struct {
  const double *a;
  int dk, dj, di; /* unit shift of the primal dimension of the advected field */
} or_flow;

static int g_nz, g_ny, g_nx; /* sizes of the current call (single-threaded) */
static const double R6 = 1.0 / 6.0; /* ultimateflux.F90:40 */

#define IDX(k, j, i) ((size_t)((k)-1) + (size_t)g_nz * ((size_t)((j)-1) + (size_t)(g_ny + 4) * (size_t)((i)-1)))
#define A(arr, k, j, i) ((arr)[IDX(k, j, i)])
#define V1(vec, k) ((vec)[(k)-1])

/* Fortran SIGN(a,b) for a>0: |a| with the sign bit of b (IEEE signed zero). */
static double fsign(double a, double b) { return copysign(fabs(a), b); }
static int fnint(double x) { return (int)lround(x); }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* The arm of the limiter the last flux_in_w / flux_in_h call took, from its
 * dmax(fncs, dmin(dmin(fnfs, fnrefs), 1)): 0 the face is fc (ftemp = fncs),
 * 1 fd (ftemp = 1), 2 ffs (ftemp = fnfs), 3 fu + fncs*hrcfl*fdels (ftemp =
 * fnrefs).  The upwind branch (|fcurvs| >= |fdels|) is arm 0.  Recorded per
 * face for the adjoint's recorded branch masks. */
static int g_arm;
static double *m_arm_x, *m_arm_y, *m_arm_z; /* arm of each face of the column (NAN: none computed) */
static void set_arm(double fnfs, double fncs, double fnrefs) {
  double m1 = dmin(fnfs, fnrefs), m2 = dmin(m1, 1.0);
#if PLANT == 10
  g_arm = fncs > m2 ? 0 : (m1 < 1.0 ? (fnfs < fnrefs ? 3 : 2) : 1);
#elif PLANT == 11
  g_arm = fncs > m2 ? 1 : (m1 < 1.0 ? (fnfs < fnrefs ? 2 : 3) : 0);
#else
  g_arm = fncs > m2 ? 0 : (m1 < 1.0 ? (fnfs < fnrefs ? 2 : 3) : 1);
#endif
}

/* stencil.F90:147-169 calculate_interpolated_cell_value: raw copy where the
 * shifted neighbour leaves [1, end+halo] or for the z shift at k=nz. */
static double flow_at(const or_flow *f, int k, int j, int i) {
  if (f->dk == 0 && f->dj == 0 && f->di == 0) return A(f->a, k, j, i);
  int jy = j + f->dj, ix = i + f->di;
  if (jy < 1 || jy > g_ny + 4 || ix < 1 || ix > g_nx + 4 || (f->dk && k == g_nz)) return A(f->a, k, j, i);
  return 0.5 * (A(f->a, k, j, i) + A(f->a, k + f->dk, jy, ix));
}
#define FL(f, k, j, i) flow_at((f), (k), (j), (i))

/* ultimateflux.F90:525-554 calculate_flux_in_w */
static double flux_in_w(int k, double wval, double dt, double sum_cfl_out, double rdc, double fdels, double fd,
                        double fc, double fu, double rdu, double fgt1, double fgt2, const double *dzn,
                        const double *rdzn) {
  double dd = V1(dzn, PLANT == 5 ? k : k + 1), rdd = V1(rdzn, k + 1);
  double cflmod = fabs(wval * V1(rdzn, k + 1) * dt);
  double hrcfl = 1.0 / (sum_cfl_out + 1.e-30);
  double hcfl = 0.5 * cflmod;
  double cflcoef = R6 * (1.0 - cflmod * cflmod) * (dd * dd * rdc);
  double rfdels = 1.0 / (fdels + 1.e-30);
  double ffs = 0.5 * (fd + fc) - hcfl * (fd - fc) - cflcoef * ((fd - fc) * rdd - (fc - fu) * rdu) - fgt1 - fgt2;
  double fnfs = (ffs - fu) * rfdels;
  double fncs = (fc - fu) * rfdels;
  double fnrefs = fncs * hrcfl;
  double ftemp = dmax(fncs, dmin(dmin(fnfs, fnrefs), 1.0));
  set_arm(fnfs, fncs, fnrefs);
  return ftemp * fdels + fu;
}

/* ultimateflux.F90:556-580 calculate_flux_in_x; :582-606 calculate_flux_in_y
 * (identical arithmetic with cx or cy) */
static double flux_in_h(double vel, double dt, double sum_cfl_out, double fdels, double fd, double fc, double fu,
                        double fgt1, double fgt2, double ch, double fcurvs) {
  double cflmod = fabs(vel * ch * dt);
  double hrcfl = 1.0 / (sum_cfl_out + 1.e-30);
  double hcfl = 0.5 * cflmod;
  double cflcoef = R6 * (1.0 - cflmod * cflmod);
  double rfdels = 1.0 / (fdels + 1.e-30);
  double ffs = 0.5 * (fd + fc) - hcfl * (fd - fc) - cflcoef * fcurvs - fgt1 - fgt2;
  double fnfs = (ffs - fu) * rfdels;
  double fncs = (fc - fu) * rfdels;
  double fnrefs = fncs * hrcfl;
  double ftemp = dmax(fncs, dmin(dmin(fnfs, fnrefs), 1.0));
  set_arm(fnfs, fncs, fnrefs);
  return ftemp * fdels + fu;
}

/* ultimateflux.F90:678-695 calculate_gradient_term_in_z; src2 may be NULL */
static double grad_z(int k, const or_flow *W, int jc1, int ic1, int jc2, int ic2, double dt, double advneg,
                     double advpos, int kdof, int kp0, int kp1, const double *F, int js1, int is1, int has_src2,
                     int js2, int is2, const double *rdz) {
  double sf = 0.125 * (FL(W, k, jc1, ic1) + FL(W, k - 1, jc1, ic1) + FL(W, k, jc2, ic2) + FL(W, k - 1, jc2, ic2));
  int kposwf = fnint(0.5 + fsign(0.5, sf));
  int ksu = kp1 * (1 - kposwf) + (k - 1) * kposwf;
  int ksl = kp0 * (1 - kposwf) + k * kposwf;
  double temp = A(F, ksl, js1, is1) - A(F, ksu, js1, is1);
  if (has_src2) temp = temp * advpos + (A(F, ksl, js2, is2) - A(F, ksu, js2, is2)) * advneg;
  return fabs(sf) * V1(rdz, k + kdof) * dt * temp;
}

/* ultimateflux.F90:697-709 calculate_gradient_term_in_x.  The four flow
 * columns are (j1,i1),(j2,i2) at level k and (j3,i3),(j4,i4) at level k+kpv. */
static double grad_x(int k, const or_flow *U, int j1, int i1, int j2, int i2, int j3, int i3, int j4, int i4,
                     double cx, double dt, double fc, double src1, double src2, int kpv) {
  double sf = 0.125 * (FL(U, k, j1, i1) + FL(U, k, j2, i2) + FL(U, k + kpv, j3, i3) + FL(U, k + kpv, j4, i4));
  double posuf = 0.5 + fsign(0.5, sf);
  return fabs(sf) * cx * dt * (fc - (src1 * (1. - posuf) + src2 * posuf));
}

/* ultimateflux.F90:711-761 calculate_gradient_term_in_y, "data_two" form used
 * by the x faces: flows V at (jd,id) and (jd2,id2), level k */
static double grad_y_two(int k, const or_flow *V, int jd, int id, int jd2, int id2, double cy, double dt, double fc,
                         const double *F, int jt1, int it1, double m1, int jt2, int it2, double m2) {
  double sf = 0.125 * (FL(V, k, jd, id) + FL(V, k, jd2, id2) + FL(V, k, jd - 1, id) + FL(V, k, jd2 - 1, id2));
  int s = fnint(fsign(1.0, sf));
  double t1 = A(F, k, jt1 - s, it1) * m1;
  double t2 = A(F, k, jt2 - s, it2) * m2;
  t1 = t1 + t2;
  return fabs(sf) * cy * dt * (fc - t1);
}

/* same function, single-source form used by the z faces: flows at levels k, k+1 */
static double grad_y_one(int k, const or_flow *V, int jd, int id, double cy, double dt, double fc, const double *F,
                         int jt, int it, int koff) {
  double sf = 0.125 * (FL(V, k, jd - 1, id) + FL(V, k, jd, id) + FL(V, k + 1, jd - 1, id) + FL(V, k + 1, jd, id));
  int s = fnint(fsign(1.0, sf));
  double t1 = A(F, k + koff, jt - s, it);
  return fabs(sf) * cy * dt * (fc - t1);
}

typedef struct {
  const or_flow *U, *V, *W;
  const double *F;
  int jf, if_; /* flow column (absolute, the reference's central (1,1)) */
  int js, is;  /* scalar column */
  int kdof;
  double dt;
  const double *rdz, *rdzn, *dzn; /* as passed to ultflx (swapped for w) */
  double cx, cy;
} or_col;

/* ultimateflux.F90:236-305 handle_vertical_fluxes_middleofcolumn */
static void vert_middle(const or_col *c, double *flux_z) {
  const or_flow *U = c->U, *V = c->V, *W = c->W;
  const double *F = c->F;
  int jf = c->jf, ifl = c->if_, js = c->js, is = c->is;
  for (int k = 2; k <= g_nz - 2; k++) {
    /* :616-676 calculate_stencil_for_w */
    double sm = fsign(0.5, FL(W, k, jf, ifl));
    double advpos = 0.5 + sm, advneg = 0.5 - sm;
    int ko = fnint(advneg);
    double fu = A(F, k - 1, js, is) * advpos + A(F, k + 2, js, is) * advneg;
    double fc = A(F, k + ko, js, is);
    double fd = A(F, k + 1, js, is) * advpos + A(F, k, js, is) * advneg;
    double fcurvs = fd - 2.0 * fc + fu;
    double fdels = fd - fu;
    if (fabs(fcurvs) >= fabs(fdels)) {
      flux_z[k + 1] = fc;
      m_arm_z[k + 1] = 0;
    } else {
      double s = 0.0;
      s = s + V1(c->rdz, PLANT == 7 ? k : k + ko) *
                  (dmax(0.0, FL(W, k + ko, jf, ifl)) + fabs(dmin(0.0, FL(W, k - 1 + ko, jf, ifl))));
      s = s + c->cy * (dmax(0.0, FL(V, k + ko, jf, ifl)) + fabs(dmin(0.0, FL(V, k + ko, jf - 1, ifl))));
      s = s + c->cx * (dmax(0.0, FL(U, k + ko, jf, ifl)) + fabs(dmin(0.0, FL(U, k + ko, jf, ifl - 1))));
      s = s * c->dt;
      double fgt1 = grad_y_one(k, V, jf, ifl, c->cy, c->dt, fc, F, js, is, ko);
      double fgt2 = grad_x(k, U, jf, ifl - 1, jf, ifl, jf, ifl - 1, jf, ifl, c->cx, c->dt, fc, A(F, k + ko, js, is + 1),
                           A(F, k + ko, js, is - 1), 1);
      double rdu = PLANT == 1 ? V1(c->rdzn, k + 2) * advpos + V1(c->rdzn, k) * advneg
                              : V1(c->rdzn, k) * advpos + V1(c->rdzn, k + 2) * advneg;
      double rdc = PLANT == 4 ? V1(c->rdz, k + 1 + c->kdof) * advpos + V1(c->rdz, k + c->kdof) * advneg
                              : V1(c->rdz, k + c->kdof) * advpos + V1(c->rdz, k + 1 + c->kdof) * advneg;
      flux_z[k + 1] = flux_in_w(k, FL(W, k, jf, ifl), c->dt, s, rdc, fdels, fd, fc, fu, rdu, fgt1, fgt2, c->dzn,
                                c->rdzn);
      m_arm_z[k + 1] = g_arm;
    }
  }
}

/* ultimateflux.F90:307-411 handle_vertical_fluxes_bottomofcolumn (w only) */
static void vert_bottom(const or_col *c, double *flux_z) {
  const or_flow *U = c->U, *V = c->V, *W = c->W;
  const double *F = c->F;
  int jf = c->jf, ifl = c->if_, js = c->js, is = c->is;
  int k = 1;
  double fu, fc, fd;
  double wk = FL(W, k, jf, ifl);
  if (wk >= 0.0) {
    fu = A(F, 1, js, is);
    fc = A(F, k, js, is);
    fd = A(F, k + 1, js, is);
  } else {
    fu = A(F, k + 2, js, is);
    fc = A(F, k + 1, js, is);
    fd = A(F, k, js, is);
  }
  double fcurvs = fd - 2.0 * fc + fu;
  double fdels = fd - fu;
  if (fabs(fcurvs) >= fabs(fdels)) {
    flux_z[k + 1] = fc;
    m_arm_z[k + 1] = 0;
    return;
  }
  double s = 0.0, fgt1, fgt2, rdu, rdc;
  if (wk >= 0.0) {
    s = s + wk * V1(c->rdz, k + c->kdof);
    s = s + c->cy * (dmax(0.0, FL(V, k, jf, ifl)) + fabs(dmin(0.0, FL(V, k, jf - 1, ifl))));
    s = s + c->cx * (dmax(0.0, FL(U, k, jf, ifl)) + fabs(dmin(0.0, FL(U, k, jf, ifl - 1))));
    s = s * c->dt;
    fgt1 = grad_y_one(k, V, jf, ifl, c->cy, c->dt, fc, F, js, is, 0);
    fgt2 = grad_x(k, U, jf, ifl - 1, jf, ifl, jf, ifl - 1, jf, ifl, c->cx, c->dt, fc, A(F, k, js, is + 1),
                  A(F, k, js, is - 1), 1);
    rdu = V1(c->rdzn, 2);
    rdc = V1(c->rdz, k + c->kdof);
  } else {
    s = s + V1(c->rdz, k + 1 + c->kdof) * (dmax(0.0, FL(W, k + 1, jf, ifl)) - wk);
    s = s + c->cy * (dmax(0.0, FL(V, k + 1, jf, ifl)) + fabs(dmin(0.0, FL(V, k + 1, jf - 1, ifl))));
    s = s + c->cx * (dmax(0.0, FL(U, k + 1, jf, ifl)) + fabs(dmin(0.0, FL(U, k + 1, jf, ifl - 1))));
    s = s * c->dt;
    fgt1 = grad_y_one(k, V, jf, ifl, c->cy, c->dt, fc, F, js, is, 1);
    fgt2 = grad_x(k, U, jf, ifl - 1, jf, ifl, jf, ifl - 1, jf, ifl, c->cx, c->dt, fc, A(F, k + 1, js, is + 1),
                  A(F, k + 1, js, is - 1), 1);
    rdu = V1(c->rdzn, PLANT == 6 ? k + 1 : k + 2);
    rdc = V1(c->rdz, PLANT == 9 ? k + c->kdof : k + 1 + c->kdof);
  }
  flux_z[k + 1] = flux_in_w(k, wk, c->dt, s, rdc, fdels, fd, fc, fu, rdu, fgt1, fgt2, c->dzn, c->rdzn);
  m_arm_z[k + 1] = g_arm;
}

/* ultimateflux.F90:413-523 handle_vertical_fluxes_topofcolumn */
static void vert_top(const or_col *c, double *flux_z) {
  const or_flow *U = c->U, *V = c->V, *W = c->W;
  const double *F = c->F;
  int jf = c->jf, ifl = c->if_, js = c->js, is = c->is;
  int nz = g_nz;
  int k = nz - 1;
  double fu, fc, fd;
  double wk = FL(W, k, jf, ifl);
  if (wk >= 0.0) {
    fu = A(F, k - 1, js, is);
    fc = A(F, k, js, is);
    fd = A(F, k + 1, js, is);
  } else {
    fu = 2.0 * A(F, k + 1, js, is) - A(F, k, js, is);
    fc = A(F, k + 1, js, is);
    fd = A(F, k, js, is);
  }
  double fcurvs = fd - 2.0 * fc + fu;
  double fdels = fd - fu;
  if (fabs(fcurvs) >= fabs(fdels)) {
    flux_z[k + 1] = fc;
    m_arm_z[k + 1] = 0;
    return;
  }
  double s = 0.0, fgt1, fgt2, rdu, rdc;
  if (wk >= 0.0) {
    s = s + V1(c->rdz, k + c->kdof) * (wk + fabs(dmin(0.0, FL(W, k - 1, jf, ifl))));
    s = s + c->cy * (dmax(0.0, FL(V, k, jf, ifl)) + fabs(dmin(0.0, FL(V, k, jf - 1, ifl))));
    s = s + c->cx * (dmax(0.0, FL(U, k, jf, ifl)) + fabs(dmin(0.0, FL(U, k, jf, ifl - 1))));
    s = s * c->dt;
    fgt1 = grad_y_one(k, V, jf, ifl, c->cy, c->dt, fc, F, js, is, 0);
    fgt2 = grad_x(k, U, jf, ifl - 1, jf, ifl, jf, ifl - 1, jf, ifl, c->cx, c->dt, fc, A(F, k, js, is + 1),
                  A(F, k, js, is - 1), 1);
    rdu = V1(c->rdzn, k);
    rdc = V1(c->rdz, k + c->kdof);
  } else {
    int kt = (k + 1 + c->kdof < nz) ? k + 1 + c->kdof : nz;
    s = s + (-(wk * V1(c->rdz, PLANT == 8 ? nz - 1 : kt)));
    s = s + c->cy * (dmax(0.0, FL(V, k + 1, jf, ifl)) + fabs(dmin(0.0, FL(V, k + 1, jf - 1, ifl))));
    s = s + c->cx * (dmax(0.0, FL(U, k + 1, jf, ifl)) + fabs(dmin(0.0, FL(U, k + 1, jf, ifl - 1))));
    s = s * c->dt;
    fgt1 = grad_y_one(k, V, jf, ifl, c->cy, c->dt, fc, F, js, is, 1);
    fgt2 = grad_x(k, U, jf, ifl - 1, jf, ifl, jf, ifl - 1, jf, ifl, c->cx, c->dt, fc, A(F, k + 1, js, is + 1),
                  A(F, k + 1, js, is - 1), 1);
    rdu = V1(c->rdzn, PLANT == 2 ? nz - 1 : nz);
    rdc = V1(c->rdz, kt);
  }
  flux_z[k + 1] = flux_in_w(k, wk, c->dt, s, rdc, fdels, fd, fc, fu, rdu, fgt1, fgt2, c->dzn, c->rdzn);
  m_arm_z[k + 1] = g_arm;
}

/* ultimateflux.F90:81-166 handle_x_direction_fluxes: right face of column is */
static void x_faces(const or_col *c, int kmax, double *flux_x) {
  const or_flow *U = c->U, *V = c->V, *W = c->W;
  const double *F = c->F;
  int jf = c->jf, ifl = c->if_, js = c->js, is = c->is;
  for (int k = 2; k <= kmax; k++) {
    int kp1 = (k + 1 < g_nz) ? k + 1 : g_nz;
    int kp0 = kp1 - 1;
    /* :616-631 calculate_stencil_for_u */
    double uk = FL(U, k, jf, ifl);
    double sm = fsign(0.5, uk);
    double advpos = 0.5 + sm, advneg = 0.5 - sm;
    double fu = A(F, k, js, is - 1) * advpos + A(F, k, js, is + 2) * advneg;
    double fc = A(F, k, js, is) * advpos + A(F, k, js, is + 1) * advneg;
    double fd = A(F, k, js, is + 1) * advpos + A(F, k, js, is) * advneg;
    double fdels = fd - fu;
    double fcurvs = fd - 2.0 * fc + fu;
    if (fabs(fcurvs) >= fabs(fdels)) {
      flux_x[k] = fc;
      m_arm_x[k] = 0;
      continue;
    }
    double s = 0.0;
    if (uk > 0.0) {
      s = s + V1(c->rdz, k) * (dmax(0.0, FL(W, k, jf, ifl)) + fabs(dmin(0.0, FL(W, k - 1, jf, ifl))));
      s = s + c->cy * (dmax(0.0, FL(V, k, jf, ifl)) + fabs(dmin(0.0, FL(V, k, jf - 1, ifl))));
      s = s + c->cx * (uk + fabs(dmin(0.0, FL(U, k, jf, ifl - 1))));
      s = s * c->dt;
    } else {
      s = s + V1(c->rdz, k) * (dmax(0.0, FL(W, k, jf, ifl + 1)) + fabs(dmin(0.0, FL(W, k - 1, jf, ifl + 1))));
      s = s + c->cy * (dmax(0.0, FL(V, k, jf, ifl + 1)) + fabs(dmin(0.0, FL(V, k, jf - 1, ifl + 1))));
      s = s + c->cx * (dmax(0.0, FL(U, k, jf, ifl + 1)) - uk);
      s = s * c->dt;
    }
    double fgt1 = grad_z(k, W, jf, ifl, jf, ifl + 1, c->dt, advneg, advpos, c->kdof, kp0, kp1, F, js, is, 1, js,
                         is + 1, c->rdz);
    double fgt2 =
        grad_y_two(k, V, jf, ifl + 1, jf, ifl, c->cy, c->dt, fc, F, js, is, advpos, js, is + 1, advneg);
    flux_x[k] = flux_in_h(uk, c->dt, s, fdels, fd, fc, fu, fgt1, fgt2, c->cx, fcurvs);
    m_arm_x[k] = g_arm;
  }
}

/* ultimateflux.F90:168-234 handle_y_direction_fluxes: north face of row js */
static void y_faces(const or_col *c, int kmax, double *flux_y) {
  const or_flow *U = c->U, *V = c->V, *W = c->W;
  const double *F = c->F;
  int jf = c->jf, ifl = c->if_, js = c->js, is = c->is;
  for (int k = 2; k <= kmax; k++) {
    int kp1 = (k + 1 < g_nz) ? k + 1 : g_nz;
    int kp0 = kp1 - 1;
    /* :633-652 calculate_stencil_for_y */
    double vk = FL(V, k, jf, ifl);
    double sm = fsign(0.5, vk);
    double advpos = 0.5 + sm, advneg = 0.5 - sm;
    int jo = fnint(advneg);
    int jp = (js > 1) ? 0 : 1;
    double fu = A(F, k, js - 1 + jp, is) * advpos + A(F, k, js + 2, is) * advneg;
    double fc = A(F, k, js, is) * advpos + A(F, k, js + 1, is) * advneg;
    double fd = A(F, k, js + 1, is) * advpos + A(F, k, js, is) * advneg;
    double fdels = fd - fu;
    double fcurvs = fd - 2.0 * fc + fu;
    if (fabs(fcurvs) >= fabs(fdels)) {
      flux_y[k] = fc;
      m_arm_y[k] = 0;
      continue;
    }
    double s = 0.0;
    s = s + V1(c->rdz, k) * (dmax(0.0, FL(W, k, jf + jo, ifl)) + fabs(dmin(0.0, FL(W, k - 1, jf + jo, ifl))));
    s = s + c->cy * (dmax(0.0, FL(V, k, jf + jo, ifl)) + fabs(dmin(0.0, FL(V, k, jf - 1 + jo + jp, ifl))));
    s = s + c->cx * (dmax(0.0, FL(U, k, jf + jo, ifl)) + fabs(dmin(0.0, FL(U, k, jf + jo, ifl - 1))));
    s = s * c->dt;
    double fgt1 = grad_z(k, W, jf, ifl, jf + 1, ifl, c->dt, advneg, advpos, c->kdof, kp0, kp1, F, js + jo, is, 0,
                         0, 0, c->rdz);
    double fgt2 = grad_x(k, U, jf, ifl - 1, jf, ifl, jf + 1, ifl - 1, jf + 1, ifl, c->cx, c->dt, fc,
                         A(F, k, js + jo, is + 1), A(F, k, js + jo, is - 1), 0);
    flux_y[k] = flux_in_h(vk, c->dt, s, fdels, fd, fc, fu, fgt1, fgt2, c->cy, fcurvs);
    m_arm_y[k] = g_arm;
  }
}

/* Per-field persistent scratch, prognostics.F90:13-20 */
typedef struct {
  double *flux_previous_x; /* (nz, ny+4) */
  double *flux_previous_y; /* (nz) */
  double *flux_y_buffer;   /* (nz) */
} or_fieldscratch;

/* Module-level scratch of tvdadvection.F90:26-54 (1-based, index 0 unused) */
static double *m_flux_x, *m_flux_y, *m_flux_z;

/* Optional record of face values per column, for the face-level checks */
typedef struct {
  double *fx, *fy, *fz; /* each (nz, ny+4, nx+4) or NULL */
  double *ax, *ay, *az; /* the faces' limiter arms (g_arm), same layout, or NULL */
} or_facerec;

/* ultimateflux.F90:23-60 ultflx on a single rank (my_coords = dim_sizes-1 = 0) */
static void ultflx(const or_col *c, int kmin, int kmax, double *flux_previous_x_col) {
  for (int k = 0; k <= g_nz + 1; k++) m_arm_x[k] = m_arm_y[k] = m_arm_z[k] = NAN;
  /* :62-79 handle_vertical_fluxes */
  if (kmin == 1) vert_bottom(c, m_flux_z);
  vert_top(c, m_flux_z);
  vert_middle(c, m_flux_z);
  if (kmin > 1) m_flux_z[2] = 0.0;
  /* :44-52: y faces skipped on the last row; the wrap supplies them */
  if (c->js != g_ny + 2) y_faces(c, kmax, m_flux_y);
  /* :54-58 */
  memcpy(m_flux_x + 1, flux_previous_x_col, sizeof(double) * (size_t)g_nz);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(g_nz + 1));
  memcpy(tmp + 1, flux_previous_x_col, sizeof(double) * (size_t)g_nz);
  x_faces(c, kmax, tmp);
  memcpy(flux_previous_x_col, tmp + 1, sizeof(double) * (size_t)g_nz);
  free(tmp);
}

/* tvdadvection.F90:911-960 differentiate_face_values */
static void differentiate(const or_col *c, double *s, const double *flux_y_previous, const double *flux_x_previous,
                          const double *t1, const double *t2, int top) {
  const or_flow *U = c->U, *V = c->V, *W = c->W;
  int jf = c->jf, ifl = c->if_, js = c->js, is = c->is;
  for (int k = 2; k <= g_nz - 1; k++) {
    double *p = &A(s, k, js, is);
    *p = *p + (FL(V, k, jf - 1, ifl) * flux_y_previous[k] - FL(V, k, jf, ifl) * m_flux_y[k]) * c->cy;
    *p = *p + 4.0 * (FL(W, k - 1, jf, ifl) * m_flux_z[k] * V1(t1, k) - FL(W, k, jf, ifl) * m_flux_z[k + 1] * V1(t2, k));
    *p = *p + (FL(U, k, jf, ifl - 1) * m_flux_x[k] - FL(U, k, jf, ifl) * flux_x_previous[k]) * c->cx;
  }
  if (top) {
    int k = g_nz;
    double *p = &A(s, k, js, is);
    *p = 0.0;
    *p = *p + (FL(V, k, jf - 1, ifl) * flux_y_previous[k] - FL(V, k, jf, ifl) * m_flux_y[k]) * c->cy;
    *p = *p + 4.0 * V1(t1, k) * FL(W, k - 1, jf, ifl) * m_flux_z[k];
    *p = *p + (FL(U, k, jf, ifl - 1) * m_flux_x[k] - FL(U, k, jf, ifl) * flux_x_previous[k]) * c->cx;
  }
}

/* tvdadvection.F90:722-762 advect_scalar_field and :765-908 advect_u/v/w share
 * this driver; `mom` selects the dual-grid flows and w's metric swap. */
static void advect_field(const or_grid *g, int x, int y, int halo_column, double dt, const or_flow *U,
                         const or_flow *V, const or_flow *W, const double *F, double *s, or_fieldscratch *fs,
                         int is_w, const or_facerec *rec) {
  or_col c;
  c.U = U;
  c.V = V;
  c.W = W;
  c.F = F;
  c.jf = y;
  c.if_ = x;
  c.js = y;
  c.is = x;
  c.dt = dt;
  c.cx = g->cx;
  c.cy = g->cy;
  int kmin, kmax;
  const double *t1, *t2;
  if (is_w) { /* :881-904 */
    c.kdof = 1;
    c.rdz = PLANT == 3 ? g->rdz : g->rdzn;
    c.rdzn = PLANT == 3 ? g->rdzn : g->rdz;
    c.dzn = PLANT == 3 ? g->dzn : g->dz;
    kmin = 1;
    kmax = g_nz - 1;
    t1 = g->tzd1;
    t2 = g->tzd2;
  } else { /* :739-757 */
    c.kdof = 0;
    c.rdz = g->rdz;
    c.rdzn = g->rdzn;
    c.dzn = g->dzn;
    kmin = 2;
    kmax = g_nz;
    t1 = g->tzc1;
    t2 = g->tzc2;
  }
  double *fpx_col = fs->flux_previous_x + (size_t)g_nz * (size_t)(y - 1);
  ultflx(&c, kmin, kmax, fpx_col);
  /* :1014-1029 complete_y_flux_wrap_recv_if_required (local copy) */
  if (y == g_ny + 2) memcpy(m_flux_y + 1, fs->flux_y_buffer, sizeof(double) * (size_t)g_nz);
  if (rec) {
    for (int k = 1; k <= g_nz; k++) {
      if (rec->fx) A(rec->fx, k, y, x) = fpx_col[k - 1];
      if (rec->fy) A(rec->fy, k, y, x) = m_flux_y[k];
      if (rec->fz) A(rec->fz, k, y, x) = m_flux_z[k];
      if (rec->ax) A(rec->ax, k, y, x) = m_arm_x[k];
      if (rec->ay) A(rec->ay, k, y, x) = m_arm_y[k];
      if (rec->az) A(rec->az, k, y, x) = m_arm_z[k];
    }
  }
  if (!halo_column) {
    /* flux_previous_y is 1-based in differentiate: shift pointer */
    double fpy[g_nz + 1], fpx[g_nz + 1];
    memcpy(fpy + 1, fs->flux_previous_y, sizeof(double) * (size_t)g_nz);
    memcpy(fpx + 1, fpx_col, sizeof(double) * (size_t)g_nz);
    differentiate(&c, s, fpy, fpx, t1, t2, !is_w);
  }
  /* :759 */
  if (y < g_ny + 2) memcpy(fs->flux_previous_y, m_flux_y + 1, sizeof(double) * (size_t)g_nz);
  /* :990-1006 register_y_flux_wrap_send_if_required (local copy) */
  if (y == 2) memcpy(fs->flux_y_buffer, m_flux_y + 1, sizeof(double) * (size_t)g_nz);
}

static or_fieldscratch scratch_new(void) {
  or_fieldscratch f;
  f.flux_previous_x = (double *)calloc((size_t)g_nz * (size_t)(g_ny + 4), sizeof(double));
  f.flux_previous_y = (double *)calloc((size_t)g_nz, sizeof(double));
  f.flux_y_buffer = (double *)calloc((size_t)g_nz, sizeof(double));
  return f;
}
static void scratch_free(or_fieldscratch *f) {
  free(f->flux_previous_x);
  free(f->flux_previous_y);
  free(f->flux_y_buffer);
}

/*
 * One call of tvdadvection's timestep_callback over every column, in the order
 * of timestepper.F90:65-85 (x outer, y inner, halos included).
 *
 * fields[] / zfields[] / sfields[]: index 0,1,2 = u,v,w; 3 = th; 4.. = the
 * n_q q fields, then the n_tracers tracers (all advected by
 * advect_scalar_field).  `advect` bit 0 = flow fields (advect_flow), bit 1 =
 * theta (advect_th), bit 2 = q fields (advect_q), bit 3 = tracers
 * (advect_tracer, which the reference sets from the theta option,
 * tvdadvection.F90:349-352).  active[] (NULL: all active) is the per-field
 * `active` flag: theta is advected only when active (:706), each q field only
 * when active (:661); tracers and the flow fields are not gated (:603-646,
 * :684-694).  The s-fields are accumulated in place (the reference's
 * `s += ...`).
 * rec_field: index of the field whose face values are written to rec (-1: none),
 * and the limiter arms of its faces to rec_a* (each may be NULL): at column
 * (k, y, x) the right x face, the north y face and flux_z(k).
 */
int oracle_tvd_timestep(int nz, int ny, int nx, double cx, double cy, const double *rdz, const double *rdzn,
                        const double *dz, const double *dzn, const double *tzc1, const double *tzc2,
                        const double *tzd1, const double *tzd2, double dtm, int field_stepping,
                        int momentum_stepping, int scalar_stepping, int advect, int n_q, int n_tracers,
                        const int *active, const double *const *fields, const double *const *zfields, double *const *sfields,
                        int rec_field, double *rec_fx, double *rec_fy, double *rec_fz, double *rec_ax,
                        double *rec_ay, double *rec_az) {
  if (nz < 4 || ny < 1 || nx < 1 || n_q < 0 || n_tracers < 0) return 1;
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  or_grid g = {nz, ny, nx, cx, cy, rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2};
  m_flux_x = (double *)calloc((size_t)nz + 2, sizeof(double));
  m_flux_y = (double *)calloc((size_t)nz + 2, sizeof(double));
  m_flux_z = (double *)calloc((size_t)nz + 2, sizeof(double));
  m_arm_x = (double *)calloc((size_t)nz + 2, sizeof(double));
  m_arm_y = (double *)calloc((size_t)nz + 2, sizeof(double));
  m_arm_z = (double *)calloc((size_t)nz + 2, sizeof(double));
  int nf = 4 + n_q + n_tracers;
  or_fieldscratch *fs = (or_fieldscratch *)malloc(sizeof(or_fieldscratch) * (size_t)nf);
  for (int n = 0; n < nf; n++) fs[n] = scratch_new();
  or_facerec rec = {rec_fx, rec_fy, rec_fz, rec_ax, rec_ay, rec_az};

  /* tvdadvection.F90:608-609, 656-657, 704-705 */
  double dt_mom = (momentum_stepping == OR_FORWARD) ? dtm : dtm * 2.0;
  double dt_sca = (scalar_stepping == OR_FORWARD) ? dtm : dtm * 2.0;
  /* :739-749, :783-796: advected field is the current one only when forward stepping */
  const double *const *adv = (field_stepping == OR_FORWARD) ? fields : zfields;

  const double *u = fields[0], *v = fields[1], *w = fields[2];
  or_flow raw_u = {u, 0, 0, 0}, raw_v = {v, 0, 0, 0}, raw_w = {w, 0, 0, 0};
  /* dual flows: u is PRIMAL in x, v in y, w in z (simplesetup.F90:70-85) */
  or_flow du_u = {u, 0, 0, 1}, du_v = {v, 0, 0, 1}, du_w = {w, 0, 0, 1};
  or_flow dv_u = {u, 0, 1, 0}, dv_v = {v, 0, 1, 0}, dv_w = {w, 0, 1, 0};
  or_flow dw_u = {u, 1, 0, 0}, dw_v = {v, 1, 0, 0}, dw_w = {w, 1, 0, 0};

  for (int x = 1; x <= nx + 4; x++) {
    for (int y = 1; y <= ny + 4; y++) {
      int halo_column = (y < 3 || x < 3 || y > ny + 2 || x > nx + 2);
      if (halo_column) { /* tvdadvection.F90:579-586 */
        int keep = (y == 2 && x <= nx + 2 && x >= 2) || (x == 2 && y >= 3 && y <= ny + 2);
        if (!keep) continue;
      }
      if (advect & 1) {
        advect_field(&g, x, y, halo_column, dt_mom, &du_u, &du_v, &du_w, adv[0], sfields[0], &fs[0], 0,
                     rec_field == 0 ? &rec : NULL);
        advect_field(&g, x, y, halo_column, dt_mom, &dv_u, &dv_v, &dv_w, adv[1], sfields[1], &fs[1], 0,
                     rec_field == 1 ? &rec : NULL);
        advect_field(&g, x, y, halo_column, dt_mom, &dw_u, &dw_v, &dw_w, adv[2], sfields[2], &fs[2], 1,
                     rec_field == 2 ? &rec : NULL);
      }
      /* :591-594: flow, theta, q fields, tracers */
      if ((advect & 2) && (!active || active[3]))
        advect_field(&g, x, y, halo_column, dt_sca, &raw_u, &raw_v, &raw_w, adv[3], sfields[3], &fs[3], 0,
                     rec_field == 3 ? &rec : NULL);
      if (advect & 4)
        for (int n = 4; n < 4 + n_q; n++)
          if (!active || active[n])
            advect_field(&g, x, y, halo_column, dt_sca, &raw_u, &raw_v, &raw_w, adv[n], sfields[n], &fs[n], 0,
                         rec_field == n ? &rec : NULL);
      if (advect & 8)
        for (int n = 4 + n_q; n < nf; n++)
          advect_field(&g, x, y, halo_column, dt_sca, &raw_u, &raw_v, &raw_w, adv[n], sfields[n], &fs[n], 0,
                       rec_field == n ? &rec : NULL);
    }
  }
  for (int n = 0; n < nf; n++) scratch_free(&fs[n]);
  free(fs);
  free(m_flux_x);
  free(m_flux_y);
  free(m_flux_z);
  free(m_arm_x);
  free(m_arm_y);
  free(m_arm_z);
  return 0;
}

/* halocommunication.F90:1121-1168 (y then x over the interior range) and
 * :1058-1114 (the four 2x2 corners) for a single rank, halo depth 2. */
void oracle_halo_fill(double *f, int nz, int ny, int nx) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  int src_y[5] = {0, ny + 1, ny + 2, 3, 4}, tgt_y[5] = {0, 1, 2, ny + 3, ny + 4};
  int src_x[5] = {0, nx + 1, nx + 2, 3, 4}, tgt_x[5] = {0, 1, 2, nx + 3, nx + 4};
  for (int h = 1; h <= 4; h++)
    for (int i = 3; i <= nx + 2; i++)
      for (int k = 1; k <= nz; k++) A(f, k, tgt_y[h], i) = A(f, k, src_y[h], i);
  for (int h = 1; h <= 4; h++)
    for (int j = 3; j <= ny + 2; j++)
      for (int k = 1; k <= nz; k++) A(f, k, j, tgt_x[h]) = A(f, k, j, src_x[h]);
  /* corners: target (ty,tx) <- source (sy,sx), each a 2x2 block */
  int cs_y[4] = {ny + 1, ny + 1, 3, 3}, cs_x[4] = {nx + 1, 3, nx + 1, 3};
  int ct_y[4] = {1, 1, ny + 3, ny + 3}, ct_x[4] = {1, nx + 3, 1, nx + 3};
  for (int c = 0; c < 4; c++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++)
        for (int k = 1; k <= nz; k++) A(f, k, ct_y[c] + dy, ct_x[c] + dx) = A(f, k, cs_y[c] + dy, cs_x[c] + dx);
}

/* gridmanager.F90:2449-2456 (level spacings) and :2048-2067 (advection
 * weights).  z, zn, rho, rhon: length nz, 1-based.  Outputs length nz; index 1
 * entries the reference never sets are returned as 0. */
void oracle_vertical_metrics(int nz, const double *z, const double *zn, const double *rho, const double *rhon,
                             double *dz, double *dzn, double *rdz, double *rdzn, double *tzc1, double *tzc2,
                             double *tzd1, double *tzd2) {
  for (int k = 0; k < nz; k++) dz[k] = dzn[k] = rdz[k] = rdzn[k] = tzc1[k] = tzc2[k] = tzd1[k] = tzd2[k] = 0.0;
  for (int k = 2; k <= nz; k++) {
    V1(dz, k) = V1(z, k) - V1(z, k - 1);
    V1(dzn, k) = V1(zn, k) - V1(zn, k - 1);
    V1(rdz, k) = 1. / V1(dz, k);
    V1(rdzn, k) = 1. / V1(dzn, k);
  }
  V1(dzn, 1) = 0.0;
  for (int k = 2; k <= nz - 1; k++) {
    V1(tzc1, k) = 0.25 * V1(rdz, k) * V1(rho, k - 1) / V1(rhon, k);
    V1(tzc2, k) = 0.25 * V1(rdz, k) * V1(rho, k) / V1(rhon, k);
  }
  for (int k = 2; k <= nz - 1; k++) {
    V1(tzd1, k) = 0.25 * V1(rdzn, k + 1) * V1(rhon, k) / V1(rho, k);
    V1(tzd2, k) = 0.25 * V1(rdzn, k + 1) * V1(rhon, k + 1) / V1(rho, k);
  }
  int k = nz;
  V1(tzc2, k) = 0.25 * V1(rdz, k) * V1(rho, k) / V1(rhon, k);
  V1(tzc1, k) = 0.25 * V1(rdz, k) * V1(rho, k - 1) / V1(rhon, k);
}

/* gridmanager.F90:2624-2633: dx = real(dxx) (default real, i.e. rounded to
 * single precision), cx = 1./dx */
double oracle_horizontal_c(double dxx) {
  double dx = (double)(float)dxx;
  return 1. / dx;
}

/* stepfields.F90:597-627 with direction == FORWARD_STEPPING and no galilean
 * shift: zfield = field + dtm * sfield for k = 1..nz, after sfield(1) = 0,
 * over the interior columns. */
void oracle_forward_step(double *znew, const double *f, double *s, int nz, int ny, int nx, double dtm) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  for (int i = 3; i <= nx + 2; i++)
    for (int j = 3; j <= ny + 2; j++) {
      A(s, 1, j, i) = 0.0;
      for (int k = 1; k <= nz; k++) A(znew, k, j, i) = A(f, k, j, i) + dtm * A(s, k, j, i);
    }
}

/* tvdadvection.F90:1131-1163: the per-level total of a 3-D tendency over the
 * interior columns, accumulated column by column in the order the timestepper
 * visits them (timestepper.F90:65-85: x outer, y inner).  t has the
 * reference's tend_3d layout (nz, ny, nx). */
void oracle_profile_total(const double *t, int nz, int ny, int nx, double *out) {
  for (int k = 0; k < nz; k++) out[k] = 0.0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) out[k] = out[k] + t[k + (size_t)nz * ((size_t)j + (size_t)ny * (size_t)i)];
}

/* stepfields.F90:553-584 step_column_in_slice and :597-627 step_field for
 * one field over the interior columns: optional time smoothing
 * (perform_timesmooth_for_field, :522-533: field = c1*field + c2*zfield),
 * then sfield(1) = 0 and, per level, sav = zfield + gal (centred stepping,
 * when sav is given), field = gal + field (flow fields), and
 * zfield = field + dtm*s (forward) or gal + zfield + 2*dtm*s (centred), with
 * gal applied only to flow fields (actual_gal, :614). */
void oracle_step_field(double *f, double *zf, double *s, double *sav, int nz, int ny, int nx, int flow_field,
                       int direction, double dtm, double gal, int do_timesmoothing, double c1, double c2) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  const double dtm_x2 = 2.0 * dtm;
  const double actual_gal = flow_field ? gal : 0.0;
  for (int i = 3; i <= nx + 2; i++)
    for (int j = 3; j <= ny + 2; j++) {
      if (do_timesmoothing)
        for (int k = 1; k <= nz; k++) A(f, k, j, i) = c1 * A(f, k, j, i) + c2 * A(zf, k, j, i);
      A(s, 1, j, i) = 0.0;
      for (int k = 1; k <= nz; k++) {
        if (sav && direction == OR_CENTRED) A(sav, k, j, i) = A(zf, k, j, i) + actual_gal;
        if (flow_field) A(f, k, j, i) = actual_gal + A(f, k, j, i);
        if (direction == OR_FORWARD)
          A(zf, k, j, i) = A(f, k, j, i) + dtm * A(s, k, j, i);
        else
          A(zf, k, j, i) = actual_gal + A(zf, k, j, i) + dtm_x2 * A(s, k, j, i);
      }
    }
}

/* swapsmooth.F90 swap_and_smooth_classic with old_smoother = .false.
 * (c1 = 1, c2 = -1) for one field over the interior columns: existing =
 * field + zfield; field = existing*c1 + field*c2; zfield = existing - field. */
void oracle_swap_classic(double *f, double *zf, int nz, int ny, int nx) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  const double c1 = 1.0, c2 = -1.0;
  for (int i = 3; i <= nx + 2; i++)
    for (int j = 3; j <= ny + 2; j++)
      for (int k = 1; k <= nz; k++) {
        const double existing = A(f, k, j, i) + A(zf, k, j, i);
        A(f, k, j, i) = existing * c1 + A(f, k, j, i) * c2;
        A(zf, k, j, i) = existing - A(f, k, j, i);
      }
}

/* swapsmooth.F90 swap_and_smooth_robert_filter for one field over the interior
 * columns, c1 = 1 - 2*tsmth, c2 = tsmth.  kind 0 (u, v): field <- zfield,
 * zfield <- c1*existing + c2*(field + sav - gal); kind 1 (w): the same without
 * gal; kind 2 (th, q, tracers): zfield <- field + tsmth*zfield, field <- the
 * old zfield. */
void oracle_swap_robert(double *f, double *zf, const double *sav, int kind, double tsmth, double gal, int nz,
                        int ny, int nx) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  const double c1 = 1.0 - 2.0 * tsmth, c2 = tsmth;
  for (int i = 3; i <= nx + 2; i++)
    for (int j = 3; j <= ny + 2; j++)
      for (int k = 1; k <= nz; k++) {
        if (kind == 2) {
          const double existing = A(zf, k, j, i);
          A(zf, k, j, i) = A(f, k, j, i) + tsmth * existing;
          A(f, k, j, i) = existing;
        } else {
          const double existing = A(f, k, j, i);
          A(f, k, j, i) = A(zf, k, j, i);
          if (kind == 0)
            A(zf, k, j, i) = c1 * existing + c2 * (A(f, k, j, i) + A(sav, k, j, i) - gal);
          else
            A(zf, k, j, i) = c1 * existing + c2 * (A(f, k, j, i) + A(sav, k, j, i));
        }
      }
}

/* thadvection.F90:124-211 at every column (1..nx+4, 1..ny+4: the registry
 * calls the component for halo columns too and it does not skip them),
 * levels 2..nz, with the factors of initialisation_callback (:60-77):
 * fcoriol_over_G = fcoriol/G, multiplicative_factor = rate*thref0*fcoriol_over_G.
 * vertical_advection_of_reference_state (:165-179, anelastic only), then
 * advection_of_mean_baroclinicity (:186-211, when baroclinicity is set: the
 * reference's l_advect_mean_baroclinicity). */
void oracle_thadv(double *sth, const double *u, const double *v, const double *w, const double *tzc1,
                  const double *tzc2, const double *dthref, const double *thref, int nz, int ny, int nx,
                  int anelastic, int baroclinicity, double thref0, double fcoriol, double g, double rate_x,
                  double rate_y, double ugal, double vgal) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  const double fcoriol_over_g = fcoriol / g;
  const double mfx = rate_x * thref0 * fcoriol_over_g, mfy = rate_y * thref0 * fcoriol_over_g;
  for (int i = 1; i <= nx + 4; i++)
    for (int j = 1; j <= ny + 4; j++) {
      if (anelastic)
        for (int k = 2; k <= nz; k++) {
          const double sctmp1 = tzc1[k - 1] * 2.0 * dthref[k - 2];
          const double sctmp2 = tzc2[k - 1] * 2.0 * dthref[k - 1];
          A(sth, k, j, i) = A(sth, k, j, i) - (sctmp1 * A(w, k - 1, j, i) + sctmp2 * A(w, k, j, i));
        }
      if (baroclinicity) {
        for (int k = 2; k <= nz; k++) {
          if (anelastic)
            A(sth, k, j, i) = A(sth, k, j, i) + thref[k - 1] * fcoriol_over_g *
                                                    ((A(v, k, j, i) + vgal) * rate_x - (A(u, k, j, i) + ugal) * rate_y);
          else
            A(sth, k, j, i) = A(sth, k, j, i) + ((A(v, k, j, i) + vgal) * mfx - (A(u, k, j, i) + ugal) * mfy);
        }
      }
    }
}

/* pwadvection.F90:258-610 over the interior columns (halo columns return at
 * :307), all three directions active (U_ACTIVE, V_ACTIVE, W_ACTIVE), with
 * l_toplevel = .true. (:21).  Flow fields (advect_flow_fields, :466-610):
 * su, sv, sw are overwritten (k = 2..nz-1; su and sv also at k = nz without
 * the W(nz) term; sw not at k = nz).  th (advect_th_field, :397-460):
 * overwritten too (the reference's own `sth = cx*...` at :407 and :437).
 * q fields (advect_q_field, :327-391): added to sq(n).  tcx = 0.25/dx,
 * tcy = 0.25/dy (gridmanager.F90:2638-2641); cx, cy the model's 1/dx, 1/dy. */
static void pw_scalar(double *s, const double *f, const double *u, const double *v, const double *w,
                      const double *tzc1, const double *tzc2, double cx, double cy, int nz, int j, int i,
                      int overwrite) {
  for (int k = 2; k <= nz; k++) {
    double x = cx * 0.5 * (A(u, k, j, i - 1) * A(f, k, j, i - 1) - A(u, k, j, i) * A(f, k, j, i + 1));
    double acc = overwrite ? x : A(s, k, j, i) + x;
    acc = acc + cy * 0.5 * (A(v, k, j - 1, i) * A(f, k, j - 1, i) - A(v, k, j, i) * A(f, k, j + 1, i));
    if (k < nz)
      acc = acc + 2.0 * (tzc1[k - 1] * A(w, k - 1, j, i) * A(f, k - 1, j, i) -
                         tzc2[k - 1] * A(w, k, j, i) * A(f, k + 1, j, i));
    else
      acc = acc + tzc1[k - 1] * 2.0 * A(w, k - 1, j, i) * A(f, k - 1, j, i);
    A(s, k, j, i) = acc;
  }
}

void oracle_pw(double *su, double *sv, double *sw, double *sth, double *const *sq, const double *u,
               const double *v, const double *w, const double *th, const double *const *q, int n_q,
               const double *tzc1, const double *tzc2, const double *tzd1, const double *tzd2, double cx,
               double cy, double tcx, double tcy, int nz, int ny, int nx, int advect_flow, int advect_th,
               int advect_q) {
  g_nz = nz;
  g_ny = ny;
  g_nx = nx;
  for (int i = 3; i <= nx + 2; i++)
    for (int j = 3; j <= ny + 2; j++) {
      if (advect_flow) {
        for (int k = 2; k <= nz; k++) {
          const int top = k == nz;
          double a = tcx * (A(u, k, j, i - 1) * (A(u, k, j, i) + A(u, k, j, i - 1)) -
                            A(u, k, j, i + 1) * (A(u, k, j, i) + A(u, k, j, i + 1)));
          a = a + tcy * (A(u, k, j - 1, i) * (A(v, k, j - 1, i) + A(v, k, j - 1, i + 1)) -
                         A(u, k, j + 1, i) * (A(v, k, j, i) + A(v, k, j, i + 1)));
          if (!top)
            a = a + (tzc1[k - 1] * A(u, k - 1, j, i) * (A(w, k - 1, j, i) + A(w, k - 1, j, i + 1)) -
                     tzc2[k - 1] * A(u, k + 1, j, i) * (A(w, k, j, i) + A(w, k, j, i + 1)));
          else
            a = a + tzc1[k - 1] * A(u, k - 1, j, i) * (A(w, k - 1, j, i) + A(w, k - 1, j, i + 1));
          A(su, k, j, i) = a;

          double b = tcy * (A(v, k, j - 1, i) * (A(v, k, j, i) + A(v, k, j - 1, i)) -
                            A(v, k, j + 1, i) * (A(v, k, j, i) + A(v, k, j + 1, i)));
          b = b + tcx * (A(v, k, j, i - 1) * (A(u, k, j, i - 1) + A(u, k, j + 1, i - 1)) -
                         A(v, k, j, i + 1) * (A(u, k, j, i) + A(u, k, j + 1, i)));
          if (!top)
            b = b + (tzc1[k - 1] * A(v, k - 1, j, i) * (A(w, k - 1, j, i) + A(w, k - 1, j + 1, i)) -
                     tzc2[k - 1] * A(v, k + 1, j, i) * (A(w, k, j, i) + A(w, k, j + 1, i)));
          else
            b = b + tzc1[k - 1] * A(v, k - 1, j, i) * (A(w, k - 1, j, i) + A(w, k - 1, j + 1, i));
          A(sv, k, j, i) = b;

          if (!top) {
            double c = (tzd1[k - 1] * A(w, k - 1, j, i) * (A(w, k, j, i) + A(w, k - 1, j, i)) -
                        tzd2[k - 1] * A(w, k + 1, j, i) * (A(w, k, j, i) + A(w, k + 1, j, i)));
            c = c + tcx * (A(w, k, j, i - 1) * (A(u, k, j, i - 1) + A(u, k + 1, j, i - 1)) -
                           A(w, k, j, i + 1) * (A(u, k, j, i) + A(u, k + 1, j, i)));
            c = c + tcy * (A(w, k, j - 1, i) * (A(v, k, j - 1, i) + A(v, k + 1, j - 1, i)) -
                           A(w, k, j + 1, i) * (A(v, k, j, i) + A(v, k + 1, j, i)));
            A(sw, k, j, i) = c;
          }
        }
      }
      if (advect_th) pw_scalar(sth, th, u, v, w, tzc1, tzc2, cx, cy, nz, j, i, 1);
      if (advect_q)
        for (int n = 0; n < n_q; n++) pw_scalar(sq[n], q[n], u, v, w, tzc1, tzc2, cx, cy, nz, j, i, 0);
    }
}
