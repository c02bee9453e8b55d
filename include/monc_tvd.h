/*
 * monc_tvd.h -- C ABI of the MI355X TVD advection library (libmonc_tvd.so).
 *
 * The library replaces MONC's `tvd_advection` component
 * (components/tvdadvection/src/tvdadvection.F90) and the single-rank periodic
 * halo fill it relies on (components/haloswapper/src/haloswapper.F90:62-71 ->
 * model_core/src/communication/halocommunication.F90:1058-1168).  The
 * reference component is a set of callbacks on `model_state_type`
 * (model_core/src/components/monc_component.F90:62-118); the entry points below
 * are the plain-pointer form of those callbacks that a Fortran shim reaches
 * through iso_c_binding.
 *
 * Arrays are Fortran column-major data(k,j,i) = (z,y,x), z fastest, with the
 * reference's depth-2 x/y halo: shape (nz, ny+4, nx+4), as allocated by
 * components/simplesetup/src/simplesetup.F90:60-137.  Vertical vectors have
 * length nz (index 1 unused where the reference leaves it unset).
 *
 * Every function returns MONC_TVD_OK (0) or a non-zero status; the message of
 * the last failure on the calling thread is returned by monc_tvd_last_error().
 * The reference aborts through log_log(LOG_ERROR) (model_core/src/utils/
 * logging.F90:24,53,86); the Fortran shim turns a non-zero status into that call.
 * One context serves one host thread; calls on it run on the context's HIP
 * stream and the host-memory entry points synchronise before returning.
 */
#ifndef MONC_TVD_H
#define MONC_TVD_H

#ifdef __cplusplus
extern "C" {
#endif

#define MONC_TVD_OK 0
#define MONC_TVD_ERR_ARG 1   /* invalid argument (sizes, halo, pointer, index) */
#define MONC_TVD_ERR_HIP 2   /* a HIP runtime call failed (allocation, launch) */
#define MONC_TVD_ERR_STATE 3 /* a field the call needs was never set or bound */

/* state.F90:15 */
#define MONC_TVD_CENTRED_STEPPING 0
#define MONC_TVD_FORWARD_STEPPING 1

/* precision of the device-resident fields */
#define MONC_TVD_FP64 0
#define MONC_TVD_FP32 1

/* advect flags: the selectors of tvdadvection.F90:349-352 resolved to tvd
 * (advect_flow, advect_th, advect_q, advect_tracer).  The reference sets
 * advect_tracer from the *theta* option (advection_theta_field, :352): a
 * caller passes MONC_TVD_ADVECT_TRACER exactly when it passes
 * MONC_TVD_ADVECT_THETA for the same configuration. */
#define MONC_TVD_ADVECT_FLOW 1
#define MONC_TVD_ADVECT_THETA 2
#define MONC_TVD_ADVECT_Q 4
#define MONC_TVD_ADVECT_TRACER 8

/* field ids: u, v, w, th, then the n_q q fields q(1..n_q) and the n_tracers
 * tracers tracer(1..n_tracers) (state.F90:86-92; all advected by
 * advect_scalar_field, tvdadvection.F90:650-695) */
#define MONC_TVD_U 0
#define MONC_TVD_V 1
#define MONC_TVD_W 2
#define MONC_TVD_TH 3
#define MONC_TVD_Q(n) (4 + (n))
#define MONC_TVD_TRACER(n_q, n) (4 + (n_q) + (n))

/* field variants: the prognostic, its z-field (time level t-1, the advected
 * field under centred stepping) and its source (tendency) field,
 * state.F90:86-92 */
#define MONC_TVD_PROG 0
#define MONC_TVD_ZFIELD 1
#define MONC_TVD_SOURCE 2
/* differentiation support (no counterpart in the reference): the adjoint of a
 * field (interior values, zero halo) and its tangent (periodic halo) */
#define MONC_TVD_ADJOINT 3
#define MONC_TVD_TANGENT 4

typedef struct monc_tvd_ctx monc_tvd_ctx;

/*
 * Replaces tvdadvection.F90:286-481 initialisation_callback (plus the metric
 * vectors of components/gridmanager/src/gridmanager.F90:2048-2067, 2449-2456
 * and cx = 1/dx, cy = 1/dy of :2624-2633).  nz, ny, nx are the local interior
 * sizes -- MONC's local_grid%size(Z_INDEX), size(Y_INDEX), size(X_INDEX)
 * (grids.F90:14: Z_INDEX = 1, Y_INDEX = 2, X_INDEX = 3) -- halo must be 2
 * (global_config halo_depth); nz >= 4, ny, nx >= 4.  n_q is
 * number_q_fields and n_tracers is n_tracers of model_state_type
 * (state.F90:126).  Every field starts active.  Copies the FP64 metric
 * vectors to the device and allocates nothing else: fields are allocated by
 * monc_tvd_set_field or supplied by monc_tvd_bind_field.
 */
int monc_tvd_create(int nz, int ny, int nx, int halo, int n_q, int n_tracers, int advect_flags, double cx,
                    double cy, const double *rdz, const double *rdzn, const double *dz, const double *dzn,
                    const double *tzc1, const double *tzc2, const double *tzd1, const double *tzd2,
                    int precision, monc_tvd_ctx **out);

/* The `active` flag of a prognostic (prognostic_field_type%active,
 * prognostics.F90:13-20) for th and the q fields: the reference advects th
 * only when th%active (tvdadvection.F90:706) and q(i) only when q(i)%active
 * (:661); the flow fields and tracers are not gated (:603-646, :684-694), so
 * other field ids are rejected.  An inactive field is left untouched by
 * monc_tvd_tendencies and the stepping entry points. */
int monc_tvd_set_field_active(monc_tvd_ctx *ctx, int field, int active);

/* Replaces tvdadvection.F90:485-524 finalisation_callback: frees the context
 * and every buffer the library allocated (bound buffers are not freed). */
int monc_tvd_destroy(monc_tvd_ctx *ctx);

/* Host <-> device staging of one field (FP64 host array of shape
 * (nz, ny+4, nx+4)); converts to/from FP32 for an FP32 context. */
int monc_tvd_set_field(monc_tvd_ctx *ctx, int field, int variant, const double *host);
int monc_tvd_get_field(monc_tvd_ctx *ctx, int field, int variant, double *host);

/* Device-resident mode: use a caller-owned device buffer of the context's
 * precision and shape for a field, or read back the library's own buffer. */
int monc_tvd_bind_field(monc_tvd_ctx *ctx, int field, int variant, void *device_ptr);
int monc_tvd_field_device_ptr(monc_tvd_ctx *ctx, int field, int variant, void **device_ptr);

/* Launch shape (no counterpart in the reference: a tuning knob of this
 * library).  kernel MONC_TVD_KERNEL_MARCH (default, columns of up to 256
 * levels): each face computed once, rows = cell rows per block (0: the
 * default for the column height; otherwise one of the shapes listed by
 * monc_tvd_march_shapes), strip = x columns per block (0: 32 on large
 * domains, narrower down to 1 until a launch has 256 blocks on small ones).
 * MONC_TVD_KERNEL_CELL: one thread per cell, every face recomputed (any
 * column height; the fallback beyond 256 levels or 4 GiB per field).
 * MONC_TVD_KERNEL_STAGE: the forward step (monc_tvd_forward_steps,
 * monc_tvd_forward_step_into, monc_tvd_forward_step_record), the drop-in call
 * (monc_tvd_tendencies) and the leapfrog step (monc_tvd_leapfrog_steps) of
 * each field kind with u, v, w (and the field the kind advects) staged in
 * LDS; rows = cell rows per block (4, 6, 8, 11 or 15; the scalars, and a
 * momentum field advecting its z-field, take 4 where 15 is set; 0: the
 * default), strip as above.  A drop-in call that records z-face diagnostics
 * marches; the other calls use the march kernel.  Without a tiling set (or
 * after kernel 0 with rows = strip = 0) these calls stage u, v, w and the
 * scalar pair in FP64, and u, v and w in FP32, where the domain gives at
 * least 512 staged blocks, and march the rest. */
#define MONC_TVD_KERNEL_MARCH 0
#define MONC_TVD_KERNEL_CELL 1
#define MONC_TVD_KERNEL_STAGE 2
int monc_tvd_set_tiling(monc_tvd_ctx *ctx, int kernel, int rows, int strip);
/* Tile-shape autotuner (no counterpart in the reference).  For each kernel
 * of the forward step (u, v, w, the scalar group) it times every compiled
 * march shape and strip length (8, 16, 32, 64) and every stage shape (rows
 * 6, 8, 11, 15 x strips 32, 64, 128) on the context's own fields, reps launches
 * each after one warm-up, writing into library scratch (the state is not
 * advanced), and keeps the fastest kernel and shape per field kind.  chosen
 * (12 ints, may be null) receives kernel, rows and strip per kind in the
 * order u, v, w, scalars.  Needs the prognostics of u, v, w and the advected
 * fields. */
int monc_tvd_autotune(monc_tvd_ctx *ctx, double dtm, int reps, int *chosen);

/* The kernel and tile shape the forward step uses per field kind, defaults
 * resolved (no counterpart in the reference): chosen (12 ints) receives
 * kernel, rows and strip for u, v, w and the scalars, as monc_tvd_autotune
 * reports them. */
int monc_tvd_get_tiling(monc_tvd_ctx *ctx, int *chosen);

/* The march shapes (cell rows per block) compiled for this context's column
 * height, default first: writes up to n of them to rows and returns how many
 * exist (0 when the march kernel cannot run this context), or a negative
 * status. */
int monc_tvd_march_shapes(monc_tvd_ctx *ctx, int *rows, int n);

/* Tendency diagnostics of the component (tvdadvection.F90:1064-1188, the
 * published fields tend_*_tvdadvection_3d_local and _profile_total_local of
 * :71-103) and the vertical face values handed to profile_diagnostics
 * (tvd_dgs_terms%adv_*_dgs, :615-644, 666-671, 712-717).  flags:
 *   MONC_TVD_DIAG_TENDENCY  every monc_tvd_tendencies call records, for each
 *       advected field, tend_3d = s_after - s_before over the interior
 *       columns (save_precomponent_tendencies / compute_component_tendencies)
 *       and its per-level total over the interior columns; with th advected
 *       and rprefrcp given (length nz, the reference's
 *       vertical%rprefrcp), also tabs = s_after*rprefrcp - s_before*rprefrcp;
 *   MONC_TVD_DIAG_ZFACE     every call records each advected field's flux_z
 *       per interior column: flux_z(k), the face between levels k-1 and k
 *       (0 at k = 1), in the field layout.
 * Tracers have neither (the reference records none for them).  flags = 0
 * turns recording off. */
#define MONC_TVD_DIAG_TENDENCY 1
#define MONC_TVD_DIAG_ZFACE 2
int monc_tvd_set_diagnostics(monc_tvd_ctx *ctx, int flags, const double *rprefrcp);

/* Read a recorded diagnostic into host memory (FP64):
 *   MONC_TVD_DIAG_3D          tend_3d of `field`, shape (nz, ny, nx), no halo;
 *   MONC_TVD_DIAG_PROFILE     its per-level total, length nz;
 *   MONC_TVD_DIAG_TABS_3D     tabs tendency (field ignored), (nz, ny, nx);
 *   MONC_TVD_DIAG_TABS_PROFILE its total, length nz;
 *   MONC_TVD_DIAG_ZFACES      flux_z of `field`, shape (nz, ny+4, nx+4).
 * Fails with MONC_TVD_ERR_STATE when that diagnostic was never recorded. */
#define MONC_TVD_DIAG_3D 0
#define MONC_TVD_DIAG_PROFILE 1
#define MONC_TVD_DIAG_TABS_3D 2
#define MONC_TVD_DIAG_TABS_PROFILE 3
#define MONC_TVD_DIAG_ZFACES 4
int monc_tvd_get_diagnostic(monc_tvd_ctx *ctx, int field, int which, double *host);

/* The same tendency diagnostics of the two sibling components on this
 * context: with MONC_TVD_DIAG_TENDENCY set, every monc_pw_tendencies call
 * records tend_3d = s_after - s_before and its per-level total for each field
 * it changes (u, v, w with advect_flow; th with advect_th and th active; each
 * q with advect_q), and the tabs tendency with rprefrcp given: the published
 * fields tend_*_pwadvection_3d_local / _profile_total_local
 * (pwadvection.F90:56-120, save_precomponent_tendencies and
 * compute_component_tendencies :619-743); every monc_thadv_tendencies call
 * records th's (and tabs') while th is active, also when no term is switched
 * on (a zero tendency): tend_th/tabs_thadvection_3d_local /
 * _profile_total_local (thadvection.F90:47-51, 84-88, 152-159, 228-270).
 * component selects whose record is read; monc_tvd_get_diagnostic reads the
 * TVD component's.  `which` as above; z-faces exist for TVD only. */
#define MONC_TVD_COMPONENT_TVD 0
#define MONC_TVD_COMPONENT_PW 1
#define MONC_TVD_COMPONENT_THADV 2
#define MONC_TVD_NCOMPONENTS 3
int monc_tvd_get_component_diagnostic(monc_tvd_ctx *ctx, int component, int field, int which, double *host);

/* Per-kernel timing (no counterpart in the reference: measurement support).
 * With profiling on, every launch is bracketed by HIP events on the context
 * stream; monc_tvd_kernel_times synchronises, folds the events into per-kind
 * totals and returns the first n of them: milliseconds and launch counts for
 * kinds 0 = u, 1 = v, 2 = w, 3 = scalars, 4 = whole-field kernels (halo
 * fill, th_advection), 5 = one-cell-per-thread kernel (all fields); kind 6
 * is reserved (always empty).
 * reset != 0 clears the totals afterwards. */
#define MONC_TVD_NKERNEL_KINDS 7
int monc_tvd_set_profiling(monc_tvd_ctx *ctx, int on);
int monc_tvd_kernel_times(monc_tvd_ctx *ctx, double *ms, long long *launches, int n, int reset);

/* Run subsequent work on this hipStream_t (NULL: the legacy default stream). */
int monc_tvd_set_stream(monc_tvd_ctx *ctx, void *hip_stream);
int monc_tvd_synchronize(monc_tvd_ctx *ctx);

/* Replaces haloswapper.F90:62-71 on a single rank: periodic fill of the depth-2
 * x/y halo and its corners (halocommunication.F90:486-498, 1058-1168) of every
 * set prognostic and z-field (haloswapper.F90:472-520). */
int monc_tvd_halo_fill(monc_tvd_ctx *ctx);

/*
 * Replaces tvdadvection.F90:528-599 timestep_callback, applied to every column
 * of the local domain (timestepper.F90:65-85): adds the TVD advection
 * tendencies to the source fields (su, sv, sw, sth, sq(:)), overwriting level
 * nz of su, sv and the scalars (:942-959).  The ultflx timestep is 2*dtm unless
 * the corresponding stepping flag is FORWARD (:608-609, :656-657, :704-705); the
 * advected field is the prognostic when field_stepping is FORWARD and its
 * z-field otherwise (:739-749).  Halos must be filled (monc_tvd_halo_fill or by
 * the caller).  Asynchronous on the context stream.
 */
int monc_tvd_tendencies(monc_tvd_ctx *ctx, double dtm, int field_stepping, int momentum_stepping,
                        int scalar_stepping);

/*
 * Device-resident forward-Euler stepping of every advected field, nsteps times:
 * per step the tendencies of all fields from the pre-step state with FORWARD
 * stepping (ultflx dt = dtm, advected field = prognostic), then
 * field <- field + dtm * s as stepfields.F90:597-627 does for FORWARD_STEPPING
 * with s(1) = 0 and nothing else in the source fields, the periodic halo of the
 * new state written by the same kernel.  The source fields are not touched.
 * Requires filled halos on entry.  Asynchronous on the context stream.
 */
int monc_tvd_forward_steps(monc_tvd_ctx *ctx, double dtm, int nsteps);

/*
 * One step of monc_tvd_forward_steps from the state in `in` into `out`:
 * arrays of one device pointer per field (u, v, w, th, q(:)), each a buffer
 * of the context's precision and field shape with filled halos; out must not
 * alias in.  Only the advected fields are written (with their periodic halo);
 * the context's own fields are not touched.  No counterpart in the reference:
 * it lets a caller keep every state of a rollout (the recompute of a
 * checkpointed adjoint) without copies.  Asynchronous.
 */
int monc_tvd_forward_step_into(monc_tvd_ctx *ctx, double dtm, void *const *in, void *const *out);

/*
 * monc_tvd_forward_step_into that also records, for the adjoint of the step,
 * the limiter arm every face took (no counterpart in the reference: the
 * branch masks of the differentiable form).  rec holds one device pointer per
 * field (null for a field the step does not advect); rec[f] is a buffer of
 * one byte per point of field f's layout (nz*(ny+4)*(nx+4)), receiving at
 * face owner (k, j, i) the arms of its right x, north y and top z faces in
 * bits 0-1, 2-3 and 4-5 (0: the face is fc, 1: fd, 2: the QUICKEST value,
 * 3: fu + (fc - fu)/S; ultimateflux.F90:556-614).  Runs the stage kernel
 * for the kinds monc_tvd_forward_step_into stages (same shapes, so the new
 * state is the same bit for bit), the march kernel in its default shape for
 * the others; columns of at most 256 levels.  Asynchronous.
 */
int monc_tvd_forward_step_record(monc_tvd_ctx *ctx, double dtm, void *const *in, void *const *out,
                                 unsigned char *const *rec);

/*
 * Binds the record buffer monc_tvd_adjoint_step reads the limiter arms of
 * `field` from (a buffer monc_tvd_forward_step_record wrote for the state the
 * adjoint step starts from); null unbinds.  When every advected field has one
 * the adjoint takes the recorded arms, otherwise it re-derives them from the
 * state.  No counterpart in the reference.
 */
int monc_tvd_bind_record(monc_tvd_ctx *ctx, int field, const unsigned char *rec);

/*
 * Device-resident MONC timesteps of an advection-only run, nsteps times: per
 * step the sources are cleared (clearsourceterms), the TVD tendencies added
 * with momentum and scalar stepping CENTRED (steppingdirection.F90:27-32:
 * ultflx takes dt = 2*dtm) and field stepping FORWARD on the first step when
 * first_forward != 0, CENTRED otherwise (:41-48), then stepfields
 * (stepfields.F90:314-351, 522-627: u and v shifted by ugal and vgal, th and
 * q time-smoothed with c1 = 1 - 2*tsmth, c2 = tsmth when centred) and
 * swap_smooth (swapsmooth.F90:44-57: the classic swap after a forward step,
 * the Robert filter after a centred one) on every field, with the periodic
 * halo of the new fields and z-fields written by the same kernels.  Replaces
 * the clearsourceterms / tvd_advection / stepfields / swap_smooth sequence of
 * one timestep (global_config:86-91) when no other component adds sources.
 * Needs the context created with all advect flags, the prognostics and, unless
 * first_forward, the z-fields (time level t-1) with filled halos.
 * Asynchronous on the context stream.
 */
int monc_tvd_leapfrog_steps(monc_tvd_ctx *ctx, double dtm, double tsmth, double ugal, double vgal, int nsteps,
                            int first_forward);

/* FP64 conservation sum of one field over the interior columns:
 * out = sum_k weights[k-1] * sum_{j,i} f(k,j,i), accumulated in FP64 on the
 * device whatever the context precision (weights: host, length nz; e.g.
 * rhon(k)*dz(k) for k >= 2 and 0 at k = 1 gives the scalar mass that
 * tvdadvection.F90:925-959 conserves).  No counterpart in the reference: the
 * reference has no global diagnostic of this path.  Synchronous. */
int monc_tvd_weighted_sum(monc_tvd_ctx *ctx, int field, int variant, const double *weights, double *out);

/* monc_tvd_weighted_sum without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * kernels and the same result, bit for bit; the weights, partial sums and
 * result live in a fixed scratch array in device memory, one per GPU, held
 * until the process ends, sized for nz <= 4096 (larger nz: MONC_TVD_ERR_ARG).
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (other host threads or streams) race.  Synchronous. */
int monc_tvd_weighted_sum_noalloc(monc_tvd_ctx *ctx, int field, int variant, const double *weights, double *out);

/*
 * Adjoint of one monc_tvd_forward_steps step (no counterpart in the
 * reference: the differentiable form of the path).  With the prognostics
 * (MONC_TVD_PROG) holding the state X the step starts from and the
 * MONC_TVD_ADJOINT fields holding the adjoint lambda' of the state after it
 * (interior values, zero halo), replaces them by lambda = lambda' +
 * dtm * J(X)^T lambda', J the Jacobian of the tendencies; halo contributions
 * are folded back onto their periodic sources.  The adjoint fields of u, v, w
 * are allocated (zero) if missing.  The limiter arms are those bound with
 * monc_tvd_bind_record (recorded by the forward step from X), or re-derived
 * from X with the forward step's arithmetic; flow signs and upwind picks are
 * re-derived from X.  Either way they are those the forward step took.
 * Asynchronous.
 */
int monc_tvd_adjoint_step(monc_tvd_ctx *ctx, double dtm);

/*
 * Adjoint of monc_tvd_halo_fill (no counterpart in the reference): adds the
 * halo entries of every MONC_TVD_ADJOINT field to their periodic sources and
 * zeroes the halo.  Applied to the adjoint of a state whose halo entries were
 * themselves used (a loss over the full arrays) before the first
 * monc_tvd_adjoint_step.  Asynchronous.
 */
int monc_tvd_adjoint_fold(monc_tvd_ctx *ctx);

/*
 * The sibling component th_advection on the same context: replaces
 * components/thadvection/src/thadvection.F90:124-211 (timestep_callback's
 * vertical_advection_of_reference_state and advection_of_mean_baroclinicity)
 * with the factors of its initialisation_callback (:60-77).  Adds to the th
 * source field (MONC_TVD_TH, MONC_TVD_SOURCE) at every column, halo columns
 * included (the component does not skip them), levels 2..nz:
 *   anelastic != 0:     sth -= tzc1(k)*2*dthref(k-1)*w(k-1) + tzc2(k)*2*dthref(k)*w(k)
 *   baroclinicity != 0: sth += thref(k)*(fcoriol/g)*((v+vgal)*rate_x - (u+ugal)*rate_y)  (anelastic)
 *                       sth += (v+vgal)*mfx - (u+ugal)*mfy, mf = rate*thref0*(fcoriol/g)  (otherwise)
 * from the prognostics u, v, w and the context's tzc1, tzc2.  dthref and
 * thref (length nz: vertical%dthref, vertical%thref) are read only when the
 * term needs them.  baroclinicity is the reference's
 * l_advect_mean_baroclinicity (baroclinicity_use_geostrophic_shear with a
 * passive q); rate_x, rate_y are rate_change_geostrophic_wind_x/_y; g is
 * science_constants G.  The tendency diagnostics of the component are not
 * produced.  Launches on the library stream.
 */
int monc_thadv_tendencies(monc_tvd_ctx *ctx, int anelastic, const double *dthref, int baroclinicity,
                          const double *thref, double thref0, double fcoriol, double g, double rate_x,
                          double rate_y, double ugal, double vgal);

/*
 * The sibling component pw_advection on the same context: replaces
 * components/pwadvection/src/pwadvection.F90:258-610 (timestep_callback with
 * advect_flow_fields, advect_th_field, advect_q_field; U_ACTIVE, V_ACTIVE,
 * W_ACTIVE, l_toplevel = .true.), the centred second-order scheme, over the
 * interior columns (halo columns are skipped, :307) and levels 2..nz.  As the
 * reference does, it overwrites the sources of u, v, w (advect_flow) and of
 * th (advect_th) and adds to those of the q fields (advect_q); level 1, the
 * halo columns and sw at k = nz are left untouched.  Reads the prognostics
 * u, v, w (and th, q when advected), the context's cx, cy and tzc1, tzc2,
 * tzd1, tzd2; tcx, tcy are the reference's horizontal%tcx = 0.25/dx and
 * tcy = 0.25/dy (gridmanager.F90:2638-2641).  The flags are the reference's
 * advect_flow/advect_th/advect_q (the field's advection option "pw" or
 * "any").  The tendency diagnostics of the component are not produced.
 * Launches on the library stream.
 */
int monc_pw_tendencies(monc_tvd_ctx *ctx, int advect_flow, int advect_th, int advect_q, double tcx, double tcy);

/*
 * Tangent-linear model of monc_tvd_forward_steps (FP64 contexts only; no
 * counterpart in the reference).  Advances the state (MONC_TVD_PROG) and its
 * tangent (MONC_TVD_TANGENT, periodic halo filled) together nsteps times:
 * dX' = dX + dtm * J(X) dX.  Asynchronous.
 */
int monc_tvd_tangent_steps(monc_tvd_ctx *ctx, double dtm, int nsteps);

/*
 * Batched ensembles (no counterpart in the reference, which runs one domain
 * per MPI job): n_members independent local domains of uneven sizes, each
 * advecting u, v, w, th and q (field ids 0..4) with forward-Euler steps of its
 * own dtm, as monc_tvd_forward_steps does for one context.  All members'
 * fields live in one device pool per variant (the state and the adjoint);
 * member m's field f starts at element offset(m) + f * field_stride(m) of the
 * pool, in the (nz, ny+4, nx+4) layout above.  Members need nz <= 256 and
 * every field under 4 GiB.  metrics holds, member after member, the eight
 * FP64 vectors rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2 of length nz(m)
 * each (the order of monc_tvd_create).
 */
typedef struct monc_tvd_ensemble monc_tvd_ensemble;
int monc_tvd_ensemble_create(int n_members, const int *nz, const int *ny, const int *nx, const double *cx,
                             const double *cy, const double *metrics, int precision, monc_tvd_ensemble **out);
int monc_tvd_ensemble_destroy(monc_tvd_ensemble *ens);
/* Pool layout: element offset and field stride of a member; member = -1
 * gives the pool's total element count in *offset. */
int monc_tvd_ensemble_layout(monc_tvd_ensemble *ens, int member, long long *offset, long long *field_stride);
/* Host staging of one member's field (FP64 host array (nz, ny+4, nx+4));
 * variant MONC_TVD_PROG (the current state) or MONC_TVD_ADJOINT. */
int monc_tvd_ensemble_set_field(monc_tvd_ensemble *ens, int member, int field, int variant, const double *host);
int monc_tvd_ensemble_get_field(monc_tvd_ensemble *ens, int member, int field, int variant, double *host);
/* Device copy of a whole pool (the current state or the adjoint) to dst, or
 * from src into it (exactly one non-null; a buffer of layout(-1) elements of
 * the ensemble's precision).  Asynchronous. */
int monc_tvd_ensemble_copy_pool(monc_tvd_ensemble *ens, int variant, void *dst, const void *src);
/* Per-member dtm (n_members values); synchronises the stream. */
int monc_tvd_ensemble_set_dt(monc_tvd_ensemble *ens, const double *dtm);
int monc_tvd_ensemble_set_stream(monc_tvd_ensemble *ens, void *hip_stream);
int monc_tvd_ensemble_synchronize(monc_tvd_ensemble *ens);
/* Periodic halo fill of every field of every member (monc_tvd_halo_fill). */
int monc_tvd_ensemble_halo_fill(monc_tvd_ensemble *ens);
/* nsteps forward-Euler steps of every member.  Each step is ONE persistent
 * launch over the march tiles of all members (u, v, w and the (th, q) pair),
 * largest members first; pairs of steps replay a captured hipGraph unless
 * graphs are switched off.  Asynchronous. */
int monc_tvd_ensemble_forward_steps(monc_tvd_ensemble *ens, int nsteps);
/* Adjoint of one step of every member (monc_tvd_adjoint_step), from the state
 * in `state` (a device buffer in pool layout; NULL: the current state), on
 * the adjoint pool. */
int monc_tvd_ensemble_adjoint_step(monc_tvd_ensemble *ens, const void *state);
/* Adjoint of the halo fill on the adjoint pool (monc_tvd_adjoint_fold). */
int monc_tvd_ensemble_adjoint_fold(monc_tvd_ensemble *ens);
/* Tile shape of the persistent step: waves per block (4, 8 or 12; at least
 * the 64-level sub-columns of the tallest member), x columns per tile, and
 * whether step pairs replay a captured graph. */
int monc_tvd_ensemble_set_tiling(monc_tvd_ensemble *ens, int waves, int strip, int use_graph);
/* Autotuner of the persistent step: times every compiled shape (waves 4, 8,
 * 12 x strip 8, 16, 32, 64) over reps steps written to library scratch (the
 * state is not advanced) and keeps the fastest; chosen (2 ints, may be null)
 * receives waves and strip. */
int monc_tvd_ensemble_autotune(monc_tvd_ensemble *ens, int reps, int *chosen);

/* Test hook (no counterpart in the reference): the library's next n device
 * allocations fail with hipErrorOutOfMemory without reaching the HIP runtime,
 * so callers can exercise their MONC_TVD_ERR_HIP paths; n <= 0 clears it.
 * Returns the count that was pending. */
int monc_tvd_test_fail_allocations(int n);

/* Message of the last failure on the calling thread ("" if none). */
const char *monc_tvd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MONC_TVD_H */
