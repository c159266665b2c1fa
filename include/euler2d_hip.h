/*
 * euler2d_hip.h - C-ABI of the MI355X (gfx950) implementation of the euler2d
 * hot path: the 2-D unsplit MUSCL-Hancock Godunov step (primitive conversion,
 * slopes, trace, approximate two-shock Riemann solver, conservative update),
 * the ghost-cell boundary fill and the CFL time-step max-reduction.
 *
 * The reference (pkestene/euler2d_cudaFortran) has no FFI: its host program
 * calls Fortran module procedures that share module-global arrays.  Each entry
 * point below replaces one of those procedures and cites it.  Every function
 * returns 0 on success and a negative E2D_ERR_* code on failure; the message
 * of the last failure on the calling thread is in e2d_last_error().  All
 * arguments are validated before any HIP call is made.
 *
 * State layout (host side, what upload/download exchange): the reference's
 * Fortran array u(isize, jsize, 4), column-major, isize = nx+4, jsize = ny+4
 * (2 ghost layers each side, serial_ref/m_run.f90:105-109), variable planes
 * in the order rho, E, rho*u, rho*v (serial_ref/m_constants.f90:11-14),
 * always float64 on the interface whatever the device precision.
 *
 * Threading: a context is not thread-safe; it owns one HIP stream.  Calls are
 * stream-ordered; only e2d_compute_dt, e2d_get_dt, e2d_download*, e2d_sums and
 * e2d_synchronize wait for the device.
 */
#ifndef EULER2D_HIP_H
#define EULER2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define E2D_VERSION 1

/* status codes */
#define E2D_OK 0
#define E2D_ERR_ARG (-1)      /* invalid argument (null pointer, bad index, bad parameter) */
#define E2D_ERR_HIP (-2)      /* a HIP runtime call failed */
#define E2D_ERR_NOMEM (-3)    /* device or host allocation failed */
#define E2D_ERR_STATE (-4)    /* call not valid in the context's current state */

/* implementation_version: cuf/m_HydroConstants.cuf:51-53 */
#define E2D_IMPL_FUSED 0        /* one fused kernel per step (LDS-free register marching) */
#define E2D_IMPL_MATERIALISED 1 /* q, slopes, face states and fluxes stored in HBM */
#define E2D_IMPL_FLUX_STORE 2   /* q and fluxes stored in HBM, slopes and face states in registers */

/* precision */
#define E2D_PREC_FP64 0        /* FP64 state and arithmetic (serial_ref built with fp_kind = dp_kind) */
#define E2D_PREC_FP32 1        /* FP32 state and arithmetic (serial_ref default kind) */
#define E2D_PREC_FP32_NATIVE 2 /* FP32 state; FP32 flux arithmetic with the native reciprocal and square
                                  root (quotients within 2 ulp, roots within 1 ulp); update added in
                                  FP64 and rounded to the FP32 state; FP64 dt and sums */
#define E2D_PREC_MIXED 3       /* FP64 state and FP64 update, FP32 slope/trace/Riemann arithmetic
                                  with the native reciprocal and square root (as fp32_native), FP64
                                  dt and sums */

/* riemann_solver: serial_ref/m_constants.f90:21-23 */
#define E2D_RIEMANN_APPROX 0
#define E2D_RIEMANN_HLLC 2

/* initial conditions for e2d_init_problem */
#define E2D_PROBLEM_IMPLODE 0 /* init_implode, serial_ref/m_run.f90:524-555 */
#define E2D_PROBLEM_DIAG 1    /* x+y<1: rho=1,p=1 else rho=0.125,p=0.1; u=v=0 */
#define E2D_PROBLEM_BUMPS 2   /* DIAG plus 16 seeded Gaussian bumps in rho and p */

/* stages kept by the materialised implementations, for e2d_download_stage
 * (version 1: all nine; version 2: Q, FX and FY) */
#define E2D_STAGE_Q 0
#define E2D_STAGE_FX 1
#define E2D_STAGE_FY 2
#define E2D_STAGE_DQX 3
#define E2D_STAGE_DQY 4
#define E2D_STAGE_QXM 5 /* XMAX face state (left state at the cell's right interface) */
#define E2D_STAGE_QXP 6 /* XMIN face state */
#define E2D_STAGE_QYM 7 /* YMAX face state */
#define E2D_STAGE_QYP 8 /* YMIN face state */

/* pointwise device numerics, for e2d_pointwise (input/output widths per item) */
#define E2D_FN_PRIM 1    /* u[4]                      -> q[4], c            */
#define E2D_FN_SLOPE 2   /* q, q+x, q-x, q+y, q-y [20] -> dqX[4], dqY[4]     */
#define E2D_FN_TRACE 3   /* q, dqX, dqY, dtdx, dtdy, face [15] -> qface[4]   */
#define E2D_FN_RIEMANN 4 /* qleft[4], qright[4]       -> qgdnv[4], flux[4]  */
#define E2D_FN_HLLC 5    /* qleft[4], qright[4]       -> qgdnv[4], flux[4]  */
#define E2D_FN_CMPFLX 6  /* qgdnv[4]                  -> flux[4]            */
#define E2D_FN_EOS 7     /* rho, eint                 -> p, c               */
#define E2D_FN_DIV 8     /* a, b                      -> a / b (the precision's flux arithmetic) */
#define E2D_FN_SQRT 9    /* x                         -> sqrt(x)            */
#define E2D_FN_DIV_ROOT 10 /* a, x                    -> a / sqrt(x), sqrt(x) for x > 0, formed as the
                              two-shock solver forms (p* - p) / w: the quotient may take its
                              reciprocal from the square root's own iteration */
#define E2D_FN_DIV_NEAR 11 /* a, b, r0                -> a / b with the reciprocal of b refined from the
                              start r0 (within about 2^-50 of 1/b), as the two-shock solver's
                              finish forms its quotients by wo^2 and by r* */
#define E2D_FN_CONV_GT 12 /* a, d                     -> 1 if a / d, rounded in the precision's arithmetic,
                              exceeds the Newton tolerance conv_tol, else 0: the two-shock solver's
                              stop test, a = |delp| >= 0, d = p* + smallpp > 0 */

/*
 * Run parameters: the namelist groups &RUN/&MESH/&HYDRO/&OTHER of the
 * reference (serial_ref/m_run.f90:71-74, cuf/m_HydroParameters.cuf:86-90)
 * minus the loop controls, which stay with the host.  Derived parameters
 * (dx, dy, smallp, smallpp, gamma6; serial_ref/m_run.f90:111-124) are
 * computed inside, in the run's precision.
 */
typedef struct e2d_params {
  int32_t nx, ny;
  int32_t boundary_type_xmin, boundary_type_xmax; /* 1 reflective, 2 absorbing, other periodic */
  int32_t boundary_type_ymin, boundary_type_ymax;
  int32_t slope_type;    /* 0 first order, 1 minmod, 2 monotonized central */
  int32_t niter_riemann; /* Newton iterations cap of the two-shock solver */
  int32_t iorder;        /* read and unused, as in the reference */
  int32_t implementation_version;
  int32_t precision;
  int32_t riemann_solver;
  double gamma0, cfl, smallr, smallc;
} e2d_params;

typedef struct e2d_ctx e2d_ctx;

/* Library version (E2D_VERSION) and whether a HIP device is usable (1/0). */
int e2d_version(void);
int e2d_device_available(void);

/* Message of the last failure on this thread ("" if none). */
const char *e2d_last_error(void);

/* A context lives on the device that is current when it is created; every
 * call on it makes that device current on the calling thread. */

/* Replaces initHydroParameters_gpu + the allocations of initHydroRun
 * (cuf/m_HydroParameters_gpu.cuf:67-103, cuf/m_HydroRun_gpu.cuf:83-120):
 * validates p, allocates the two state buffers (u = buffer 0, u2 = buffer 1)
 * and, for implementation versions 1 and 2, the stage arrays.  Buffer 1 starts at
 * zero, as the reference's never-initialised u2 (serial_ref/m_run.f90:155).
 * Besides the ranges of the enumerations, p needs gamma0 > 1, cfl > 0,
 * smallr > 0, smallc > 0 and floors that stay normal doubles:
 * gamma0 smallr^2 smallp (1 - gamma6) >= 1e-290 (the Riemann solver's
 * radicands) and smallr smallp >= 2^-949 (its Newton stop test), with
 * smallp = smallc^2 / gamma0 and gamma6 = (gamma0 + 1) / (2 gamma0). */
int e2d_create(const e2d_params *p, e2d_ctx **out);

/* Replaces cleanupHydroRun (cuf/m_HydroRun_gpu.cuf:123-140). */
int e2d_destroy(e2d_ctx *ctx);

/* Replaces the initial condition + upload of initHydroRun
 * (serial_ref/m_run.f90:160-167, cuf/m_HydroRun_gpu.cuf:105-118), computed on
 * the device in the run's precision; zeroes buffer 1, resets the running
 * invDt and the step counter.  seed is used by E2D_PROBLEM_BUMPS only. */
int e2d_init_problem(e2d_ctx *ctx, int problem, uint64_t seed);

/* d_u = h_u (cuf/m_HydroRun_gpu.cuf:117) and h_u = d_u / d_u2
 * (cuf/main_gpu.cuf:58-62) for buffer 0 (u) or 1 (u2). */
int e2d_upload(e2d_ctx *ctx, int buf, const double *h_u);
int e2d_download(e2d_ctx *ctx, int buf, double *h_u);

/* Replaces compute_dt (serial_ref/m_run.f90:248-312, cuf/m_HydroRun_gpu.cuf:198-300):
 * useU selects buffer 0 (useU == 0) or 1, the reduction window is
 * [3, size-3] (1-based), the running maximum of invDt over all calls is kept
 * in the context (the SAVE'd local of the reference) and dt = cfl / invDt is
 * returned.  Synchronous. */
int e2d_compute_dt(e2d_ctx *ctx, int useU, double *dt);

/* Replaces make_boundaries_gpu (cuf/m_HydroRun_gpu.cuf:1041-1132) /
 * make_boundaries (serial_ref/m_run.f90:561-648): fills the non-corner ghost
 * cells of buffer buf; corners are never written. */
int e2d_make_boundaries(e2d_ctx *ctx, int buf);

/* Replaces godunov_unsplit (serial_ref/m_run.f90:187-243,
 * cuf/m_HydroRun_gpu.cuf:306-452): one step with the host's dt, buffer 0 -> 1
 * for even nStep, 1 -> 0 for odd; boundaries of the input buffer are filled
 * first; nothing is copied from the input to the output buffer. */
int e2d_godunov_unsplit(e2d_ctx *ctx, int nStep, double dt);

/* The main-loop body of the reference (serial_ref/main.f90:66-74,
 * cuf/main_gpu.cuf:67-73) nsteps times without host round trips: per step
 * dt = cfl / invDt from the device-resident running maximum, the Godunov
 * step, and invDt updated from the new state.  *nStep is the step counter of
 * the host loop (selects the ping-pong direction) and is advanced by nsteps.
 * Equivalent, bit for bit, to calling e2d_compute_dt + e2d_godunov_unsplit
 * nsteps times. */
int e2d_run_steps(e2d_ctx *ctx, int nsteps, int *nStep);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the loop of e2d_run_steps (implementationVersion 0, any
 * precision and solver) on caller-owned state buffers, without a context.
 * d_u and d_u2 are buffers 0 and 1 in the device layout of e2d_device_buffer
 * (element (i,j,v) at [(v*jsize + j)*pitch + i], storage type of the
 * precision, pitch >= nx+4); nStep selects the input buffer (nStep % 2) as in
 * e2d_run_steps; *d_invdt (one device double) holds the running invDt on
 * entry and receives the new running maximum.  Everything is enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) and nothing waits.
 * Equivalent, bit for bit, to e2d_run_steps on a context whose buffers and
 * running invDt hold the same values.  It calls no allocating routine: its
 * only scratch (the two running-invDt slots, 8 KB) is a static device array
 * that each device keeps from the first call on it until the process ends,
 * so calls that overlap on one device race.  Grids up to the e2d_create
 * limit (pitch * (ny+4) * 4 < 2^31 elements); implementation versions 1 and
 * 2 are rejected (their stage arrays would need allocations). */
int e2d_run_steps_noalloc(const e2d_params *p, void *d_u, void *d_u2, int64_t pitch, int nsteps, int nStep,
                          double *d_invdt, void *stream);

/* dt that the next step will use (cfl / running invDt); synchronous. */
int e2d_get_dt(e2d_ctx *ctx, double *dt);

/* Running invDt (the only state the reference loop carries besides u, u2):
 * read/set for checkpoint and restart. */
int e2d_get_invdt(e2d_ctx *ctx, double *invdt);
int e2d_set_invdt(e2d_ctx *ctx, double invdt);

/* FP64 sums over the interior of buffer buf of (rho, E, rho*u, rho*v),
 * in a fixed order (deterministic). */
int e2d_sums(e2d_ctx *ctx, int buf, double out[4]);

/* Implementation versions 1 and 2 only: download a stage array of the last step
 * (layout as the state; q, dq and face arrays (isize,jsize,4), fx/fy as the
 * reference's fx/fy arrays, serial_ref/m_run.f90:17). */
int e2d_download_stage(e2d_ctx *ctx, int stage, double *h);

/* Wait for all work of the context. */
int e2d_synchronize(e2d_ctx *ctx);

/* GPU timers, replacing the CudaTimer event pairs of the reference
 * (cuf/m_Monitoring_gpu.cuf:14-129; d_godunov_timer, d_boundaries_timer,
 * cuf/m_HydroRun_gpu.cuf:365-451).  While enabled, every launch made by
 * e2d_godunov_unsplit / e2d_run_steps / e2d_compute_dt is bracketed by hipEvents
 * recorded on the context's stream.  e2d_get_timers synchronises, returns the
 * accumulated milliseconds and launch counts per category since the last call,
 * and resets them. */
#define E2D_TIMER_GODUNOV 0    /* the step kernel(s) */
#define E2D_TIMER_BOUNDARIES 1 /* ghost fill (a separate launch in version 1 only) */
#define E2D_TIMER_DT 2         /* standalone compute_dt reductions */
#define E2D_NUM_TIMERS 3
int e2d_enable_timers(e2d_ctx *ctx, int on);
int e2d_get_timers(e2d_ctx *ctx, double ms[E2D_NUM_TIMERS], int64_t count[E2D_NUM_TIMERS]);

/* Device pointer and row pitch (in elements) of a state buffer; element
 * (i, j, v) (0-based) sits at ptr[(v * jsize + j) * pitch + i], in the
 * precision's storage type (double for FP64, float for FP32/MIXED). */
int e2d_device_buffer(e2d_ctx *ctx, int buf, void **dptr, int64_t *pitch);

/*
 * Tangent-linear and adjoint of the step (no counterpart in the reference).
 * Two-shock solver only; the tangent for FP64, the adjoint for every
 * precision (its reverse passes run in the step's arithmetic type, FP32 for
 * the FP32 and mixed variants, and add each step's contribution to lam in
 * FP64).  The step is linearised as executed (floors, limiter branches,
 * sampling branches and Newton iteration counts of the forward pass,
 * |x|' = +1 at 0) with dt frozen.  Vector fields are FP64 device arrays in
 * the state's device layout (element (i,j,v) at [(v*jsize + j)*pitch + i],
 * see e2d_device_buffer) for every precision; only interior entries are read
 * and written.
 */

/* v_out = J v_in for the step whose input state is buffer buf (the state is
 * not modified apart from its ghost cells, which are filled). */
int e2d_tangent_step(e2d_ctx *ctx, int buf, double dt, const double *d_v_in, double *d_v_out);

/* lam_out = J^T lam_in for the same step; lam_out may alias lam_in. */
int e2d_adjoint_step(e2d_ctx *ctx, int buf, double dt, const double *d_lam_in, double *d_lam_out);

/* Gradient of an nsteps rollout from the state in buffer nStep0 % 2 with the
 * frozen dt sequence h_dts[0..nsteps-1]: lam_0 = (dU_n / dU_0)^T lam_n.
 * Binomial checkpointing (Griewank and Walther's revolve schedule: the least
 * number of forward steps for nsnap kept states) keeps at most nsnap state
 * snapshots in device memory; *n_forward returns the number of forward steps
 * recomputed.  On return the starting state is back in its
 * buffer. */
int e2d_adjoint_rollout(e2d_ctx *ctx, int nStep0, int nsteps, const double *h_dts, int nsnap,
                        const double *d_lam_n, double *d_lam_0, int64_t *n_forward);

/* Forward half of a checkpointed rollout (no counterpart in the reference,
 * which has no adjoint): advances the state in buffer nStep0 % 2 by nsteps
 * frozen-dt steps, as nsteps calls of e2d_godunov_unsplit would, and keeps in
 * the nsnap snapshots the starting state and the states that the reverse
 * sweep of e2d_adjoint_rollout stores on its first descent.  The next
 * e2d_adjoint_rollout call with the same nStep0, nsteps, dts and nsnap, made
 * with the same starting state back in buffer nStep0 % 2 (checked bit for
 * bit on the device), starts from them and recomputes that many fewer
 * forward steps; the gradient is the same bit for bit.  Any other adjoint
 * rollout call discards the recording. */
int e2d_rollout_forward(e2d_ctx *ctx, int nStep0, int nsteps, const double *h_dts, int nsnap);

/* FP64 device vector fields in the state's layout (zero-filled on allocation),
 * and their transfer from/to host arrays in the layout of e2d_upload. */
int e2d_field_alloc(e2d_ctx *ctx, double **d_field);
int e2d_field_free(e2d_ctx *ctx, double *d_field);
int e2d_field_upload(e2d_ctx *ctx, double *d_field, const double *h);
int e2d_field_download(e2d_ctx *ctx, const double *d_field, double *h);

/* Device-to-device copy of a whole state buffer (device layout, storage
 * type) from/to caller memory, e.g. torch tensors of shape (4, jsize, pitch). */
int e2d_copy_to_buffer(e2d_ctx *ctx, int buf, const void *d_src);
int e2d_copy_from_buffer(e2d_ctx *ctx, int buf, void *d_dst);

/* Pointwise device numerics over a batch, for known-answer tests of the
 * device routines against the reference's (serial_ref/m_utils.f90).
 * x holds n items of the input width of func, y receives n items of its
 * output width (see E2D_FN_*); float64 host arrays; synchronous.  p is
 * validated as by e2d_create. */
int e2d_pointwise(const e2d_params *p, int func, int n, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif /* EULER2D_HIP_H */
