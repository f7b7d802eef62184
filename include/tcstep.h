/*
 * tcstep.h -- handle-based C-ABI of the MI355X RK3 step library (libtcstep.so).
 *
 * The reference keeps one solver per process in module statics that are initialised on the
 * first call (src/fluid/compute_potential.c:505-514, src/fluid/predict/ux.c:455-474,
 * src/decide_dt.c:21-23).  This library keeps the same state in an explicit handle, so a
 * process can hold several independent solvers (the ensemble driver does).  The
 * reference-shaped entry points (integrate.h, fluid_solver.h, decide_dt.h) are thin
 * wrappers over a default handle.
 *
 * Conventions (the reference's, src/*.c): every function returns 0 on success and non-zero
 * on failure; tc_last_error() then returns a message.  Fields cross this boundary in the
 * reference's NPY layout (src/array.c:82-222): ux is (nz, ny, nx+1) with the x walls at
 * i = 1 and nx+1, uy/uz/p/psi/t are (nz, ny, nx+2) with the walls at i = 0 and nx+1, no
 * y/z halos, C order, float64.  Device buffers stay resident between calls; nothing moves
 * between host and device except through tc_set_field / tc_get_field and the scalar dt.
 */
#ifndef TCSTEP_H
#define TCSTEP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_FP64 0
/* no reference counterpart: FP64 state (u, p), divergence, radial Poisson solve, projection
 * and diagnostics; FP32 right-hand sides, RK sources, implicit-solve increments, FFTs and
 * potential */
#define TC_MIXED 1
/* no reference counterpart: TC_MIXED plus one defect-correction pass at the end of every
 * step -- the divergence left by the FP32 potential is measured in FP64, solved for once more
 * and removed from the velocity -- so that the step leaves FP64-level divergence */
#define TC_MIXED_REFINED 2

typedef struct {
  int nx, ny, nz;           /* glsizes (domain.c:40) */
  int is_curved;            /* domain.c:37 */
  double lengths[3];        /* domain.c:43 */
  double Re, Pr;            /* fluid/init.c:118-119 */
  int implicit[3];          /* param/implicit.c:4-6, default {1, 0, 0} */
  int scalar;               /* 1: advance T (reference), 0: T is left untouched */
  double uy_xm, uy_xp;      /* param/boundary-condition.c:3-4, default 1, 0 */
  double t_xm, t_xp;        /* param/boundary-condition.c:6-7, default +0.5, -0.5 */
  double coef_dt_adv;       /* decide_dt.c:128 */
  double coef_dt_dif;       /* decide_dt.c:129 */
  double dt_fixed;          /* > 0: decide_dt returns this value (no reference counterpart) */
  int precision;            /* TC_FP64, TC_MIXED or TC_MIXED_REFINED */
  int device;               /* HIP device ordinal */
} tc_params_t;

typedef struct tc_solver_s tc_solver_t;
typedef struct tc_cnn_s tc_cnn_t;

/* defaults of the reference (src/param/*.c), Re = Pr = 1, coefficients 1 */
void tc_params_default(tc_params_t * p);

/* xf: nx+1 face positions XF(1..nx+1), xc: nx+2 centre positions XC(0..nx+1)
 * (domain.c:46-53).  Fields start at zero with boundary values applied. */
int tc_create(const tc_params_t * p, const double * xf, const double * xc, tc_solver_t ** out);
int tc_destroy(tc_solver_t * s);
const char * tc_last_error(void);

/* name: "ux" "uy" "uz" "p" "psi" "t"; data in the NPY layout above.  tc_set_field applies the
 * field's boundary condition and periodic halos (fluid/init.c:112-116). */
int tc_set_field(tc_solver_t * s, const char * name, const double * data);
int tc_get_field(tc_solver_t * s, const char * name, double * data);
/* the same in the reference's host array layout, with the y/z halo planes:
 * (nz+2, ny+2, nx+1) for ux, (nz+2, ny+2, nx+2) otherwise (src/array.c:10-46) */
int tc_set_field_ref(tc_solver_t * s, const char * name, const double * data);
int tc_get_field_ref(tc_solver_t * s, const char * name, double * data);
/* decide_dt coefficients (decide_dt.c:128-129) and a fixed step (> 0) */
int tc_set_dt_params(tc_solver_t * s, double coef_dt_adv, double coef_dt_dif, double dt_fixed);

/* one time step, integrate() (integrate.c:9-35); *dt receives the step size used */
int tc_step(tc_solver_t * s, double * dt);
/* nsteps steps queued on the solver's stream with no host synchronisation */
int tc_step_async(tc_solver_t * s, int nsteps);
int tc_synchronize(tc_solver_t * s);
int tc_get_dt(tc_solver_t * s, double * dt);

/* decide_dt (decide_dt.c:125-160) on the current field */
int tc_decide_dt(tc_solver_t * s, double * dt);
/* the four stages of one RK substep with a host-given dt (fluid_solver.h:34-60) */
int tc_predict(tc_solver_t * s, int rkstep, double dt);
int tc_potential(tc_solver_t * s, int rkstep, double dt);
int tc_correct(tc_solver_t * s, int rkstep, double dt);
int tc_pressure(tc_solver_t * s, int rkstep, double dt);

/* diagnostics in FP64 (src/logging/*.c).  Injection and dissipation use the velocity
 * gradients of the last predict call, as the reference's logging reads the tensors it left
 * (injection.c:49-51); before any step they use the current field. */
int tc_divergence(tc_solver_t * s, double * divmax);
int tc_injection(tc_solver_t * s, double out[2]);
int tc_dissipation(tc_solver_t * s, double out[2]);
int tc_total_energy(tc_solver_t * s, double out[4]);
/* on: every following step records the injection and dissipation of the field at the start
 * of its last substep (the reference's logging reads the tensors that substep left) */
int tc_arm_diagnostics(tc_solver_t * s, int on);

/* running statistics, replacing statistics.init/collect/output (src/statistics.c:56-307):
 * device accumulators in the reference's layouts, ux1 ux2 adv dif with (nx+1) x ny x nz and
 * uy1 uy2 uz1 uz2 with (nx+2) x ny x nz doubles (i fastest, no halos).  tc_stat_arm(s, 1)
 * allocates and zeroes them on first use and makes every following step keep the
 * LYX0+LYX1+LXY tensors of its last substep (statistics.c:221-234 reads those);
 * tc_stat_collect adds one sample of the current field; tc_stat_get copies one accumulator
 * ("ux1", "ux2", "uy1", "uy2", "uz1", "uz2", "adv", "dif") to host and/or returns the sample
 * count */
int tc_stat_arm(tc_solver_t * s, int on);
int tc_stat_collect(tc_solver_t * s);
int tc_stat_reset(tc_solver_t * s);
int tc_stat_get(tc_solver_t * s, const char * name, double * host, size_t * num);

/* no reference counterpart: launch shape of a cell kernel ("predict" or "correct"): kb planes
 * marched per block (1..nz) and ty j rows per block of 64 x ty threads (1, 2 or 4).  Kernel
 * "predict_window" with kb 1 or 0 switches the predictor's form that stages the fields'
 * planes in LDS (32 x 8 tiles, taken when ty is 4; the default) on or off; "fft_exact" with
 * kb a mask (1: the axial transforms, 4: the inverse theta transform: their exact-tiling
 * forms where the tiles cover the spectrum exactly; 2: the divergence + theta transform's
 * twiddle tables in LDS; 7, the default: all three) the forms of the persistent FP64
 * transforms; "psolve_lean" with kb a mask (1: the periodic y and z solves' waves of interior
 * full-length segments past the factors' convergence row skip the row tests; 2: their folds
 * read LDS four segments at a time; 3, the default: both) the forms of the periodic solves.
 * Results are identical for every shape and form; only the speed changes */
int tc_tune_set(tc_solver_t * s, const char * kernel, int kb, int ty);
/* no reference counterpart: times the predictor for every candidate shape (kb in 2..32,
 * ty in 1, 2, 4) on the solver's own buffers, reps launches each, and keeps the fastest for
 * the predictor and the correction.  Only the next-field and RK buffers are written, which
 * the next step overwrites, so the flow itself is unchanged. */
int tc_autotune(tc_solver_t * s, int reps, int * kb, int * ty, double * best_ms);

/* no reference counterpart: n independent flows (any sizes, parameters and precisions, one
 * device) advanced together, spread over a few streams; the first tc_ensemble_step takes one
 * eager step and records each RK substep of the whole batch into a hipGraph (for both
 * parities of the members' double buffers), every later step is three graph launches.
 * Members are ordinary solvers for set/get/diagnostics (tc_ensemble_member).  A step records
 * again by itself when a member changed in a way the graphs cannot see (fields replaced, a
 * closure coupled, uncoupled or destroyed, launch shapes or dt parameters set, a member
 * stepped on its own); tc_ensemble_invalidate drops the graphs explicitly. */
typedef struct tc_ensemble_s tc_ensemble_t;
int tc_ensemble_create(int n, const tc_params_t * params, const double * const * xf,
    const double * const * xc, tc_ensemble_t ** out);
tc_solver_t * tc_ensemble_member(tc_ensemble_t * e, int m);
/* tc_autotune on every member (members of the same shape share the result); kb, ty: optional
 * outputs of n entries */
int tc_ensemble_autotune(tc_ensemble_t * e, int reps, int * kb, int * ty);
int tc_ensemble_step(tc_ensemble_t * e, int nsteps);
int tc_ensemble_synchronize(tc_ensemble_t * e);
int tc_ensemble_invalidate(tc_ensemble_t * e);
int tc_ensemble_destroy(tc_ensemble_t * e);

/* no reference counterpart: the learned subgrid closure.  A 3-D CNN on a grid of
 * nx x ny x nz cells (3x3x3 convolutions 5 -> 32 -> 32 -> 4 channels, ReLU, BF16 operands on
 * the matrix cores with FP32 accumulation; x zero-padded, y and z periodic), trained on the
 * device: inputs x [batch][nz][ny][nx][5] (u, v, w, T, r at cell centres, normalised by the
 * caller), targets y [batch][nz][ny][nx][4] (corrections of the ux, uy, uz, T right-hand
 * sides), mean-square loss, Adam.  Layer l weights are [27][co][ci] with co, ci padded to
 * (32, 16), (32, 32), (16, 32); tap = dx+1 + 3 (dy+1) + 9 (dz+1). */
int tc_cnn_create(int nx, int ny, int nz, int batch, unsigned long long seed, int device, tc_cnn_t ** out);
int tc_cnn_destroy(tc_cnn_t * c);
int tc_cnn_set_layer(tc_cnn_t * c, int layer, const float * w, const float * b);
/* weights, biases and the gradients of the last tc_cnn_gradients / tc_cnn_train_step */
int tc_cnn_get_layer(tc_cnn_t * c, int layer, float * w, float * b, float * gw, float * gb);
int tc_cnn_predict(tc_cnn_t * c, const float * x, float * y);
int tc_cnn_gradients(tc_cnn_t * c, const float * x, const float * y, double * loss);
int tc_cnn_train_step(tc_cnn_t * c, const float * x, const float * y, float lr, double * loss);
/* couple a batch-1 network of the solver's grid (null: detach): at the start of every step
 * the inputs are formed from the current fields ((q - in_mean) / in_std), the network runs on
 * the solver's stream, and out_scale times its outputs are added to the explicit right-hand
 * sides (ux at faces as the mean of the two neighbouring centres) in every substep.  A
 * network is coupled to one solver at a time (its activations are that solver's scratch);
 * destroying either side uncouples them */
int tc_closure_attach(tc_solver_t * s, tc_cnn_t * c, const float in_mean[5], const float in_std[5],
    const float out_scale[4]);

/* events around the kernels of one named stage ("predict", "poisson", ...) while timing */
int tc_profile_enable(tc_solver_t * s, int on);
int tc_profile_read(tc_solver_t * s, const char * stage, double * total_ms, long long * launches);
size_t tc_device_bytes(const tc_solver_t * s);
/* measured copy bandwidth (GB/s, read + write bytes) over nbytes, iters copies */
int tc_probe_copy_bandwidth(tc_solver_t * s, size_t nbytes, int iters, double * gbps);

/* seeded initial condition (tests/icgen.py states the same definition in numpy):
 * tanh-clustered grid (stretch s, s = 0 uniform) and laminar Couette + uniform noise */
int tc_ic_grid(int nx, double lx, int is_curved, double stretch, double * xf, double * xc);
int tc_ic_fields(tc_solver_t * s, unsigned long long seed, double amp, double t_amp, int nmodes, double mode_amp);
/* no reference counterpart: tc_ic_fields for callers that cannot allocate device memory
 * (serving loops that reseed flows, harnesses that forbid hipMalloc).  The same fields, bit
 * for bit, with no allocation: tc_ic_fields stages its per-index tables in a device buffer it
 * allocates and frees on every call; this variant stages them in a device array the library
 * holds for the whole process (4 tables of TC_IC_NOALLOC_N = 4096 doubles, 128 KiB per
 * device) and its host work space on the stack (64 KiB).  Grids with nx + 2, ny or nz above
 * 4096 are refused with an error.  All calls on one device share that array: calls that
 * overlap on one device (threads, or solvers on different streams) race and must be
 * serialised by the caller; a call leaves the array holding its own tables. */
int tc_ic_fields_noalloc(tc_solver_t * s, unsigned long long seed, double amp, double t_amp, int nmodes, double mode_amp);

#ifdef __cplusplus
}
#endif
#endif
