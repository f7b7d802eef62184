/*
 * tc_internal.h -- the solver handle (private to the host library).
 */
#ifndef TC_INTERNAL_H
#define TC_INTERNAL_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>
#include "tcstep.h"
#include "../tc_shared.h"

enum { TC_FUX = 0, TC_FUY = 1, TC_FUZ = 2, TC_FT = 3, TC_FP = 4, TC_FPSI = 5 };

enum { TC_S_DT, TC_S_FACTOR, TC_S_PREDICT, TC_S_XSOLVE, TC_S_YSOLVE, TC_S_ZSOLVE,
  TC_S_DIVY, TC_S_ZFFT, TC_S_PXSOLVE, TC_S_YIFFT, TC_S_CORRECT, TC_S_DIAG, TC_NSTAGE };

#define TC_MAXALLOC 512

/* tunable launch shapes: {planes per block, j rows per block} of the cell kernels */
enum { TC_TUNE_PREDICT, TC_TUNE_CORRECT, TC_NTUNE };
#define TC_MAXEV 64

/* line-solve factor tables of one field (see tck_linesolve.hip) */
typedef struct {
  double * xr; double * xv; double * xl;
  double * yr; double * yv; double * yq1; double * ymeta;
  double * zr; double * zv; double * zq1; double * zmeta;
} tc_ls_t;

typedef struct {
  hipEvent_t ev[TC_MAXEV][2];
  int nrec;
  double total_ms;
  long long launches;
} tc_prof_t;

struct tc_solver_s {
  tc_params_t p;
  tc_geom_t g;
  tc_coef_t coef;
  tc_flags_t flags;
  hipStream_t stream;
  int own_stream;         /* 0: the stream belongs to an ensemble */
  void * h_stage;         /* pinned staging of field transfers (whole planes) */
  size_t h_stage_bytes;
  int tp;                 /* pitch of per-i tables: nx + 2 */
  int nky;
  double hz, nu, kappa;
  double dif_limit;       /* min of the explicit diffusive dt limits */
  /* device state */
  double * u[TC_NFIELD][2];
  int cur[TC_NFIELD];
  int mixed;              /* precision TC_MIXED: FP32 temporaries (src, increment, spectrum,
                           * psi), FP64 state and projection arithmetic */
  int refine;             /* TC_MIXED_REFINED: a second potential and velocity correction
                           * at the last substep */
  size_t esz;             /* bytes per temporary element: 8 or 4 */
  void * src[TC_NFIELD][2];
  void * dlt[TC_NFIELD];  /* the increment of the mixed step (FP64: the next field buffer) */
  float * d_psi32;        /* psi of the mixed step */
  double * d_p;
  double * d_psi;
  double * d_tab;
  void * d_spec;
  double * d_pr; double * d_pl; double * d_pu;
  double * d_evy; double * d_evz;
  void * d_twy; void * d_twyn; void * d_twz;   /* double2, or float2 in the mixed step */
  double * d_dt;
  unsigned long long * d_dtmin;
  double * d_part;
  double * d_diag;
  tc_ls_t ls[TC_NFIELD][3];   /* per field and RK substep */
  /* host copies */
  double * h_tab; double * h_xf; double * h_xc; double * h_pl; double * h_pu;
  double h_dt;
  double h_diag_snap[10];
  double diag_vals[10];
  int diag_valid, diag_pending, diag_armed;
  int dt_valid;
  /* running statistics: ux1 ux2 uy1 uy2 uz1 uz2 adv dif, then the tensor sum of the last
   * substep (allocated on first use) */
  double * d_stat[9];
  size_t stat_num;
  int stat_armed;
  /* bookkeeping */
  void * allocs[TC_MAXALLOC];
  int nalloc;
  size_t dev_bytes;
  int prof_on;
  tc_prof_t prof[TC_NSTAGE];
  int tune[TC_NTUNE][2];
  int predict_window;     /* 1: the predictor variant with the fields' planes staged in LDS */
  int fft_exact;          /* forms of the persistent transforms where they apply: 1 axial exact, 2 divergence twiddles in LDS, 4 inverse theta exact */
  int psolve_lean;        /* periodic line solves: 1 interior waves skip the row tests, 2 batched LDS folds */
  /* bumped by every host-side change that work recorded earlier (an ensemble's graphs)
   * cannot see: fields replaced, a closure coupled, uncoupled or destroyed, launch shapes,
   * dt parameters, diagnostics or profiling armed, and every launch outside a replay */
  unsigned long long gen;
  /* learned closure (tc_cnn.c): network, corrections in the field layout (float), inputs */
  tc_cnn_t * closure;
  void * d_corr[TC_NFIELD];
  double * d_cnn_xc;
  float cnn_mean[5], cnn_inv_std[5], cnn_scale[4];
};

int tc_fail(const char * fmt, ...);
/* return 1 with a message from the enclosing function when a HIP call or launcher fails */
#define HIPOK(expr) do { hipError_t e_ = (expr); if(e_ != hipSuccess) \
  return tc_fail("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); } while(0)
#define KOK(expr) do { int e_ = (expr); if(e_ != 0) \
  return tc_fail("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString((hipError_t)e_)); } while(0)
int tc_apply_bc(tc_solver_t * s, int field, double * a);
double * tc_field_ptr(tc_solver_t * s, int f);
int tc_enqueue_step(tc_solver_t * s);
/* substep r of a step (r = 0 also decides dt and the implicit factors) */
int tc_enqueue_substep(tc_solver_t * s, int r);
/* the host-side bookkeeping of one step (buffer parity, CFL state) without launching: for
 * replays of a recorded step */
void tc_advance_host_state(tc_solver_t * s);
/* make the solver enqueue on a stream it does not own */
int tc_share_stream(tc_solver_t * s, hipStream_t stream);
void tc_closure_release(tc_solver_t * s);
/* the attached network on the current fields -> corrections (tc_cnn.c) */
int tc_closure_eval(tc_solver_t * s);

int tc_finish_field(tc_solver_t * s, int f);

void tc_prof_begin(tc_solver_t * s, int stage);
void tc_prof_end(tc_solver_t * s, int stage);

#endif
