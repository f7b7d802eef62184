/*
 * tc_solver.c -- host orchestration of the MI355X RK3 step (C11).
 *
 * One tc_solver_t owns the device-resident state of one flow (the reference's fluid_t and
 * the module statics of its solvers) and queues the kernels of integrate() (integrate.c:9-35)
 * on one HIP stream:
 *
 *   decide_dt      CFL minimum (fused into the previous step's last correction, or its own
 *                  reduction) + finalise; dt never leaves the device inside a step
 *   3 x substep    line-solve factor tables (from dt on the device)
 *                  predict (one fused sweep for ux uy uz T)
 *                  implicit line solves x -> y -> z, the last one adding the increment
 *                  Poisson: div + theta rFFT, z FFT, radial solves, z iFFT, theta iFFT -> psi
 *                  correction + pressure update (+ CFL minimum of the next step)
 *
 * Velocity and scalar fields are double-buffered (the predictor reads neighbours of the old
 * field while writing the new one); the RK source of substep s lives in buffer s % 2 and is
 * read back at substep s+1, which replaces the reference's pointer swap (predict/main.c:8-25)
 * by a fixed assignment.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#include <math.h>
#include <float.h>
#include <hip/hip_runtime_api.h>
#include "tcstep.h"
#include "tc_kernels.h"
#include "tc_internal.h"

static __thread char g_err[512];

const char * tc_last_error(void){
  return g_err;
}

int tc_fail(const char * fmt, ...){
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return 1;
}


void tc_params_default(tc_params_t * p){
  memset(p, 0, sizeof(*p));
  p->is_curved = 1;
  p->lengths[0] = p->lengths[1] = p->lengths[2] = 1.;
  p->Re = 1.;
  p->Pr = 1.;
  p->implicit[0] = 1;
  p->scalar = 1;
  p->uy_xm = 1.;
  p->uy_xp = 0.;
  p->t_xm = + 0.5;
  p->t_xp = - 0.5;
  p->coef_dt_adv = 1.;
  p->coef_dt_dif = 1.;
  p->dt_fixed = 0.;
  p->precision = TC_FP64;
  p->device = 0;
}

/* ---------------------------------------------------------------------------------------
 * device memory
 * ------------------------------------------------------------------------------------- */
static int dalloc(tc_solver_t * s, double ** p, size_t n){
  if(s->nalloc >= TC_MAXALLOC) return tc_fail("too many device allocations");
  const size_t bytes = sizeof(double) * (n ? n : 1);
  HIPOK(hipMalloc((void **)p, bytes));
  HIPOK(hipMemsetAsync(*p, 0, bytes, s->stream));
  s->allocs[s->nalloc++] = *p;
  s->dev_bytes += bytes;
  return 0;
}

static int dalloc_bytes(tc_solver_t * s, void ** p, size_t bytes){
  double * d = NULL;
  if(dalloc(s, &d, (bytes + sizeof(double) - 1) / sizeof(double))) return 1;
  *p = d;
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * metrics and per-i coefficient tables (domain.c:259-357, predict/*.c init_laplacians,
 * compute_potential.c:173-211), evaluated with the operand order of the reference
 * ------------------------------------------------------------------------------------- */

static int build_tables(tc_solver_t * s, const double * xf_in, const double * xc_in){
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
  const int tp = nx + 2;
  s->tp = tp;
  double * xf = calloc((size_t)nx + 2, sizeof(double));   /* XF(i) at [i] */
  double * xc = calloc((size_t)nx + 2, sizeof(double));
  double * tab = calloc((size_t)T_NTAB * tp, sizeof(double));
  if(!xf || !xc || !tab) return tc_fail("host allocation failed");
  for(int i = 1; i <= nx + 1; i++) xf[i] = xf_in[i - 1];
  for(int i = 0; i <= nx + 1; i++) xc[i] = xc_in[i];
  #define TB(t, i) tab[(size_t)(t) * tp + (i)]
  const double ly = s->p.lengths[1];
  const double hz = s->p.lengths[2] / nz;
  s->hz = hz;
  for(int i = 1; i <= nx + 1; i++) TB(T_HXXF, i) = xc[i] - xc[i - 1];
  TB(T_HXXC, 0) = 0.5 * xf[2] - 0.5 * xf[1];
  for(int i = 1; i <= nx; i++) TB(T_HXXC, i) = xf[i + 1] - xf[i];
  TB(T_HXXC, nx + 1) = 0.5 * xf[nx + 1] - 0.5 * xf[nx];
  for(int i = 1; i <= nx + 1; i++) TB(T_HYXF, i) = s->p.is_curved ? xf[i] * ly / ny : ly / ny;
  for(int i = 0; i <= nx + 1; i++) TB(T_HYXC, i) = s->p.is_curved ? xc[i] * ly / ny : ly / ny;
  for(int i = 1; i <= nx + 1; i++) TB(T_JDXF, i) = TB(T_HXXF, i) * TB(T_HYXF, i) * hz;
  for(int i = 0; i <= nx + 1; i++) TB(T_JDXC, i) = TB(T_HXXC, i) * TB(T_HYXC, i) * hz;
  const double nu = 1. / s->p.Re;
  const double kappa = 1. / s->p.Re / s->p.Pr;
  s->nu = nu;
  s->kappa = kappa;
  #define HXXF(i) TB(T_HXXF, i)
  #define HXXC(i) TB(T_HXXC, i)
  #define HYXF(i) TB(T_HYXF, i)
  #define HYXC(i) TB(T_HYXC, i)
  #define JDXF(i) TB(T_JDXF, i)
  #define JDXC(i) TB(T_JDXC, i)
  for(int i = 1; i <= nx + 1; i++){
    TB(T_AXF, i) = 0.5 * JDXF(i) / HXXF(i);
    TB(T_IJDXF, i) = 1. / JDXF(i);
    TB(T_IHXF, i) = 1. / HXXF(i);
    TB(T_IHYF, i) = 1. / HYXF(i);
    TB(T_JHXF, i) = JDXF(i) / HXXF(i);
    TB(T_JHYF, i) = JDXF(i) / HYXF(i);
    TB(T_JHZF, i) = JDXF(i) / hz;
    /* lyx.c:72-79: - 1. / jd_x0 * djdhx, djdhx = - JDXC(i-1)/HXXC(i-1) + JDXC(i)/HXXC(i) */
    const double jdhx_xm = JDXC(i - 1) / HXXC(i - 1);
    const double jdhx_xp = JDXC(i) / HXXC(i);
    const double djdhx = - jdhx_xm + jdhx_xp;
    TB(T_L1F, i) = - 1. / JDXF(i) * djdhx;
    TB(T_NUJF, i) = nu / JDXF(i);
    TB(T_N2JF, i) = 2. * nu / JDXF(i);
  }
  for(int i = 2; i <= nx; i++){
    /* ux.c:206-215 and :358-365 */
    const double djdhx_xm = - JDXF(i - 1) / HXXF(i - 1) + JDXF(i) / HXXF(i);
    const double djdhx_xp = - JDXF(i) / HXXF(i) + JDXF(i + 1) / HXXF(i + 1);
    TB(T_D1M, i) = 0.5 * djdhx_xm;
    TB(T_D1P, i) = 0.5 * djdhx_xp;
    /* ux.c:45-51, :59-65 */
    const double l = 1. / JDXF(i) * JDXC(i - 1) / HXXC(i - 1) / HXXC(i - 1);
    const double u = 1. / JDXF(i) * JDXC(i) / HXXC(i) / HXXC(i);
    TB(T_LXU_L, i) = l;
    TB(T_LXU_C, i) = - l - u;
    TB(T_LXU_U, i) = u;
    const double ly_ = 1. / HYXF(i) / HYXF(i);
    const double uy_ = 1. / HYXF(i) / HYXF(i);
    TB(T_LYU_L, i) = ly_;
    TB(T_LYU_C, i) = - ly_ - uy_;
    TB(T_LYU_U, i) = uy_;
  }
  for(int i = 0; i <= nx + 1; i++){
    TB(T_AYC, i) = 0.5 * JDXC(i) / HYXC(i);
    TB(T_AZC, i) = 0.5 * JDXC(i) / hz;
    TB(T_IJDXC, i) = 1. / JDXC(i);
    TB(T_IHYC, i) = 1. / HYXC(i);
    TB(T_JHYC, i) = JDXC(i) / HYXC(i);
    TB(T_JHZC, i) = JDXC(i) / hz;
    TB(T_NUJC, i) = nu / JDXC(i);
    TB(T_N2JC, i) = 2. * nu / JDXC(i);
  }
  for(int i = 1; i <= nx; i++){
    /* lyy.c:72-79 and uy.c:205-208 */
    const double jdhx_xm = JDXF(i) / HXXF(i);
    const double jdhx_xp = JDXF(i + 1) / HXXF(i + 1);
    const double djdhx = - jdhx_xm + jdhx_xp;
    TB(T_L1C, i) = 1. / JDXC(i) * djdhx;
    const double djdhx2 = - JDXF(i) / HXXF(i) + JDXF(i + 1) / HXXF(i + 1);
    TB(T_HDC, i) = 0.5 * djdhx2;
    /* uy.c:384-385 */
    const double djdhx_xm = - JDXC(i - 1) / HXXC(i - 1) + JDXC(i) / HXXC(i);
    const double djdhx_xp = - JDXC(i) / HXXC(i) + JDXC(i + 1) / HXXC(i + 1);
    TB(T_DAM, i) = 0.5 * djdhx_xm;
    TB(T_DAP, i) = 0.5 * djdhx_xp;
    /* uy.c:46-52, :58-64 */
    const double l = 1. / JDXC(i) * JDXF(i) / HXXF(i) / HXXF(i);
    const double u = 1. / JDXC(i) * JDXF(i + 1) / HXXF(i + 1) / HXXF(i + 1);
    TB(T_LXC_L, i) = l;
    TB(T_LXC_C, i) = - l - u;
    TB(T_LXC_U, i) = u;
    const double ly_ = 1. / HYXC(i) / HYXC(i);
    const double uy_ = 1. / HYXC(i) / HYXC(i);
    TB(T_LYC_L, i) = ly_;
    TB(T_LYC_C, i) = - ly_ - uy_;
    TB(T_LYC_U, i) = uy_;
    /* update_pressure.c:59-60, :91-92, :122-123 */
    TB(T_PLX_L, i) = JDXF(i) / HXXF(i) / HXXF(i);
    TB(T_PLX_U, i) = JDXF(i + 1) / HXXF(i + 1) / HXXF(i + 1);
    TB(T_PLY, i) = JDXC(i) / HYXC(i) / HYXC(i);
    TB(T_PLZ, i) = JDXC(i) / hz / hz;
  }
  /* decide_dt_dif (decide_dt.c:89-115) for the explicit directions */
  {
    const double diffusivity = fmax(nu, kappa);
    double grid[3];
    grid[0] = DBL_MAX;
    for(int i = 0; i <= nx + 1; i++) grid[0] = fmin(grid[0], HXXC(i));
    grid[1] = HYXC(1);
    grid[2] = hz;
    double lim = INFINITY;
    for(int d = 0; d < 3; d++){
      const double dif = s->p.coef_dt_dif / diffusivity * 0.5 / 3 * pow(grid[d], 2.);
      if(!s->p.implicit[d]) lim = fmin(lim, dif);
    }
    s->dif_limit = lim;
  }
  /* Poisson tables (compute_potential.c:205-209, :311-349) */
  s->h_pl = calloc((size_t)nx, sizeof(double));
  s->h_pu = calloc((size_t)nx, sizeof(double));
  for(int i = 1; i <= nx; i++){
    s->h_pl[i - 1] = 1. / JDXC(i) * JDXF(i) / HXXF(i) / HXXF(i);
    s->h_pu[i - 1] = 1. / JDXC(i) * JDXF(i + 1) / HXXF(i + 1) / HXXF(i + 1);
  }
  #undef HXXF
  #undef HXXC
  #undef HYXF
  #undef HYXC
  #undef JDXF
  #undef JDXC
  /* device copy */
  if(dalloc(s, &s->d_tab, (size_t)T_NTAB * tp)) return 1;
  HIPOK(hipMemcpyAsync(s->d_tab, tab, sizeof(double) * T_NTAB * tp, hipMemcpyHostToDevice, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  s->h_tab = tab;
  s->h_xf = xf;
  s->h_xc = xc;
  #define DT(t) (s->d_tab + (size_t)(t) * tp)
  tc_coef_t * c = &s->coef;
  c->hxxf = DT(T_HXXF); c->hxxc = DT(T_HXXC); c->hyxf = DT(T_HYXF); c->hyxc = DT(T_HYXC);
  c->jdxf = DT(T_JDXF); c->jdxc = DT(T_JDXC);
  c->axf = DT(T_AXF); c->ijdxf = DT(T_IJDXF); c->ihxf = DT(T_IHXF); c->ihyf = DT(T_IHYF);
  c->jhxf = DT(T_JHXF); c->jhyf = DT(T_JHYF); c->jhzf = DT(T_JHZF); c->l1f = DT(T_L1F);
  c->nujf = DT(T_NUJF); c->n2jf = DT(T_N2JF); c->d1m = DT(T_D1M); c->d1p = DT(T_D1P);
  c->ayc = DT(T_AYC); c->azc = DT(T_AZC); c->ijdxc = DT(T_IJDXC); c->ihyc = DT(T_IHYC);
  c->jhyc = DT(T_JHYC); c->jhzc = DT(T_JHZC); c->l1c = DT(T_L1C); c->hdc = DT(T_HDC);
  c->nujc = DT(T_NUJC); c->n2jc = DT(T_N2JC); c->dam = DT(T_DAM); c->dap = DT(T_DAP);
  c->lxu_l = DT(T_LXU_L); c->lxu_c = DT(T_LXU_C); c->lxu_u = DT(T_LXU_U);
  c->lyu_l = DT(T_LYU_L); c->lyu_c = DT(T_LYU_C); c->lyu_u = DT(T_LYU_U);
  c->lxc_l = DT(T_LXC_L); c->lxc_c = DT(T_LXC_C); c->lxc_u = DT(T_LXC_U);
  c->lyc_l = DT(T_LYC_L); c->lyc_c = DT(T_LYC_C); c->lyc_u = DT(T_LYC_U);
  c->plx_l = DT(T_PLX_L); c->plx_u = DT(T_PLX_U); c->ply = DT(T_PLY); c->plz = DT(T_PLZ);
  c->tab = s->d_tab;
  c->tp = tp;
  #undef DT
  /* predict/ux.c:70-78 */
  c->lapz_l = 1. / hz / hz;
  c->lapz_u = 1. / hz / hz;
  c->lapz_c = - c->lapz_l - c->lapz_u;
  c->hz = hz;
  c->ihz = 1. / hz;
  c->nu = nu;
  c->kappa = kappa;
  #undef TB
  return 0;
}

/* FFT twiddles exp(-2 pi i t / n), t = 0..cnt-1, evaluated in FP64 (rounded to FP32 for the
 * mixed step) */
static int upload_twiddles(tc_solver_t * s, int n, int cnt, void ** dptr){
  double * h = malloc(sizeof(double) * 2 * (size_t)cnt);
  float * hf = malloc(sizeof(float) * 2 * (size_t)cnt);
  if(!h || !hf){ free(h); free(hf); return tc_fail("host allocation failed"); }
  const double two_pi = 6.283185307179586476925286766559;
  for(int t = 0; t < cnt; t++){
    const double ang = - two_pi * (double)t / (double)n;
    h[2 * t] = cos(ang);
    h[2 * t + 1] = sin(ang);
    hf[2 * t] = (float)h[2 * t];
    hf[2 * t + 1] = (float)h[2 * t + 1];
  }
  const size_t bytes = 2 * s->esz * (size_t)cnt;
  if(dalloc_bytes(s, dptr, bytes)){ free(h); free(hf); return 1; }
  hipError_t e = hipMemcpyAsync(*dptr, s->mixed ? (void *)hf : (void *)h, bytes, hipMemcpyHostToDevice, s->stream);
  if(e == hipSuccess) e = hipStreamSynchronize(s->stream);
  free(h);
  free(hf);
  if(e != hipSuccess) return tc_fail("twiddle upload: %s", hipGetErrorString(e));
  return 0;
}

static int setup_poisson(tc_solver_t * s){
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
  const int nky = ny / 2 + 1;
  s->nky = nky;
  const double g_pi = 3.14159265358979324;
  double * evy = malloc(sizeof(double) * nky);
  double * evz = malloc(sizeof(double) * nz);
  if(!evy || !evz) return tc_fail("host allocation failed");
  for(int n = 0; n < nky; n++) evy[n] = - 4. * pow(sin(g_pi * n / ny), 2.);
  for(int n = 0; n < nz; n++) evz[n] = - 4. * pow(sin(g_pi * n / nz), 2.);
  if(dalloc(s, &s->d_evy, nky) || dalloc(s, &s->d_evz, nz)) return 1;
  if(dalloc(s, &s->d_pl, nx) || dalloc(s, &s->d_pu, nx)) return 1;
  HIPOK(hipMemcpyAsync(s->d_evy, evy, sizeof(double) * nky, hipMemcpyHostToDevice, s->stream));
  HIPOK(hipMemcpyAsync(s->d_evz, evz, sizeof(double) * nz, hipMemcpyHostToDevice, s->stream));
  HIPOK(hipMemcpyAsync(s->d_pl, s->h_pl, sizeof(double) * nx, hipMemcpyHostToDevice, s->stream));
  HIPOK(hipMemcpyAsync(s->d_pu, s->h_pu, sizeof(double) * nx, hipMemcpyHostToDevice, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  free(evy);
  free(evz);
  const size_t nspec = (size_t)nx * nky * nz;
  if(dalloc_bytes(s, &s->d_spec, 2 * s->esz * nspec)) return 1;
  if(dalloc(s, &s->d_pr, nspec)) return 1;
  KOK(tck_launch_pfactor(nx, nky, nz, s->d_pl, s->d_pu, s->coef.hyxc, s->d_evy, s->d_evz,
      s->hz, s->d_pr, s->mixed, s->stream));
  if(upload_twiddles(s, ny / 2, ny / 2, &s->d_twy)) return 1;
  if(upload_twiddles(s, ny, ny / 2 + 1, &s->d_twyn)) return 1;
  if(upload_twiddles(s, nz, nz, &s->d_twz)) return 1;
  HIPOK(hipStreamSynchronize(s->stream));
  return 0;
}

/* per-field, per-substep line-solve tables */
static int setup_linesolve(tc_solver_t * s){
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
  const int tp = s->tp;
  for(int f = 0; f < TC_NFIELD; f++){
    for(int r = 0; r < 3; r++){
      tc_ls_t * L = &s->ls[f][r];
      if(s->p.implicit[0]){
        if(dalloc(s, &L->xr, nx + 1) || dalloc(s, &L->xv, nx + 1) || dalloc(s, &L->xl, nx + 1)) return 1;
      }
      if(s->p.implicit[1]){
        if(dalloc(s, &L->yr, (size_t)ny * tp) || dalloc(s, &L->yv, (size_t)ny * tp) ||
           dalloc(s, &L->yq1, (size_t)ny * tp) || dalloc(s, &L->ymeta, (size_t)7 * tp)) return 1;
      }
      if(s->p.implicit[2]){
        if(dalloc(s, &L->zr, nz) || dalloc(s, &L->zv, nz) || dalloc(s, &L->zq1, nz) ||
           dalloc(s, &L->zmeta, 7)) return 1;
      }
    }
  }
  return 0;
}

/* launch shapes of the cell kernels (planes per block kb, rows per block ty): 32 x 4 for both
 * the predictor and the correction, shortened until the grid has at least four blocks
 * per CU, so that small grids fill the device (tc_autotune refines them).  The predictor
 * stages the fields' planes in LDS whenever its tiles have 4 rows. */
static void default_shapes(tc_solver_t * s){
  int ncu = 0;
  if(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->p.device) != hipSuccess || ncu <= 0) ncu = 256;
  const long long want = 4LL * ncu;
  const long long nbx = (s->p.nx + 63) / 64;
  const int start[2][2] = {{32, 4}, {32, 4}};
  const int which[2] = {TC_TUNE_PREDICT, TC_TUNE_CORRECT};
  for(int q = 0; q < 2; q++){
    int kb = start[q][0], ty = start[q][1];
    for(;;){
      const long long nb = nbx * ((s->p.ny + ty - 1) / ty) * ((s->p.nz + kb - 1) / kb);
      if(nb >= want) break;
      if(kb > 1) kb /= 2;
      else if(ty > 1) ty /= 2;
      else break;
    }
    s->tune[which[q]][0] = kb;
    s->tune[which[q]][1] = ty;
  }
  s->predict_window = 1;
  s->fft_exact = 7;
  s->psolve_lean = 3;
}

int tc_create(const tc_params_t * p, const double * xf, const double * xc, tc_solver_t ** out){
  *out = NULL;
  if(p->nx < 3 || p->ny < 4 || p->nz < 3) return tc_fail("grid too small: %d x %d x %d", p->nx, p->ny, p->nz);
  if(p->ny % 2 != 0) return tc_fail("ny must be even (real theta transform)");
  if(!tck_fft_supported(p->ny / 2) || !tck_fft_supported(p->nz))
    return tc_fail("ny/2 = %d and nz = %d must be products of 2, 3, 5 and 7", p->ny / 2, p->nz);
  if(p->nx > 1024) return tc_fail("nx = %d exceeds 1024 (radial line solver)", p->nx);
  if(p->ny / 2 > 4096 || p->nz > 4096) return tc_fail("transform length exceeds 4096");
  if(p->precision < TC_FP64 || p->precision > TC_MIXED_REFINED) return tc_fail("precision %d not supported", p->precision);
  tc_solver_t * s = calloc(1, sizeof(tc_solver_t));
  if(!s) return tc_fail("host allocation failed");
  s->p = *p;
  s->mixed = p->precision == TC_MIXED || p->precision == TC_MIXED_REFINED;
  s->refine = p->precision == TC_MIXED_REFINED;
  s->esz = s->mixed ? sizeof(float) : sizeof(double);
  if(hipSetDevice(p->device) != hipSuccess){
    free(s);
    return tc_fail("hipSetDevice(%d) failed", p->device);
  }
  default_shapes(s);
  if(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess){
    free(s);
    return tc_fail("hipStreamCreate failed");
  }
  s->own_stream = 1;
  tc_geom_t * g = &s->g;
  g->nx = p->nx; g->ny = p->ny; g->nz = p->nz;
  g->px = ((p->nx + 2 + TC_XOFF) + 15) / 16 * 16;
  g->py = p->ny + 2;
  g->pz = p->nz + 2;
  g->plane = (long long)g->px * g->py;
  g->n3 = g->plane * g->pz;
  s->flags.implicit[0] = p->implicit[0];
  s->flags.implicit[1] = p->implicit[1];
  s->flags.implicit[2] = p->implicit[2];
  s->flags.scalar = p->scalar;
  s->flags.uy_xm = p->uy_xm; s->flags.uy_xp = p->uy_xp;
  s->flags.t_xm = p->t_xm; s->flags.t_xp = p->t_xp;
  int rc = build_tables(s, xf, xc);
  const size_t n3 = (size_t)g->n3;
  for(int f = 0; f < TC_NFIELD && !rc; f++){
    rc |= dalloc(s, &s->u[f][0], n3);
    rc |= dalloc(s, &s->u[f][1], n3);
    rc |= dalloc_bytes(s, &s->src[f][0], s->esz * n3);
    rc |= dalloc_bytes(s, &s->src[f][1], s->esz * n3);
    if(s->mixed) rc |= dalloc_bytes(s, &s->dlt[f], s->esz * n3);
  }
  if(!rc) rc |= dalloc(s, &s->d_p, n3);
  if(!rc && !s->mixed) rc |= dalloc(s, &s->d_psi, n3);
  if(!rc && s->mixed) rc |= dalloc_bytes(s, (void **)&s->d_psi32, sizeof(float) * n3);
  if(!rc) rc |= dalloc(s, &s->d_dt, 4);
  if(!rc) rc |= dalloc(s, (double **)&s->d_dtmin, (size_t)TC_DTSLOTS * TC_DTSTRIDE);
  if(!rc){
    const long long nb = tck_diag_cells_blocks(g) > tck_diag_walls_blocks(g) ? tck_diag_cells_blocks(g) : tck_diag_walls_blocks(g);
    rc |= dalloc(s, &s->d_part, (size_t)nb * 8);
    rc |= dalloc(s, &s->d_diag, 16);
  }
  if(!rc) rc = setup_poisson(s);
  if(!rc) rc = setup_linesolve(s);
  if(!rc){
    /* boundary values on every buffer (fields start at zero) */
    for(int f = 0; f < TC_NFIELD && !rc; f++){
      for(int b = 0; b < 2 && !rc; b++) rc = tc_apply_bc(s, f, s->u[f][b]);
    }
    if(!rc) rc = tc_apply_bc(s, TC_FP, s->d_p);
    if(!rc && !s->mixed) rc = tc_apply_bc(s, TC_FPSI, s->d_psi);
  }
  if(!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = tc_fail("initialisation failed");
  if(rc){
    tc_destroy(s);
    return 1;
  }
  s->dt_valid = 0;
  s->diag_valid = 0;
  *out = s;
  return 0;
}

int tc_destroy(tc_solver_t * s){
  if(!s) return 0;
  hipSetDevice(s->p.device);
  if(s->stream) hipStreamSynchronize(s->stream);
  tc_closure_release(s);
  for(int n = 0; n < s->nalloc; n++) hipFree(s->allocs[n]);
  for(int n = 0; n < TC_NSTAGE; n++){
    for(int e = 0; e < TC_MAXEV; e++){
      if(s->prof[n].ev[e][0]){ hipEventDestroy(s->prof[n].ev[e][0]); hipEventDestroy(s->prof[n].ev[e][1]); }
    }
  }
  for(int f = 0; f < TC_NFIELD; f++) if(s->d_corr[f]) hipFree(s->d_corr[f]);
  if(s->d_cnn_xc) hipFree(s->d_cnn_xc);
  if(s->h_stage) hipHostFree(s->h_stage);
  if(s->stream && s->own_stream) hipStreamDestroy(s->stream);
  free(s->h_tab); free(s->h_xf); free(s->h_xc); free(s->h_pl); free(s->h_pu);
  free(s);
  return 0;
}

size_t tc_device_bytes(const tc_solver_t * s){
  return s ? s->dev_bytes : 0;
}

/* ---------------------------------------------------------------------------------------
 * boundary values and host <-> device copies
 * ------------------------------------------------------------------------------------- */
int tc_apply_bc(tc_solver_t * s, int field, double * a){
  int kind;
  double wm = 0., wp = 0.;
  switch(field){
    case TC_FUX: kind = 0; break;
    case TC_FUY: kind = 1; wm = s->p.uy_xm; wp = s->p.uy_xp; break;
    case TC_FUZ: kind = 2; break;
    case TC_FT: kind = 3; wm = s->p.t_xm; wp = s->p.t_xp; break;
    default: kind = 4; break;
  }
  KOK(tck_launch_bc(&s->g, a, kind, wm, wp, s->stream));
  return 0;
}

static int field_id(const char * name){
  if(!strcmp(name, "ux")) return TC_FUX;
  if(!strcmp(name, "uy")) return TC_FUY;
  if(!strcmp(name, "uz")) return TC_FUZ;
  if(!strcmp(name, "t")) return TC_FT;
  if(!strcmp(name, "p")) return TC_FP;
  if(!strcmp(name, "psi")) return TC_FPSI;
  return -1;
}

double * tc_field_ptr(tc_solver_t * s, int f){
  if(f >= 0 && f < TC_NFIELD) return s->u[f][s->cur[f]];
  if(f == TC_FP) return s->d_p;
  if(f == TC_FPSI) return s->d_psi;
  return NULL;
}

/* NPY x extent and first device slot */
static void npy_x(const tc_solver_t * s, int f, int * nxs, int * i0){
  if(f == TC_FUX){ *nxs = s->p.nx + 1; *i0 = 1; }
  else { *nxs = s->p.nx + 2; *i0 = 0; }
}

/* ---------------------------------------------------------------------------------------
 * host transfers of fields: whole planes through a pinned staging buffer (at most
 * TC_STAGE_BYTES, at least one plane), rows repacked on the host.  One contiguous copy per
 * chunk of planes instead of one strided copy per row keeps PCIe busy at any grid size.
 * ------------------------------------------------------------------------------------- */
#define TC_STAGE_BYTES ((size_t)64 << 20)

static int stage_buffer(tc_solver_t * s, size_t plane_bytes, size_t total, void ** out){
  size_t want = total < TC_STAGE_BYTES ? total : TC_STAGE_BYTES;
  if(want < plane_bytes) want = plane_bytes;
  if(s->h_stage_bytes < want){
    if(s->h_stage) hipHostFree(s->h_stage);
    s->h_stage = NULL;
    s->h_stage_bytes = 0;
    HIPOK(hipHostMalloc(&s->h_stage, want, hipHostMallocDefault));
    s->h_stage_bytes = want;
  }
  *out = s->h_stage;
  return 0;
}

/* rows j0..j1 of planes k0..k1 of a device field (elements of esz bytes, nxs values from
 * slot i0) <-> host rows at host_row(j, k) * nxs.  to_host: device -> host (widened from
 * FP32 to FP64 when widen); otherwise host -> device, the planes read first so that what
 * lies between the rows (halos, padding) is kept. */
typedef size_t (*host_row_fn)(const tc_geom_t * g, int j, int k);
static size_t row_npy(const tc_geom_t * g, int j, int k){ return (size_t)(k - 1) * g->ny + (size_t)(j - 1); }
static size_t row_ref(const tc_geom_t * g, int j, int k){ return (size_t)k * (g->ny + 2) + (size_t)j; }

static int xfer_rows(tc_solver_t * s, void * dev, size_t esz, int nxs, int i0, int j0, int j1,
    int k0, int k1, host_row_fn host_row, void * host, int to_host, int widen){
  const tc_geom_t * g = &s->g;
  const size_t plane = (size_t)g->px * g->py;
  const size_t plane_bytes = plane * esz;
  void * stage = NULL;
  if(stage_buffer(s, plane_bytes, plane_bytes * (size_t)(k1 - k0 + 1), &stage)) return 1;
  const int nk = (int)(s->h_stage_bytes / plane_bytes);
  const size_t hesz = widen ? sizeof(double) : esz;
  for(int kc = k0; kc <= k1; kc += nk){
    const int nc = kc + nk - 1 <= k1 ? nk : k1 - kc + 1;
    char * dp = (char *)dev + (size_t)kc * plane_bytes;
    HIPOK(hipMemcpyAsync(stage, dp, (size_t)nc * plane_bytes, hipMemcpyDeviceToHost, s->stream));
    HIPOK(hipStreamSynchronize(s->stream));
    for(int k = kc; k < kc + nc; k++){
      for(int j = j0; j <= j1; j++){
        char * sr = (char *)stage + ((size_t)(k - kc) * plane + (size_t)j * g->px + (size_t)(i0 + TC_XOFF)) * esz;
        char * hr = (char *)host + host_row(g, j, k) * (size_t)nxs * hesz;
        if(!to_host) memcpy(sr, hr, (size_t)nxs * esz);
        else if(!widen) memcpy(hr, sr, (size_t)nxs * esz);
        else for(int q = 0; q < nxs; q++) ((double *)hr)[q] = (double)((const float *)sr)[q];
      }
    }
    if(!to_host){
      HIPOK(hipMemcpyAsync(dp, stage, (size_t)nc * plane_bytes, hipMemcpyHostToDevice, s->stream));
      HIPOK(hipStreamSynchronize(s->stream));
    }
  }
  return 0;
}

/* layout 0: NPY (nz, ny, nxs); layout 1: the reference's host array with one halo plane on
 * each side in y and z, (nz+2, ny+2, nxs) (include/array_macros/fluid/*.h) */
static int set_field_layout(tc_solver_t * s, const char * name, const double * data, int layout){
  const int f = field_id(name);
  if(f < 0) return tc_fail("unknown field '%s'", name);
  if(f == TC_FPSI && s->mixed) return tc_fail("psi is an output of the mixed step and cannot be set");
  HIPOK(hipSetDevice(s->p.device));
  int nxs, i0;
  npy_x(s, f, &nxs, &i0);
  const tc_geom_t * g = &s->g;
  if(xfer_rows(s, tc_field_ptr(s, f), sizeof(double), nxs, i0, 1, g->ny, 1, g->nz,
      layout == 0 ? row_npy : row_ref, (void *)data, 0, 0)) return 1;
  return tc_finish_field(s, f);
}

int tc_set_field(tc_solver_t * s, const char * name, const double * data){
  return set_field_layout(s, name, data, 0);
}

int tc_set_field_ref(tc_solver_t * s, const char * name, const double * data){
  return set_field_layout(s, name, data, 1);
}

/* after the interior of a field was replaced: boundary values and halos, the second buffer
 * of a double-buffered field, and invalidation of what was derived from the old field */
int tc_finish_field(tc_solver_t * s, int f){
  const tc_geom_t * g = &s->g;
  double * dst = tc_field_ptr(s, f);
  if(tc_apply_bc(s, f, dst)) return 1;
  if(f < TC_NFIELD){
    HIPOK(hipMemcpyAsync(s->u[f][1 - s->cur[f]], dst, sizeof(double) * g->n3, hipMemcpyDeviceToDevice, s->stream));
  }
  HIPOK(hipStreamSynchronize(s->stream));
  s->dt_valid = 0;
  s->diag_valid = 0;
  s->gen++;
  return 0;
}

static int get_field_layout(tc_solver_t * s, const char * name, double * data, int layout){
  const int f = field_id(name);
  if(f < 0) return tc_fail("unknown field '%s'", name);
  HIPOK(hipSetDevice(s->p.device));
  int nxs, i0;
  npy_x(s, f, &nxs, &i0);
  const tc_geom_t * g = &s->g;
  /* psi of the mixed step is FP32, widened on the host */
  const int w32 = f == TC_FPSI && s->mixed;
  void * src = w32 ? (void *)s->d_psi32 : (void *)tc_field_ptr(s, f);
  const size_t esz = w32 ? sizeof(float) : sizeof(double);
  if(layout == 0) return xfer_rows(s, src, esz, nxs, i0, 1, g->ny, 1, g->nz, row_npy, data, 1, w32);
  return xfer_rows(s, src, esz, nxs, i0, 0, g->py - 1, 0, g->pz - 1, row_ref, data, 1, w32);
}

int tc_get_field(tc_solver_t * s, const char * name, double * data){
  return get_field_layout(s, name, data, 0);
}

int tc_get_field_ref(tc_solver_t * s, const char * name, double * data){
  return get_field_layout(s, name, data, 1);
}

/* decide_dt coefficients (decide_dt.c:128-129) and the fixed step */
int tc_set_dt_params(tc_solver_t * s, double coef_dt_adv, double coef_dt_dif, double dt_fixed){
  s->p.coef_dt_adv = coef_dt_adv;
  s->p.coef_dt_dif = coef_dt_dif;
  s->p.dt_fixed = dt_fixed;
  const double diffusivity = fmax(s->nu, s->kappa);
  const int nx = s->p.nx;
  double grid[3];
  grid[0] = DBL_MAX;
  for(int i = 0; i <= nx + 1; i++) grid[0] = fmin(grid[0], s->h_tab[(size_t)T_HXXC * s->tp + i]);
  grid[1] = s->h_tab[(size_t)T_HYXC * s->tp + 1];
  grid[2] = s->hz;
  double lim = INFINITY;
  for(int d = 0; d < 3; d++){
    const double dif = coef_dt_dif / diffusivity * 0.5 / 3 * pow(grid[d], 2.);
    if(!s->p.implicit[d]) lim = fmin(lim, dif);
  }
  s->dif_limit = lim;
  s->dt_valid = 0;
  s->gen++;
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * profiling: events around the kernels of each stage
 * ------------------------------------------------------------------------------------- */
static const char * g_stage_names[TC_NSTAGE] = {"dt", "factor", "predict", "xsolve", "ysolve", "zsolve",
  "div_yfft", "zfft", "pxsolve", "yifft", "correct", "diag"};

int tc_profile_enable(tc_solver_t * s, int on){
  s->prof_on = on;
  s->gen++;
  if(on){
    for(int n = 0; n < TC_NSTAGE; n++){
      s->prof[n].nrec = 0;
      s->prof[n].total_ms = 0.;
      s->prof[n].launches = 0;
    }
  }
  return 0;
}

static void prof_flush(tc_solver_t * s, int n){
  tc_prof_t * P = &s->prof[n];
  if(P->nrec == 0) return;
  hipEventSynchronize(P->ev[P->nrec - 1][1]);
  for(int e = 0; e < P->nrec; e++){
    float ms = 0.f;
    hipEventElapsedTime(&ms, P->ev[e][0], P->ev[e][1]);
    P->total_ms += ms;
    P->launches++;
  }
  P->nrec = 0;
}

void tc_prof_begin(tc_solver_t * s, int stage){
  if(!s->prof_on) return;
  tc_prof_t * P = &s->prof[stage];
  if(P->nrec == TC_MAXEV) prof_flush(s, stage);
  if(!P->ev[P->nrec][0]){
    hipEventCreate(&P->ev[P->nrec][0]);
    hipEventCreate(&P->ev[P->nrec][1]);
  }
  hipEventRecord(P->ev[P->nrec][0], s->stream);
}

void tc_prof_end(tc_solver_t * s, int stage){
  if(!s->prof_on) return;
  tc_prof_t * P = &s->prof[stage];
  hipEventRecord(P->ev[P->nrec][1], s->stream);
  P->nrec++;
}

int tc_profile_read(tc_solver_t * s, const char * stage, double * total_ms, long long * launches){
  for(int n = 0; n < TC_NSTAGE; n++){
    if(!strcmp(stage, g_stage_names[n])){
      prof_flush(s, n);
      *total_ms = s->prof[n].total_ms;
      *launches = s->prof[n].launches;
      return 0;
    }
  }
  return tc_fail("unknown stage '%s'", stage);
}

#define PROF(stage, call) do { tc_prof_begin(s, stage); int e__ = (call); tc_prof_end(s, stage); \
  if(e__) return tc_fail("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString((hipError_t)e__)); } while(0)

/* ---------------------------------------------------------------------------------------
 * the step
 * ------------------------------------------------------------------------------------- */
static int launch_dt(tc_solver_t * s){
  if(s->p.dt_fixed > 0.){
    PROF(TC_S_DT, tck_launch_dt_finalise(s->d_dtmin, s->d_dt, s->p.coef_dt_adv, s->dif_limit, s->p.dt_fixed, s->stream));
    return 0;
  }
  if(!s->dt_valid){
    PROF(TC_S_DT, tck_launch_dt(&s->g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
        tc_field_ptr(s, TC_FUZ), s->d_dtmin, s->stream));
  }
  PROF(TC_S_DT, tck_launch_dt_finalise(s->d_dtmin, s->d_dt, s->p.coef_dt_adv, s->dif_limit, 0., s->stream));
  s->dt_valid = 0;
  return 0;
}

/* stale-tensor diagnostics of the reference's logging: gradients at the start of the last
 * predict (injection.c:49-51, dissipation.c:36-229) */
static int snapshot_diag(tc_solver_t * s){
  PROF(TC_S_DIAG, tck_launch_diag_walls(&s->g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FT), s->p.scalar, s->d_part, s->d_diag + 8, s->stream));
  PROF(TC_S_DIAG, tck_launch_diag_cells(&s->g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FUZ), tc_field_ptr(s, TC_FT), s->p.scalar, s->d_part, s->d_diag, s->stream));
  HIPOK(hipMemcpyAsync(s->h_diag_snap, s->d_diag, sizeof(double) * 10, hipMemcpyDeviceToHost, s->stream));
  s->diag_pending = 1;
  return 0;
}

/* LU factors of the implicit-diffusion systems of substeps r0..r1 for every advected field,
 * from the dt on the device (one launch) */
static int factor_tables(tc_solver_t * s, int r0, int r1){
  if(!(s->p.implicit[0] || s->p.implicit[1] || s->p.implicit[2])) return 0;
  const int nf = s->p.scalar ? 4 : 3;
  /* fac: 2 on the diagonal stress direction (ux.c:499-506, uy.c:539-544, uz.c:449-454)
   * and for T in y (t.c:346-349) */
  static const double facs[4][3] = {{2., 1., 1.}, {1., 2., 1.}, {1., 1., 2.}, {1., 2., 1.}};
  int rk[12], i0[12], n[12];
  double diff[12], fac[12][3];
  const double * lx[12][3];
  const double * ly[12][3];
  double * xt[12][3];
  double * yt[12][4];
  double * zt[12][4];
  const tc_coef_t * c = &s->coef;
  int nj = 0;
  for(int r = r0; r <= r1; r++){
    for(int f = 0; f < nf; f++){
      const int ux = f == TC_FUX;
      const tc_ls_t * L = &s->ls[f][r];
      rk[nj] = r;
      diff[nj] = f == TC_FT ? s->kappa : s->nu;
      for(int d = 0; d < 3; d++) fac[nj][d] = facs[f][d];
      i0[nj] = ux ? 2 : 1;
      n[nj] = ux ? s->p.nx - 1 : s->p.nx;
      lx[nj][0] = ux ? c->lxu_l : c->lxc_l; lx[nj][1] = ux ? c->lxu_c : c->lxc_c; lx[nj][2] = ux ? c->lxu_u : c->lxc_u;
      ly[nj][0] = ux ? c->lyu_l : c->lyc_l; ly[nj][1] = ux ? c->lyu_c : c->lyc_c; ly[nj][2] = ux ? c->lyu_u : c->lyc_u;
      xt[nj][0] = L->xr; xt[nj][1] = L->xv; xt[nj][2] = L->xl;
      yt[nj][0] = L->yr; yt[nj][1] = L->yv; yt[nj][2] = L->yq1; yt[nj][3] = L->ymeta;
      zt[nj][0] = L->zr; zt[nj][1] = L->zv; zt[nj][2] = L->zq1; zt[nj][3] = L->zmeta;
      nj++;
    }
  }
  PROF(TC_S_FACTOR, tck_launch_factor(s->d_dt, s->p.implicit, s->p.nx, s->p.ny, s->p.nz, s->tp,
      c->lapz_l, c->lapz_c, c->lapz_u, nj, rk, diff, (const double (*)[3])fac, i0, n,
      (const double * const (*)[3])lx, (const double * const (*)[3])ly,
      (double * const (*)[3])xt, (double * const (*)[4])yt, (double * const (*)[4])zt, s->stream));
  return 0;
}

static int predict(tc_solver_t * s, int rkstep){
  const int nf = s->p.scalar ? 4 : 3;
  const int any_implicit = s->p.implicit[0] || s->p.implicit[1] || s->p.implicit[2];
  if(s->diag_armed && rkstep == 2){
    if(snapshot_diag(s)) return 1;
  }
  if(s->stat_armed && rkstep == 2){
    /* the tensors that statistics.c:221-234 reads are those of the last substep */
    PROF(TC_S_DIAG, tck_launch_stat(&s->g, &s->coef, 0, tc_field_ptr(s, TC_FUX),
        tc_field_ptr(s, TC_FUY), tc_field_ptr(s, TC_FUZ), s->d_stat, s->stream));
  }
  const double * cur[4];
  double * nxt[4];
  void * dlt[4];
  const void * srcb[4];
  void * srca[4];
  for(int f = 0; f < 4; f++){
    cur[f] = s->u[f][s->cur[f]];
    nxt[f] = s->u[f][1 - s->cur[f]];
    dlt[f] = s->mixed ? s->dlt[f] : (void *)nxt[f];
    srcb[f] = rkstep > 0 ? s->src[f][(rkstep - 1) & 1] : NULL;
    srca[f] = rkstep < 2 ? s->src[f][rkstep & 1] : NULL;
  }
  const float * corr[4] = {NULL, NULL, NULL, NULL};
  if(s->closure) for(int f = 0; f < 4; f++) corr[f] = (const float *)s->d_corr[f];
  PROF(TC_S_PREDICT, tck_launch_predict(&s->g, &s->coef, &s->flags, cur, s->d_p, nxt, dlt, srcb, srca,
      corr, s->d_dt, rkstep, !any_implicit, s->mixed, s->tune[TC_TUNE_PREDICT][0], s->tune[TC_TUNE_PREDICT][1], s->predict_window, s->stream));
  if(any_implicit){
    int last = 0;
    for(int d = 0; d < 3; d++) if(s->p.implicit[d]) last = d;
    /* every field's solves run x, y, z in order; the fields are independent, so each
     * direction's solves of all fields go in one launch */
    void * xd[4]; double * xo[4]; const double * xc[4];
    const double * xr[4]; const double * xv[4]; const double * xl[4]; const double * xq[4];
    const double * xm[4];
    int xi0[4], xn[4];
    for(int f = 0; f < nf; f++){
      xd[f] = dlt[f]; xo[f] = nxt[f];
      xi0[f] = f == TC_FUX ? 2 : 1;
      xn[f] = f == TC_FUX ? s->p.nx - 1 : s->p.nx;
    }
    if(s->p.implicit[0]){
      for(int f = 0; f < nf; f++){
        tc_ls_t * L = &s->ls[f][rkstep];
        xc[f] = last == 0 ? cur[f] : NULL;
        xr[f] = L->xr; xv[f] = L->xv; xl[f] = L->xl;
      }
      PROF(TC_S_XSOLVE, tck_launch_xsolve(&s->g, nf, xd, xo, xc, xr, xv, xl, xi0, xn, s->mixed, s->stream));
    }
    if(s->p.implicit[1]){
      for(int f = 0; f < nf; f++){
        tc_ls_t * L = &s->ls[f][rkstep];
        xc[f] = last == 1 ? cur[f] : NULL;
        xr[f] = L->yr; xv[f] = L->yv; xq[f] = L->yq1; xm[f] = L->ymeta;
      }
      PROF(TC_S_YSOLVE, tck_launch_psolve(&s->g, 1, nf, xd, xo, xc, xr, xv, xq, xm, s->tp, 1, xi0, xn,
          s->mixed, s->psolve_lean, s->stream));
    }
    if(s->p.implicit[2]){
      for(int f = 0; f < nf; f++){
        tc_ls_t * L = &s->ls[f][rkstep];
        xc[f] = cur[f];
        xr[f] = L->zr; xv[f] = L->zv; xq[f] = L->zq1; xm[f] = L->zmeta;
      }
      PROF(TC_S_ZSOLVE, tck_launch_psolve(&s->g, 2, nf, xd, xo, xc, xr, xv, xq, xm, 1, 0, xi0, xn,
          s->mixed, s->psolve_lean, s->stream));
    }
  }
  for(int f = 0; f < nf; f++) s->cur[f] = 1 - s->cur[f];
  return 0;
}

static int potential(tc_solver_t * s, int rkstep){
  const tc_geom_t * g = &s->g;
  const int mx = s->mixed;
  void * psi = mx ? (void *)s->d_psi32 : (void *)s->d_psi;
  PROF(TC_S_DIVY, tck_launch_div_yfft(g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FUZ), s->d_dt, rkstep, s->d_spec, s->d_twy, s->d_twyn, mx, s->fft_exact & 2, s->stream));
  PROF(TC_S_ZFFT, tck_launch_zfft(g->nx, s->nky, g->nz, s->d_spec, s->d_twz, -1, mx, s->fft_exact & 1, s->stream));
  PROF(TC_S_PXSOLVE, tck_launch_pxsolve(g->nx, s->nky * g->nz, s->d_spec, s->d_pr, s->d_pl, s->d_pu, mx, s->stream));
  PROF(TC_S_ZFFT, tck_launch_zfft(g->nx, s->nky, g->nz, s->d_spec, s->d_twz, +1, mx, s->fft_exact & 1, s->stream));
  PROF(TC_S_YIFFT, tck_launch_yifft_psi(g, s->d_spec, psi, s->d_twy, s->d_twyn, mx, s->fft_exact & 4, s->stream));
  return 0;
}

/* mode: 1 velocity correction (correct/*.c), 2 pressure update (update_pressure.c), 3 both */
static int correct(tc_solver_t * s, int rkstep, int mode, int want_cfl){
  PROF(TC_S_CORRECT, tck_launch_correct(&s->g, &s->coef, &s->flags, tc_field_ptr(s, TC_FUX),
      tc_field_ptr(s, TC_FUY), tc_field_ptr(s, TC_FUZ), s->d_p,
      s->mixed ? (const void *)s->d_psi32 : (const void *)s->d_psi, s->d_dt, rkstep, mode,
      want_cfl ? s->d_dtmin : NULL, s->mixed, s->tune[TC_TUNE_CORRECT][0], s->tune[TC_TUNE_CORRECT][1], s->stream));
  return 0;
}

int tc_enqueue_substep(tc_solver_t * s, int r){
  s->gen++;
  if(r == 0){
    if(launch_dt(s)) return 1;
    if(factor_tables(s, 0, 2)) return 1;
    /* the learned correction of this step, from the fields at its start */
    if(s->closure && tc_closure_eval(s)) return 1;
  }
  if(predict(s, r)) return 1;
  if(potential(s, r)) return 1;
  /* the last correction also yields the CFL minimum of the next step (decide_dt.c:32-80
   * reads exactly the velocity this kernel writes) */
  const int cfl = (r == 2) && !(s->p.dt_fixed > 0.);
  const int refine = s->refine && r == 2;
  if(correct(s, r, 3, cfl && !refine)) return 1;
  if(refine){
    /* defect correction: the FP64 divergence of the corrected velocity, its potential and a
     * second velocity correction (the pressure keeps the first potential) */
    if(potential(s, r)) return 1;
    if(correct(s, r, 1, cfl)) return 1;
  }
  if(cfl) s->dt_valid = 1;
  return 0;
}

int tc_enqueue_step(tc_solver_t * s){
  for(int r = 0; r < 3; r++){
    if(tc_enqueue_substep(s, r)) return 1;
  }
  return 0;
}

void tc_advance_host_state(tc_solver_t * s){
  const int nf = s->p.scalar ? 4 : 3;
  for(int r = 0; r < 3; r++){
    for(int f = 0; f < nf; f++) s->cur[f] = 1 - s->cur[f];
  }
  s->dt_valid = !(s->p.dt_fixed > 0.);
  s->diag_valid = 0;
}

int tc_share_stream(tc_solver_t * s, hipStream_t stream){
  HIPOK(hipSetDevice(s->p.device));
  HIPOK(hipStreamSynchronize(s->stream));
  if(s->own_stream) HIPOK(hipStreamDestroy(s->stream));
  s->stream = stream;
  s->own_stream = 0;
  return 0;
}

int tc_autotune(tc_solver_t * s, int reps, int * kb_out, int * ty_out, double * best_ms){
  static const int kbs[5] = {2, 4, 8, 16, 32};
  static const int tys[3] = {1, 2, 4};
  HIPOK(hipSetDevice(s->p.device));
  const int any_implicit = s->p.implicit[0] || s->p.implicit[1] || s->p.implicit[2];
  const double * cur[4];
  double * nxt[4];
  void * dlt[4];
  const void * srcb[4];
  void * srca[4];
  for(int f = 0; f < 4; f++){
    cur[f] = s->u[f][s->cur[f]];
    nxt[f] = s->u[f][1 - s->cur[f]];
    dlt[f] = s->mixed ? s->dlt[f] : (void *)nxt[f];
    srcb[f] = s->src[f][0];
    srca[f] = s->src[f][1];
  }
  hipEvent_t e0, e1;
  HIPOK(hipEventCreate(&e0));
  HIPOK(hipEventCreate(&e1));
  double best = 1e300;
  int bkb = s->tune[TC_TUNE_PREDICT][0], bty = s->tune[TC_TUNE_PREDICT][1];
  int rc = 0;
  for(int a = 0; a < 5 && !rc; a++){
    if(kbs[a] > s->p.nz) continue;
    for(int b = 0; b < 3 && !rc; b++){
      /* one untimed launch, then reps timed ones */
      for(int q = 0; q <= reps && !rc; q++){
        if(q == 1 && hipEventRecord(e0, s->stream) != hipSuccess) rc = 1;
        if(!rc) rc = tck_launch_predict(&s->g, &s->coef, &s->flags, cur, s->d_p, nxt, dlt, srcb, srca,
            NULL, s->d_dt, 1, !any_implicit, s->mixed, kbs[a], tys[b], s->predict_window, s->stream);
      }
      float ms = 0.f;
      if(!rc && (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) rc = 1;
      if(!rc && ms / reps < best){
        best = ms / reps;
        bkb = kbs[a];
        bty = tys[b];
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if(rc) return tc_fail("tc_autotune: launch or timing failed");
  s->tune[TC_TUNE_PREDICT][0] = bkb; s->tune[TC_TUNE_PREDICT][1] = bty;
  s->tune[TC_TUNE_CORRECT][0] = bkb; s->tune[TC_TUNE_CORRECT][1] = bty;
  s->gen++;
  if(kb_out) *kb_out = bkb;
  if(ty_out) *ty_out = bty;
  if(best_ms) *best_ms = best;
  return 0;
}

int tc_step(tc_solver_t * s, double * dt){
  HIPOK(hipSetDevice(s->p.device));
  if(tc_enqueue_step(s)) return 1;
  HIPOK(hipMemcpyAsync(&s->h_dt, s->d_dt, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  if(s->diag_pending){
    memcpy(s->diag_vals, s->h_diag_snap, sizeof(s->diag_vals));
    s->diag_valid = 1;
    s->diag_pending = 0;
  }
  if(dt) *dt = s->h_dt;
  return 0;
}

int tc_step_async(tc_solver_t * s, int nsteps){
  HIPOK(hipSetDevice(s->p.device));
  for(int n = 0; n < nsteps; n++){
    if(tc_enqueue_step(s)) return 1;
  }
  return 0;
}

int tc_synchronize(tc_solver_t * s){
  HIPOK(hipSetDevice(s->p.device));
  HIPOK(hipStreamSynchronize(s->stream));
  if(s->diag_pending){
    memcpy(s->diag_vals, s->h_diag_snap, sizeof(s->diag_vals));
    s->diag_valid = 1;
    s->diag_pending = 0;
  }
  return 0;
}

int tc_get_dt(tc_solver_t * s, double * dt){
  HIPOK(hipMemcpyAsync(&s->h_dt, s->d_dt, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  *dt = s->h_dt;
  return 0;
}

int tc_decide_dt(tc_solver_t * s, double * dt){
  HIPOK(hipSetDevice(s->p.device));
  s->gen++;
  s->dt_valid = 0;
  if(launch_dt(s)) return 1;
  return tc_get_dt(s, dt);
}

static int set_dt(tc_solver_t * s, double dt){
  s->h_dt = dt;
  HIPOK(hipMemcpyAsync(s->d_dt, &s->h_dt, sizeof(double), hipMemcpyHostToDevice, s->stream));
  return 0;
}

int tc_predict(tc_solver_t * s, int rkstep, double dt){
  if(rkstep < 0 || rkstep > 2) return tc_fail("rkstep %d out of range", rkstep);
  HIPOK(hipSetDevice(s->p.device));
  s->gen++;
  if(set_dt(s, dt) || factor_tables(s, rkstep, rkstep) || predict(s, rkstep)) return 1;
  s->dt_valid = 0;
  return tc_synchronize(s);
}

int tc_potential(tc_solver_t * s, int rkstep, double dt){
  if(rkstep < 0 || rkstep > 2) return tc_fail("rkstep %d out of range", rkstep);
  HIPOK(hipSetDevice(s->p.device));
  s->gen++;
  if(set_dt(s, dt) || potential(s, rkstep)) return 1;
  return tc_synchronize(s);
}

int tc_correct(tc_solver_t * s, int rkstep, double dt){
  if(rkstep < 0 || rkstep > 2) return tc_fail("rkstep %d out of range", rkstep);
  HIPOK(hipSetDevice(s->p.device));
  s->gen++;
  if(set_dt(s, dt) || correct(s, rkstep, 1, 0)) return 1;
  s->dt_valid = 0;
  return tc_synchronize(s);
}

int tc_pressure(tc_solver_t * s, int rkstep, double dt){
  if(rkstep < 0 || rkstep > 2) return tc_fail("rkstep %d out of range", rkstep);
  HIPOK(hipSetDevice(s->p.device));
  s->gen++;
  if(set_dt(s, dt) || correct(s, rkstep, 2, 0)) return 1;
  return tc_synchronize(s);
}

/* ---------------------------------------------------------------------------------------
 * diagnostics
 * ------------------------------------------------------------------------------------- */
static int cells_now(tc_solver_t * s, double out[8]){
  HIPOK(hipSetDevice(s->p.device));
  KOK(tck_launch_diag_cells(&s->g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FUZ), tc_field_ptr(s, TC_FT), s->p.scalar, s->d_part, s->d_diag, s->stream));
  HIPOK(hipMemcpyAsync(out, s->d_diag, sizeof(double) * 8, hipMemcpyDeviceToHost, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  return 0;
}

int tc_divergence(tc_solver_t * s, double * divmax){
  double q[8];
  if(cells_now(s, q)) return 1;
  *divmax = q[0];
  return 0;
}

int tc_total_energy(tc_solver_t * s, double out[4]){
  double q[8];
  if(cells_now(s, q)) return 1;
  out[0] = q[3]; out[1] = q[4]; out[2] = q[5]; out[3] = q[6];
  return 0;
}

/* values of the reference's logging: the velocity terms come from the tensors left by the
 * last predict (snapshot), the scalar terms read the current T (injection.c:75-87,
 * dissipation.c:243-318) */
static int logging_values(tc_solver_t * s, double vals[10]){
  HIPOK(hipSetDevice(s->p.device));
  KOK(tck_launch_diag_walls(&s->g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FT), s->p.scalar, s->d_part, s->d_diag + 8, s->stream));
  KOK(tck_launch_diag_cells(&s->g, &s->coef, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FUZ), tc_field_ptr(s, TC_FT), s->p.scalar, s->d_part, s->d_diag, s->stream));
  HIPOK(hipMemcpyAsync(vals, s->d_diag, sizeof(double) * 10, hipMemcpyDeviceToHost, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  if(s->diag_valid){
    vals[1] = s->diag_vals[1];
    vals[8] = s->diag_vals[8];
  }
  return 0;
}

int tc_injection(tc_solver_t * s, double out[2]){
  double v[10];
  if(logging_values(s, v)) return 1;
  out[0] = s->nu * v[8];
  out[1] = s->kappa * v[9];
  return 0;
}

int tc_dissipation(tc_solver_t * s, double out[2]){
  double v[10];
  if(logging_values(s, v)) return 1;
  out[0] = s->nu * v[1];
  out[1] = s->kappa * v[2];
  return 0;
}

/* launch shapes of the cell kernels ("predict", "correct"): planes per block and j rows per
 * block; results do not depend on them.  "predict_window" (kb 0 or 1) switches between forms
 * of the predictor, "fft_exact" (kb a mask: 1 the axial transforms, 2 the divergence + theta
 * transform's twiddles in LDS, 4 the inverse theta transform) between forms of the persistent
 * transforms, "psolve_lean" (kb a mask: 1 the periodic line solves' interior waves skip the row
 * tests, 2 their folds read LDS four segments at a time; default 3) between forms of the
 * periodic line solves; each pair computes the same bits */
int tc_tune_set(tc_solver_t * s, const char * kernel, int kb, int ty){
  if(kernel && strcmp(kernel, "predict_window") == 0){
    s->predict_window = kb ? 1 : 0;
    s->gen++;
    return 0;
  }
  if(kernel && strcmp(kernel, "fft_exact") == 0){
    s->fft_exact = kb & 7;
    s->gen++;
    return 0;
  }
  if(kernel && strcmp(kernel, "psolve_lean") == 0){
    s->psolve_lean = kb & 3;
    s->gen++;
    return 0;
  }
  int n = -1;
  if(kernel && strcmp(kernel, "predict") == 0) n = TC_TUNE_PREDICT;
  if(kernel && strcmp(kernel, "correct") == 0) n = TC_TUNE_CORRECT;
  if(n < 0) return tc_fail("tc_tune_set: unknown kernel '%s'", kernel ? kernel : "(null)");
  if(kb < 1 || kb > s->p.nz || !(ty == 1 || ty == 2 || ty == 4)) return tc_fail("tc_tune_set: kb %d, ty %d out of range", kb, ty);
  s->tune[n][0] = kb;
  s->tune[n][1] = ty;
  s->gen++;
  return 0;
}

int tc_arm_diagnostics(tc_solver_t * s, int on){
  s->diag_armed = on;
  s->gen++;
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * running statistics (src/statistics.c:56-307) on the device
 * ------------------------------------------------------------------------------------- */
static const char * const g_stat_names[8] = {"ux1", "ux2", "uy1", "uy2", "uz1", "uz2", "adv", "dif"};

static size_t stat_count(const tc_solver_t * s, int n){
  const size_t rows = (size_t)s->p.ny * (size_t)s->p.nz;
  const int face = (n == 0 || n == 1 || n == 6 || n == 7 || n == 8);
  return rows * (size_t)(face ? s->p.nx + 1 : s->p.nx + 2);
}

int tc_stat_arm(tc_solver_t * s, int on){
  HIPOK(hipSetDevice(s->p.device));
  if(on && !s->d_stat[0]){
    for(int n = 0; n < 9; n++){
      if(dalloc(s, &s->d_stat[n], stat_count(s, n))) return 1;
    }
    s->stat_num = 0;
  }
  s->stat_armed = on;
  s->gen++;
  return 0;
}

int tc_stat_reset(tc_solver_t * s){
  if(!s->d_stat[0]) return 0;
  HIPOK(hipSetDevice(s->p.device));
  for(int n = 0; n < 8; n++){
    HIPOK(hipMemsetAsync(s->d_stat[n], 0, sizeof(double) * stat_count(s, n), s->stream));
  }
  s->stat_num = 0;
  return 0;
}

int tc_stat_collect(tc_solver_t * s){
  if(!s->d_stat[0]) return tc_fail("tc_stat_collect: statistics are not armed (tc_stat_arm)");
  HIPOK(hipSetDevice(s->p.device));
  KOK(tck_launch_stat(&s->g, &s->coef, 1, tc_field_ptr(s, TC_FUX), tc_field_ptr(s, TC_FUY),
      tc_field_ptr(s, TC_FUZ), s->d_stat, s->stream));
  s->stat_num += 1;
  return 0;
}

int tc_stat_get(tc_solver_t * s, const char * name, double * host, size_t * num){
  if(!s->d_stat[0]) return tc_fail("tc_stat_get: statistics are not armed (tc_stat_arm)");
  int n = -1;
  for(int m = 0; m < 8; m++) if(name && strcmp(name, g_stat_names[m]) == 0) n = m;
  if(n < 0) return tc_fail("tc_stat_get: unknown statistic '%s'", name ? name : "(null)");
  HIPOK(hipSetDevice(s->p.device));
  if(host){
    HIPOK(hipMemcpyAsync(host, s->d_stat[n], sizeof(double) * stat_count(s, n), hipMemcpyDeviceToHost, s->stream));
    HIPOK(hipStreamSynchronize(s->stream));
  }
  if(num) *num = s->stat_num;
  return 0;
}

/* the copy ceiling: dst = src over nbytes, the best of 1, 2, 4 and 8 independent 16 B chunks
 * per lane (tck_copy), iters copies each after one untimed one; read + write bytes per second */
int tc_probe_copy_bandwidth(tc_solver_t * s, size_t nbytes, int iters, double * gbps){
  HIPOK(hipSetDevice(s->p.device));
  if(iters < 1 || nbytes < 4096) return tc_fail("tc_probe_copy_bandwidth: nbytes %zu, iters %d", nbytes, iters);
  double * a = NULL;
  double * b = NULL;
  HIPOK(hipMalloc((void **)&a, nbytes));
  if(hipMalloc((void **)&b, nbytes) != hipSuccess){ hipFree(a); return tc_fail("hipMalloc failed"); }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const long long n = (long long)(nbytes / sizeof(double)) & ~1LL;
  int rc = 0;
  double best = 0.;
  static const int chunks[4] = {1, 2, 4, 8};
  for(int c = 0; c < 4 && !rc; c++){
    rc = tck_launch_copy(a, b, n, chunks[c], s->stream);
    hipEventRecord(e0, s->stream);
    for(int it = 0; it < iters && !rc; it++) rc = tck_launch_copy(it & 1 ? b : a, it & 1 ? a : b, n, chunks[c], s->stream);
    hipEventRecord(e1, s->stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    const double v = 2.0 * (double)n * sizeof(double) * iters / (ms * 1e-3) / 1e9;
    if(!rc && v > best) best = v;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if(rc) return tc_fail("copy kernel failed");
  *gbps = best;
  return 0;
}
