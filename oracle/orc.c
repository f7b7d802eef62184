/*
 * orc.c -- CPU restatement ("oracle") of NaokiHori/SimpleTCSolver's RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  C99 + OpenMP over the outermost loop of each
 * sweep.  Arithmetic is written with the same operand order as the reference so that, with
 * -ffp-contract=off, each stencil term rounds exactly as the reference's does.  The sweep
 * structure (one pass per term, as in src/fluid/predict/*.c) is kept on purpose: this file is
 * also timed as the CPU baseline, and its cost should look like the reference's.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <stdbool.h>
#include "orc.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double l, c, u; } lap_t;

struct orc_t_ {
  orc_config_t c;
  int nx, ny, nz;
  double * xf; double * xc;
  double * hxxf; double * hxxc; double * hyxf; double * hyxc; double * jdxf; double * jdxc;
  double hz;
  double * ux; double * uy; double * uz; double * p; double * psi; double * t;
  double * srcux[3]; double * srcuy[3]; double * srcuz[3]; double * srct[3];
  double * lxx; double * lyx0; double * lyx1; double * lxy; double * lzx; double * lxz;
  double * lyy0; double * lyy1; double * lzy; double * lyz; double * lzz;
  lap_t * lapx_ux; lap_t * lapy_ux;     /* predict/ux.c:34-81, i = 2..nx  */
  lap_t * lapx_c;  lap_t * lapy_c;      /* predict/{uy,uz,t}.c init_laplacians, i = 1..nx */
  lap_t lapz;
  double * delta;                       /* x1 pencil of linear_system_t (stores the increment) */
  /* Poisson solver (fluid/compute_potential.c) */
  double * eval_y; double * eval_z;     /* :311-349 */
  double * ptdm_l; double * ptdm_u;     /* :173-211 */
  double * spec;                        /* complex nx x (ny/2+1) x nz, i fastest */
  /* statistics.c accumulators (allocated by the first orc_stat_collect) */
  double * st[8];
  size_t st_num;
};

/* ---- layouts: include/array_macros/fluid/*.h, include/array_macros/domain/*.h ---- */
#define NX (o->nx)
#define NY (o->ny)
#define NZ (o->nz)
#define IU(i, j, k) ((size_t)(i) + (size_t)(NX + 2) * ((size_t)(j) + (size_t)(NY + 2) * (size_t)(k)))
#define IF(i, j, k) ((size_t)((i) - 1) + (size_t)(NX + 1) * ((size_t)(j) + (size_t)(NY + 2) * (size_t)(k)))
#define IC(i, j, k) ((size_t)((i) - 1) + (size_t)(NX) * ((size_t)(j) + (size_t)(NY + 2) * (size_t)(k)))
#define UX(i, j, k) (o->ux[IF(i, j, k)])
#define UY(i, j, k) (o->uy[IU(i, j, k)])
#define UZ(i, j, k) (o->uz[IU(i, j, k)])
#define P(i, j, k) (o->p[IU(i, j, k)])
#define PSI(i, j, k) (o->psi[IU(i, j, k)])
#define T(i, j, k) (o->t[IU(i, j, k)])
#define LXX(i, j, k) (o->lxx[IC(i, j, k)])
#define LYY0(i, j, k) (o->lyy0[IC(i, j, k)])
#define LYY1(i, j, k) (o->lyy1[IC(i, j, k)])
#define LZY(i, j, k) (o->lzy[IC(i, j, k)])
#define LYZ(i, j, k) (o->lyz[IC(i, j, k)])
#define LZZ(i, j, k) (o->lzz[IC(i, j, k)])
#define LXY(i, j, k) (o->lxy[IF(i, j, k)])
#define LXZ(i, j, k) (o->lxz[IF(i, j, k)])
#define LYX0(i, j, k) (o->lyx0[IF(i, j, k)])
#define LYX1(i, j, k) (o->lyx1[IF(i, j, k)])
#define LZX(i, j, k) (o->lzx[IF(i, j, k)])
#define XF(i) (o->xf[(i) - 1])
#define XC(i) (o->xc[(i)])
#define HXXF(i) (o->hxxf[(i) - 1])
#define HXXC(i) (o->hxxc[(i)])
#define HYXF(i) (o->hyxf[(i) - 1])
#define HYXC(i) (o->hyxc[(i)])
#define JDXF(i) (o->jdxf[(i) - 1])
#define JDXC(i) (o->jdxc[(i)])
/* RK buffers: SRCUX (i = 2..nx) and SRCU{Y,Z}/SRCT (i = 1..nx), no halos */
#define CNTX(i, j, k) ((size_t)((i) - 2) + (size_t)(NX - 1) * ((size_t)((j) - 1) + (size_t)NY * (size_t)((k) - 1)))
#define CNTC(i, j, k) ((size_t)((i) - 1) + (size_t)(NX) * ((size_t)((j) - 1) + (size_t)NY * (size_t)((k) - 1)))

#define KLOOP _Pragma("omp parallel for schedule(static)") for(int k = 1; k <= NZ; k++)

/* rkcoefs, src/runge_kutta.c:8-15 */
static const double RKA[3] = {32. / 60., 25. / 60., 45. / 60.};
static const double RKB[3] = { 0. / 60., -17. / 60., -25. / 60.};
static double rk_g(int s){ return RKA[s] + RKB[s]; }

void orc_rkcoefs(double out[9]){
  for(int s = 0; s < 3; s++){
    out[3 * s + 0] = RKA[s];
    out[3 * s + 1] = RKB[s];
    out[3 * s + 2] = RKA[s] + RKB[s];
  }
}

/* planted mistake (orc_set_mutation); ORC_MUT_NONE in every use but the pin tests */
static int mut = ORC_MUT_NONE;
void orc_set_mutation(int m){ mut = m; }

static double nu_of(const orc_t * o){ return 1. / o->c.Re; }               /* compute_diffusivity.c:3-7 */
static double kappa_of(const orc_t * o){ return 1. / o->c.Re / o->c.Pr; }  /* compute_diffusivity.c:9-13 */

static void * zalloc(size_t n, size_t s){
  void * p = calloc(n ? n : 1, s);
  return p;
}

/* ---------------------------------------------------------------------------------------
 * FFT: recursive mixed-radix decimation in time, out of place, interleaved complex.
 * X[f] = sum_n x[n] exp(sign 2 pi i n f / N).  Stands in for FFTW's plans
 * (compute_potential.c:247-309), which are unnormalised.
 * ------------------------------------------------------------------------------------- */
static int smallest_factor(int n){
  if(n % 4 == 0) return 4;
  if(n % 2 == 0) return 2;
  for(int f = 3; f * f <= n; f += 2){
    if(n % f == 0) return f;
  }
  return n;
}

/* twiddle table w[m] = exp(-2 pi i m / n) and a combine scratch of n complex values */
typedef struct { int n; double * w; double * scratch; double * tmp; } fftplan_t;

static void plan_init(fftplan_t * P, int n){
  const double two_pi = 6.283185307179586476925286766559;
  P->n = n;
  P->w = malloc(sizeof(double) * 2 * (size_t)n);
  P->scratch = malloc(sizeof(double) * 2 * (size_t)n);
  P->tmp = malloc(sizeof(double) * 2 * (size_t)n);
  for(int m = 0; m < n; m++){
    const double ang = - two_pi * (double)m / (double)n;
    P->w[2 * m + 0] = cos(ang);
    P->w[2 * m + 1] = sin(ang);
  }
}

static void plan_free(fftplan_t * P){
  free(P->w);
  free(P->scratch);
  free(P->tmp);
}

/* W_n^e for a sub-transform of length n dividing P->n; sign > 0 conjugates */
static inline void twiddle(const fftplan_t * P, int n, long e, int sign, double * wr, double * wi){
  const long m = (e % n) * (P->n / n);
  *wr = P->w[2 * m];
  *wi = sign > 0 ? - P->w[2 * m + 1] : P->w[2 * m + 1];
}

static void fft_rec(const fftplan_t * P, int n, int sign, const double * x, size_t stride, double * y){
  if(n == 1){
    y[0] = x[0];
    y[1] = x[1];
    return;
  }
  const int r = smallest_factor(n);
  const int m = n / r;
  for(int q = 0; q < r; q++){
    fft_rec(P, m, sign, x + 2 * stride * (size_t)q, stride * (size_t)r, y + 2 * (size_t)q * m);
  }
  double * tmp = P->tmp;
  double * out = P->scratch;
  for(int f = 0; f < m; f++){
    for(int q = 0; q < r; q++){
      double wr, wi;
      twiddle(P, n, (long)q * f, sign, &wr, &wi);
      const double ar = y[2 * ((size_t)q * m + f) + 0];
      const double ai = y[2 * ((size_t)q * m + f) + 1];
      tmp[2 * q + 0] = ar * wr - ai * wi;
      tmp[2 * q + 1] = ar * wi + ai * wr;
    }
    if(r == 2){
      out[2 * (size_t)f + 0] = tmp[0] + tmp[2];
      out[2 * (size_t)f + 1] = tmp[1] + tmp[3];
      out[2 * ((size_t)f + m) + 0] = tmp[0] - tmp[2];
      out[2 * ((size_t)f + m) + 1] = tmp[1] - tmp[3];
    }else if(r == 4){
      /* exact quarter-turn rotations by sign * i */
      const double s = (double)sign;
      const double a0r = tmp[0] + tmp[4], a0i = tmp[1] + tmp[5];
      const double a1r = tmp[0] - tmp[4], a1i = tmp[1] - tmp[5];
      const double b0r = tmp[2] + tmp[6], b0i = tmp[3] + tmp[7];
      const double b1r = tmp[2] - tmp[6], b1i = tmp[3] - tmp[7];
      out[2 * (size_t)f + 0] = a0r + b0r;
      out[2 * (size_t)f + 1] = a0i + b0i;
      out[2 * ((size_t)f + 2 * (size_t)m) + 0] = a0r - b0r;
      out[2 * ((size_t)f + 2 * (size_t)m) + 1] = a0i - b0i;
      out[2 * ((size_t)f + (size_t)m) + 0] = a1r - s * b1i;
      out[2 * ((size_t)f + (size_t)m) + 1] = a1i + s * b1r;
      out[2 * ((size_t)f + 3 * (size_t)m) + 0] = a1r + s * b1i;
      out[2 * ((size_t)f + 3 * (size_t)m) + 1] = a1i - s * b1r;
    }else{
      for(int s = 0; s < r; s++){
        double sr = 0., si = 0.;
        for(int q = 0; q < r; q++){
          double wr, wi;
          twiddle(P, r, (long)q * s, sign, &wr, &wi);
          sr += tmp[2 * q] * wr - tmp[2 * q + 1] * wi;
          si += tmp[2 * q] * wi + tmp[2 * q + 1] * wr;
        }
        out[2 * ((size_t)f + (size_t)m * s) + 0] = sr;
        out[2 * ((size_t)f + (size_t)m * s) + 1] = si;
      }
    }
  }
  memcpy(y, out, sizeof(double) * 2 * (size_t)n);
}

static void plan_c2c(const fftplan_t * P, int sign, const double * in, double * out){
  fft_rec(P, P->n, sign < 0 ? -1 : 1, in, 1, out);
}

static void plan_r2c(const fftplan_t * P, const double * in, double * x, double * out){
  const int n = P->n;
  for(int i = 0; i < n; i++){ x[2 * i] = in[i]; x[2 * i + 1] = 0.; }
  fft_rec(P, n, -1, x, 1, out);
}

/* c2r as FFTW_BACKWARD (sign +1) on Hermitian input: imaginary parts of the DC and (for
 * even n) Nyquist bins are not part of a Hermitian spectrum and are ignored */
static void plan_c2r(const fftplan_t * P, const double * in, double * x, double * y, double * out){
  const int n = P->n;
  const int h = n / 2 + 1;
  for(int f = 0; f < h; f++){
    x[2 * f] = in[2 * f];
    x[2 * f + 1] = in[2 * f + 1];
  }
  x[1] = 0.;
  if(n % 2 == 0) x[2 * (n / 2) + 1] = 0.;
  for(int f = h; f < n; f++){
    x[2 * f] = in[2 * (n - f)];
    x[2 * f + 1] = - in[2 * (n - f) + 1];
  }
  fft_rec(P, n, +1, x, 1, y);
  for(int i = 0; i < n; i++) out[i] = y[2 * i];
}

void orc_fft_c2c(int n, int sign, const double * in, double * out){
  fftplan_t P;
  plan_init(&P, n);
  plan_c2c(&P, sign, in, out);
  plan_free(&P);
}

void orc_fft_r2c(int n, const double * in, double * out){
  fftplan_t P;
  plan_init(&P, n);
  double * x = malloc(sizeof(double) * 2 * (size_t)n);
  double * y = malloc(sizeof(double) * 2 * (size_t)n);
  plan_r2c(&P, in, x, y);
  memcpy(out, y, sizeof(double) * 2 * (size_t)(n / 2 + 1));
  free(x);
  free(y);
  plan_free(&P);
}

void orc_fft_c2r(int n, const double * in, double * out){
  fftplan_t P;
  plan_init(&P, n);
  double * x = malloc(sizeof(double) * 2 * (size_t)n);
  double * y = malloc(sizeof(double) * 2 * (size_t)n);
  plan_c2r(&P, in, x, y, out);
  free(x);
  free(y);
  plan_free(&P);
}

/* ---------------------------------------------------------------------------------------
 * Thomas algorithm, src/tdm.c:17-50 (gtsv) and :66-106 (periodic, Sherman-Morrison)
 * ------------------------------------------------------------------------------------- */
static void gtsv_r(int n, const double * l, const double * c, const double * u,
    double * v, double * q, size_t s){
  v[0] = u[0] / c[0];
  q[0] = q[0] / c[0];
  for(int i = 1; i < n - 1; i++){
    const double val = 1. / (c[i] - l[i] * v[i - 1]);
    v[i] = val * (u[i]);
    q[s * i] = val * (q[s * i] - l[i] * q[s * (i - 1)]);
  }
  const double val = c[n - 1] - l[n - 1] * v[n - 2];
  if(fabs(val) > DBL_EPSILON){
    q[s * (n - 1)] = 1. / val * (q[s * (n - 1)] - l[n - 1] * q[s * (n - 2)]);
  }else{
    q[s * (n - 1)] = 0.;
  }
  for(int i = n - 2; i >= 0; i--){
    q[s * i] -= v[i] * q[s * (i + 1)];
  }
}

/* complex right-hand side; real and imaginary parts see the same real operations
 * (C99 real * complex multiplies component-wise) */
static void gtsv_c(int n, const double * l, const double * c, const double * u,
    double * v, double * q){
  v[0] = u[0] / c[0];
  q[0] = q[0] / c[0];
  q[1] = q[1] / c[0];
  for(int i = 1; i < n - 1; i++){
    const double val = 1. / (c[i] - l[i] * v[i - 1]);
    v[i] = val * (u[i]);
    q[2 * i + 0] = val * (q[2 * i + 0] - l[i] * q[2 * (i - 1) + 0]);
    q[2 * i + 1] = val * (q[2 * i + 1] - l[i] * q[2 * (i - 1) + 1]);
  }
  const double val = c[n - 1] - l[n - 1] * v[n - 2];
  if(fabs(val) > DBL_EPSILON){
    q[2 * (n - 1) + 0] = 1. / val * (q[2 * (n - 1) + 0] - l[n - 1] * q[2 * (n - 2) + 0]);
    q[2 * (n - 1) + 1] = 1. / val * (q[2 * (n - 1) + 1] - l[n - 1] * q[2 * (n - 2) + 1]);
  }else{
    q[2 * (n - 1) + 0] = 0.;
    q[2 * (n - 1) + 1] = 0.;
  }
  for(int i = n - 2; i >= 0; i--){
    q[2 * i + 0] -= v[i] * q[2 * (i + 1) + 0];
    q[2 * i + 1] -= v[i] * q[2 * (i + 1) + 1];
  }
}

/* periodic system on a strided real line (tdm.c:78-99); q1 and v are scratch of size n */
static void tdm_periodic_r(int n, const double * l, const double * c, const double * u,
    double * v, double * q1, double * q, size_t s){
  for(int i = 0; i < n - 1; i++){
    q1[i] = i == 0 ? -l[i] : i == n - 2 ? -u[i] : 0.;
  }
  gtsv_r(n - 1, l, c, u, v, q1, 1);
  gtsv_r(n - 1, l, c, u, v, q, s);
  const double num = q[s * (n - 1)] - u[n - 1] * q[0] - l[n - 1] * q[s * (n - 2)];
  const double den = c[n - 1] + u[n - 1] * q1[0] + l[n - 1] * q1[n - 2];
  q[s * (n - 1)] = fabs(den) < DBL_EPSILON ? 0. : num / den;
  for(int i = 0; i < n - 1; i++){
    q[s * i] = q[s * i] + q[s * (n - 1)] * q1[i];
  }
}

int orc_gtsv(int n, const double * l, const double * c, const double * u, double * q){
  double * v = zalloc((size_t)n, sizeof(double));
  gtsv_r(n, l, c, u, v, q, 1);
  free(v);
  return 0;
}

int orc_gtsv_complex(int n, const double * l, const double * c, const double * u, double * q){
  double * v = zalloc((size_t)n, sizeof(double));
  gtsv_c(n, l, c, u, v, q);
  free(v);
  return 0;
}

int orc_tdm_periodic(int n, const double * l, const double * c, const double * u, double * q){
  double * v = zalloc((size_t)n, sizeof(double));
  double * q1 = zalloc((size_t)n, sizeof(double));
  tdm_periodic_r(n, l, c, u, v, q1, q, 1);
  free(q1);
  free(v);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * periodic halos, src/halo.c:18-142: y planes over the full x extent and every k, then z
 * planes over the full (x, y) extent.  isz = stored x extent of the array.
 * ------------------------------------------------------------------------------------- */
static void halo_yz(const orc_t * o, double * a, int isz){
  const size_t row = (size_t)isz;
  const size_t plane = row * (size_t)(NY + 2);
  #pragma omp parallel for schedule(static)
  for(int k = 0; k <= NZ + 1; k++){
    double * base = a + plane * (size_t)k;
    memcpy(base + row * 0, base + row * (size_t)NY, sizeof(double) * row);
    memcpy(base + row * (size_t)(NY + 1), base + row * 1, sizeof(double) * row);
  }
  memcpy(a + plane * 0, a + plane * (size_t)NZ, sizeof(double) * plane);
  memcpy(a + plane * (size_t)(NZ + 1), a + plane * 1, sizeof(double) * plane);
}

/* ---------------------------------------------------------------------------------------
 * boundary conditions, src/fluid/boundary/*.c
 * ------------------------------------------------------------------------------------- */
static void bc_ux(orc_t * o){
  KLOOP for(int j = 1; j <= NY; j++){ UX(1, j, k) = 0.; UX(NX + 1, j, k) = 0.; }
  halo_yz(o, o->ux, NX + 1);
}
static void bc_uy(orc_t * o){
  KLOOP for(int j = 1; j <= NY; j++){ UY(0, j, k) = o->c.uy_xm; UY(NX + 1, j, k) = o->c.uy_xp; }
  halo_yz(o, o->uy, NX + 2);
}
static void bc_uz(orc_t * o){
  KLOOP for(int j = 1; j <= NY; j++){ UZ(0, j, k) = 0.; UZ(NX + 1, j, k) = 0.; }
  halo_yz(o, o->uz, NX + 2);
}
static void bc_neumann(orc_t * o, double * a){
  KLOOP for(int j = 1; j <= NY; j++){
    a[IU(0, j, k)] = a[IU(1, j, k)];
    a[IU(NX + 1, j, k)] = a[IU(NX, j, k)];
  }
  halo_yz(o, a, NX + 2);
}
static void bc_t(orc_t * o){
  KLOOP for(int j = 1; j <= NY; j++){ T(0, j, k) = o->c.t_xm; T(NX + 1, j, k) = o->c.t_xp; }
  halo_yz(o, o->t, NX + 2);
}

int orc_update_boundaries(orc_t * o){
  bc_ux(o);
  bc_uy(o);
  bc_uz(o);
  bc_neumann(o, o->p);
  bc_t(o);
  bc_neumann(o, o->psi);   /* boundary/psi.c:13-45 */
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * construction: metrics src/domain.c:259-357, laplacians predict/*.c init_laplacians,
 * Poisson tables compute_potential.c:173-211, :311-349
 * ------------------------------------------------------------------------------------- */
orc_t * orc_create(const orc_config_t * cfg, const double * xf, const double * xc){
  if(cfg->nx < 3 || cfg->ny < 3 || cfg->nz < 3) return NULL;
  orc_t * o = zalloc(1, sizeof(orc_t));
  o->c = *cfg;
  o->nx = cfg->nx;
  o->ny = cfg->ny;
  o->nz = cfg->nz;
#ifdef _OPENMP
  if(cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
  const int nx = NX, ny = NY, nz = NZ;
  o->xf = zalloc((size_t)nx + 1, sizeof(double));
  o->xc = zalloc((size_t)nx + 2, sizeof(double));
  memcpy(o->xf, xf, sizeof(double) * (size_t)(nx + 1));
  memcpy(o->xc, xc, sizeof(double) * (size_t)(nx + 2));
  o->hxxf = zalloc((size_t)nx + 1, sizeof(double));
  o->hxxc = zalloc((size_t)nx + 2, sizeof(double));
  o->hyxf = zalloc((size_t)nx + 1, sizeof(double));
  o->hyxc = zalloc((size_t)nx + 2, sizeof(double));
  o->jdxf = zalloc((size_t)nx + 1, sizeof(double));
  o->jdxc = zalloc((size_t)nx + 2, sizeof(double));
  const double ly = cfg->lengths[1];
  for(int i = 1; i <= nx + 1; i++) HXXF(i) = XC(i) - XC(i - 1);
  HXXC(0) = 0.5 * XF(2) - 0.5 * XF(1);
  for(int i = 1; i <= nx; i++) HXXC(i) = XF(i + 1) - XF(i);
  HXXC(nx + 1) = 0.5 * XF(nx + 1) - 0.5 * XF(nx);
  for(int i = 1; i <= nx + 1; i++) HYXF(i) = cfg->is_curved ? XF(i) * ly / ny : ly / ny;
  for(int i = 0; i <= nx + 1; i++) HYXC(i) = cfg->is_curved ? XC(i) * ly / ny : ly / ny;
  o->hz = cfg->lengths[2] / nz;
  const double hz = o->hz;
  for(int i = 1; i <= nx + 1; i++) JDXF(i) = HXXF(i) * HYXF(i) * hz;
  for(int i = 0; i <= nx + 1; i++) JDXC(i) = HXXC(i) * HYXC(i) * hz;
  const size_t nu_ = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  const size_t nf_ = (size_t)(nx + 1) * (ny + 2) * (nz + 2);
  const size_t nc_ = (size_t)(nx) * (ny + 2) * (nz + 2);
  o->ux = zalloc(nf_, sizeof(double));
  o->uy = zalloc(nu_, sizeof(double));
  o->uz = zalloc(nu_, sizeof(double));
  o->p = zalloc(nu_, sizeof(double));
  o->psi = zalloc(nu_, sizeof(double));
  o->t = zalloc(nu_, sizeof(double));
  for(int n = 0; n < 3; n++){
    o->srcux[n] = zalloc((size_t)(nx - 1) * ny * nz, sizeof(double));
    o->srcuy[n] = zalloc((size_t)nx * ny * nz, sizeof(double));
    o->srcuz[n] = zalloc((size_t)nx * ny * nz, sizeof(double));
    o->srct[n] = zalloc((size_t)nx * ny * nz, sizeof(double));
  }
  o->lxx = zalloc(nc_, sizeof(double));
  o->lyy0 = zalloc(nc_, sizeof(double));
  o->lyy1 = zalloc(nc_, sizeof(double));
  o->lzy = zalloc(nc_, sizeof(double));
  o->lyz = zalloc(nc_, sizeof(double));
  o->lzz = zalloc(nc_, sizeof(double));
  o->lxy = zalloc(nf_, sizeof(double));
  o->lxz = zalloc(nf_, sizeof(double));
  o->lyx0 = zalloc(nf_, sizeof(double));
  o->lyx1 = zalloc(nf_, sizeof(double));
  o->lzx = zalloc(nf_, sizeof(double));
  o->delta = zalloc((size_t)nx * ny * nz, sizeof(double));
  /* predict/ux.c:34-81 */
  o->lapx_ux = zalloc((size_t)nx - 1, sizeof(lap_t));
  o->lapy_ux = zalloc((size_t)nx - 1, sizeof(lap_t));
  for(int i = 2; i <= nx; i++){
    const double l = 1. / JDXF(i) * JDXC(i - 1) / HXXC(i - 1) / HXXC(i - 1);
    const double u = 1. / JDXF(i) * JDXC(i) / HXXC(i) / HXXC(i);
    o->lapx_ux[i - 2] = (lap_t){l, - l - u, u};
    const double ly_ = 1. / HYXF(i) / HYXF(i);
    const double uy_ = 1. / HYXF(i) / HYXF(i);
    o->lapy_ux[i - 2] = (lap_t){ly_, - ly_ - uy_, uy_};
  }
  /* predict/uy.c:34-82 (identical in uz.c:29-77 and t.c:29-77) */
  o->lapx_c = zalloc((size_t)nx, sizeof(lap_t));
  o->lapy_c = zalloc((size_t)nx, sizeof(lap_t));
  for(int i = 1; i <= nx; i++){
    const double l = 1. / JDXC(i) * JDXF(i) / HXXF(i) / HXXF(i);
    const double u = 1. / JDXC(i) * JDXF(i + 1) / HXXF(i + 1) / HXXF(i + 1);
    o->lapx_c[i - 1] = (lap_t){l, - l - u, u};
    const double ly_ = 1. / HYXC(i) / HYXC(i);
    const double uy_ = 1. / HYXC(i) / HYXC(i);
    o->lapy_c[i - 1] = (lap_t){ly_, - ly_ - uy_, uy_};
  }
  {
    const double l = 1. / hz / hz;
    const double u = 1. / hz / hz;
    o->lapz = (lap_t){l, - l - u, u};
  }
  /* compute_potential.c:173-211, :311-349 */
  const double g_pi = 3.14159265358979324;
  o->ptdm_l = zalloc((size_t)nx, sizeof(double));
  o->ptdm_u = zalloc((size_t)nx, sizeof(double));
  for(int i = 1; i <= nx; i++){
    o->ptdm_l[i - 1] = 1. / JDXC(i) * JDXF(i) / HXXF(i) / HXXF(i);
    o->ptdm_u[i - 1] = 1. / JDXC(i) * JDXF(i + 1) / HXXF(i + 1) / HXXF(i + 1);
  }
  const int nky = ny / 2 + 1;
  o->eval_y = zalloc((size_t)nky, sizeof(double));
  o->eval_z = zalloc((size_t)nz, sizeof(double));
  for(int n = 0; n < nky; n++) o->eval_y[n] = - 4. * pow(sin(g_pi * n / ny), 2.);
  for(int n = 0; n < nz; n++) o->eval_z[n] = - 4. * pow(sin(g_pi * n / nz), 2.);
  o->spec = zalloc((size_t)2 * nx * nky * nz, sizeof(double));
  return o;
}

void orc_destroy(orc_t * o){
  if(!o) return;
  double * bufs[] = {o->xf, o->xc, o->hxxf, o->hxxc, o->hyxf, o->hyxc, o->jdxf, o->jdxc,
    o->ux, o->uy, o->uz, o->p, o->psi, o->t, o->lxx, o->lyx0, o->lyx1, o->lxy, o->lzx, o->lxz,
    o->lyy0, o->lyy1, o->lzy, o->lyz, o->lzz, o->delta, o->eval_y, o->eval_z, o->ptdm_l,
    o->ptdm_u, o->spec};
  for(size_t n = 0; n < sizeof(bufs) / sizeof(bufs[0]); n++) free(bufs[n]);
  for(int n = 0; n < 3; n++){ free(o->srcux[n]); free(o->srcuy[n]); free(o->srcuz[n]); free(o->srct[n]); }
  free(o->lapx_ux); free(o->lapy_ux); free(o->lapx_c); free(o->lapy_c);
  for(int n = 0; n < 8; n++) free(o->st[n]);
  free(o);
}

double orc_hz(const orc_t * o){ return o->hz; }

double * orc_array(orc_t * o, const char * name, size_t * nitems){
  const size_t nx = NX, ny = NY, nz = NZ;
  const size_t nu_ = (nx + 2) * (ny + 2) * (nz + 2);
  const size_t nf_ = (nx + 1) * (ny + 2) * (nz + 2);
  const size_t nc_ = nx * (ny + 2) * (nz + 2);
  struct { const char * s; double * p; size_t n; } tab[] = {
    {"ux", o->ux, nf_}, {"uy", o->uy, nu_}, {"uz", o->uz, nu_}, {"p", o->p, nu_},
    {"psi", o->psi, nu_}, {"t", o->t, nu_},
    {"lxx", o->lxx, nc_}, {"lyy0", o->lyy0, nc_}, {"lyy1", o->lyy1, nc_}, {"lzy", o->lzy, nc_},
    {"lyz", o->lyz, nc_}, {"lzz", o->lzz, nc_}, {"lxy", o->lxy, nf_}, {"lxz", o->lxz, nf_},
    {"lyx0", o->lyx0, nf_}, {"lyx1", o->lyx1, nf_}, {"lzx", o->lzx, nf_},
    {"srcux0", o->srcux[0], (nx - 1) * ny * nz}, {"srcux1", o->srcux[1], (nx - 1) * ny * nz},
    {"srcux2", o->srcux[2], (nx - 1) * ny * nz},
    {"srcuy0", o->srcuy[0], nx * ny * nz}, {"srcuy1", o->srcuy[1], nx * ny * nz}, {"srcuy2", o->srcuy[2], nx * ny * nz},
    {"srcuz0", o->srcuz[0], nx * ny * nz}, {"srcuz1", o->srcuz[1], nx * ny * nz}, {"srcuz2", o->srcuz[2], nx * ny * nz},
    {"srct0", o->srct[0], nx * ny * nz}, {"srct1", o->srct[1], nx * ny * nz}, {"srct2", o->srct[2], nx * ny * nz},
    {"xf", o->xf, nx + 1}, {"xc", o->xc, nx + 2}, {"hxxf", o->hxxf, nx + 1}, {"hxxc", o->hxxc, nx + 2},
    {"hyxf", o->hyxf, nx + 1}, {"hyxc", o->hyxc, nx + 2}, {"jdxf", o->jdxf, nx + 1}, {"jdxc", o->jdxc, nx + 2},
    {"ux1", o->st[0], (nx + 1) * ny * nz}, {"ux2", o->st[1], (nx + 1) * ny * nz},
    {"uy1", o->st[2], (nx + 2) * ny * nz}, {"uy2", o->st[3], (nx + 2) * ny * nz},
    {"uz1", o->st[4], (nx + 2) * ny * nz}, {"uz2", o->st[5], (nx + 2) * ny * nz},
    {"adv", o->st[6], (nx + 1) * ny * nz}, {"dif", o->st[7], (nx + 1) * ny * nz},
  };
  for(size_t n = 0; n < sizeof(tab) / sizeof(tab[0]); n++){
    if(0 == strcmp(tab[n].s, name)){
      if(nitems) *nitems = tab[n].n;
      return tab[n].p;
    }
  }
  return NULL;
}

/* ---------------------------------------------------------------------------------------
 * decide_dt, src/decide_dt.c:32-160
 * ------------------------------------------------------------------------------------- */
int orc_decide_dt(orc_t * o, double * dt){
  if(o->c.dt_fixed > 0.){
    *dt = o->c.dt_fixed;
    return 0;
  }
  const double small = 1.e-8;
  const double hz = o->hz;
  double adv = 1.;
  #pragma omp parallel for reduction(min:adv) schedule(static)
  for(int k = 1; k <= NZ; k++){
    double loc = 1.;
    for(int j = 1; j <= NY; j++){
      for(int i = 2; i <= NX; i++){
        const double vel = fabs(UX(i, j, k)) + small;
        loc = fmin(loc, HXXF(i) / vel);
      }
      for(int i = 1; i <= NX; i++){
        const double vel = fabs(UY(i, j, k)) + small;
        loc = fmin(loc, HYXC(i) / vel);
      }
      for(int i = 1; i <= NX; i++){
        const double vel = fabs(UZ(i, j, k)) + small;
        loc = fmin(loc, hz / vel);
      }
    }
    adv = fmin(adv, loc);
  }
  adv *= o->c.coef_dt_adv;
  const double diffusivity = fmax(nu_of(o), kappa_of(o));
  double grid[3];
  grid[0] = DBL_MAX;
  for(int i = 0; i <= NX + 1; i++) grid[0] = fmin(grid[0], HXXC(i));
  grid[1] = HYXC(1);
  grid[2] = hz;
  double dif[3];
  for(int d = 0; d < 3; d++){
    dif[d] = o->c.coef_dt_dif / diffusivity * 0.5 / 3 * pow(grid[d], 2.);
  }
  *dt = adv;
  for(int d = 0; d < 3; d++){
    if(!o->c.implicit[d]) *dt = fmin(*dt, dif[d]);
  }
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * velocity-gradient tensor, src/fluid/predict/l*.c
 * ------------------------------------------------------------------------------------- */
int orc_compute_tensors(orc_t * o){
  const double hz = o->hz;
  /* lxx.c:19-49 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double hx = HXXC(i);
    LXX(i, j, k) = 1. / hx * (- UX(i, j, k) + UX(i + 1, j, k));
  }
  halo_yz(o, o->lxx, NX);
  /* lxy.c:19-49 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double hx = HXXF(i);
    LXY(i, j, k) = 1. / hx * (- UY(i - 1, j, k) + UY(i, j, k));
  }
  halo_yz(o, o->lxy, NX + 1);
  /* lxz.c:10-44 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double hx = HXXF(i);
    LXZ(i, j, k) = 1. / hx * (- UZ(i - 1, j, k) + UZ(i, j, k));
  }
  halo_yz(o, o->lxz, NX + 1);
  /* lyx.c:24-54 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double hy = HYXF(i);
    LYX0(i, j, k) = 1. / hy * (- UX(i, j - 1, k) + UX(i, j, k));
  }
  halo_yz(o, o->lyx0, NX + 1);
  /* lyx.c:56-95 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double hx_xm = HXXC(i - 1);
    const double hx_xp = HXXC(i);
    const double jd_xm = JDXC(i - 1);
    const double jd_x0 = JDXF(i);
    const double jd_xp = JDXC(i);
    const double jdhx_xm = jd_xm / hx_xm;
    const double jdhx_xp = jd_xp / hx_xp;
    const double djdhx = - jdhx_xm + jdhx_xp;
    LYX1(i, j, k) = - 1. / jd_x0 * djdhx * (
        + (1 == i ? 0. : 0.5) * UY(i - 1, j, k)
        + (NX + 1 == i ? 0. : 0.5) * UY(i, j, k));
  }
  halo_yz(o, o->lyx1, NX + 1);
  /* lyy.c:24-54 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double hy = HYXC(i);
    LYY0(i, j, k) = 1. / hy * (- UY(i, j, k) + UY(i, j + 1, k));
  }
  halo_yz(o, o->lyy0, NX);
  /* lyy.c:56-95 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double hx_xm = HXXF(i);
    const double hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i);
    const double jd_x0 = JDXC(i);
    const double jd_xp = JDXF(i + 1);
    const double jdhx_xm = jd_xm / hx_xm;
    const double jdhx_xp = jd_xp / hx_xp;
    const double djdhx = - jdhx_xm + jdhx_xp;
    LYY1(i, j, k) = 1. / jd_x0 * djdhx * (+ 0.5 * UX(i, j, k) + 0.5 * UX(i + 1, j, k));
  }
  halo_yz(o, o->lyy1, NX);
  /* lyz.c:10-44 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double hy = HYXC(i);
    LYZ(i, j, k) = 1. / hy * (- UZ(i, j - 1, k) + UZ(i, j, k));
  }
  halo_yz(o, o->lyz, NX);
  /* lzx.c:9-42 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    LZX(i, j, k) = 1. / hz * (- UX(i, j, k - 1) + UX(i, j, k));
  }
  halo_yz(o, o->lzx, NX + 1);
  /* lzy.c:9-42 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    LZY(i, j, k) = 1. / hz * (- UY(i, j, k - 1) + UY(i, j, k));
  }
  halo_yz(o, o->lzy, NX);
  /* lzz.c:9-42 */
  KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    LZZ(i, j, k) = 1. / hz * (- UZ(i, j, k) + UZ(i, j, k + 1));
  }
  halo_yz(o, o->lzz, NX);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * right-hand sides; one sweep per term, in the order of compute_rhs_* (ux.c:402-439,
 * uy.c:427-465, uz.c:326-357, t.c:247-272)
 * ------------------------------------------------------------------------------------- */
#define XLOOP KLOOP for(int j = 1; j <= NY; j++) for(int i = 2; i <= NX; i++)
#define CLOOP KLOOP for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++)

static void rhs_ux(orc_t * o){
  const double hz = o->hz;
  const double diffusivity = nu_of(o);
  double * srca = o->srcux[0];
  double * srcg = o->srcux[2];
  /* advection_x ux.c:92-124 */
  XLOOP {
    const double hx_xm = HXXF(i - 1), hx_x0 = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i - 1), jd_x0 = JDXF(i), jd_xp = JDXF(i + 1);
    const double ux_xm = + 0.5 * jd_xm / hx_xm * UX(i - 1, j, k) + 0.5 * jd_x0 / hx_x0 * UX(i, j, k);
    const double ux_xp = + 0.5 * jd_x0 / hx_x0 * UX(i, j, k) + 0.5 * jd_xp / hx_xp * UX(i + 1, j, k);
    const double l = - 0.5 * ux_xm;
    const double u = + 0.5 * ux_xp;
    const double c = - l - u;
    srca[CNTX(i, j, k)] -= 1. / jd_x0 * (+ l * UX(i - 1, j, k) + c * UX(i, j, k) + u * UX(i + 1, j, k));
  }
  /* advection_y ux.c:126-159 */
  XLOOP {
    const double hy_xm = HYXC(i - 1), hy_xp = HYXC(i);
    const double jd_xm = JDXC(i - 1), jd_x0 = JDXF(i), jd_xp = JDXC(i);
    const double uy_ym = + 0.5 * jd_xm / hy_xm * UY(i - 1, j, k) + 0.5 * jd_xp / hy_xp * UY(i, j, k);
    const double uy_yp = + 0.5 * jd_xm / hy_xm * UY(i - 1, j + 1, k) + 0.5 * jd_xp / hy_xp * UY(i, j + 1, k);
    const double l = - 0.5 * uy_ym;
    const double u = + 0.5 * uy_yp;
    const double c = - l - u;
    srca[CNTX(i, j, k)] -= 1. / jd_x0 * (+ l * UX(i, j - 1, k) + c * UX(i, j, k) + u * UX(i, j + 1, k));
  }
  /* advection_z ux.c:161-192 */
  XLOOP {
    const double jd_xm = JDXC(i - 1), jd_x0 = JDXF(i), jd_xp = JDXC(i);
    const double uz_zm = + 0.5 * jd_xm / hz * UZ(i - 1, j, k) + 0.5 * jd_xp / hz * UZ(i, j, k);
    const double uz_zp = + 0.5 * jd_xm / hz * UZ(i - 1, j, k + 1) + 0.5 * jd_xp / hz * UZ(i, j, k + 1);
    const double l = - 0.5 * uz_zm;
    const double u = + 0.5 * uz_zp;
    const double c = - l - u;
    srca[CNTX(i, j, k)] -= 1. / jd_x0 * (+ l * UX(i, j, k - 1) + c * UX(i, j, k) + u * UX(i, j, k + 1));
  }
  /* advection_1 ux.c:194-226 */
  XLOOP {
    const double hx_xm = HXXF(i - 1), hx_x0 = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i - 1), jd_x0 = JDXF(i), jd_xp = JDXF(i + 1);
    const double djdhx_xm = - jd_xm / hx_xm + jd_x0 / hx_x0;
    const double djdhx_xp = - jd_x0 / hx_x0 + jd_xp / hx_xp;
    const double uy_xm = + 0.5 * UY(i - 1, j, k) + 0.5 * UY(i - 1, j + 1, k);
    const double uy_xp = + 0.5 * UY(i, j, k) + 0.5 * UY(i, j + 1, k);
    srca[CNTX(i, j, k)] += 1. / jd_x0 * (+ 0.5 * djdhx_xm * uy_xm * uy_xm + 0.5 * djdhx_xp * uy_xp * uy_xp);
  }
  /* diffusion_x ux.c:228-248 */
  {
    double * src = o->c.implicit[0] ? srcg : srca;
    XLOOP {
      const lap_t lp = o->lapx_ux[i - 2];
      src[CNTX(i, j, k)] += 2. * diffusivity * (+ lp.l * UX(i - 1, j, k) + lp.c * UX(i, j, k) + lp.u * UX(i + 1, j, k));
    }
  }
  /* diffusion_y0 ux.c:250-269 */
  {
    double * src = o->c.implicit[1] ? srcg : srca;
    XLOOP {
      const lap_t lp = o->lapy_ux[i - 2];
      src[CNTX(i, j, k)] += diffusivity * (+ lp.l * UX(i, j - 1, k) + lp.c * UX(i, j, k) + lp.u * UX(i, j + 1, k));
    }
  }
  /* diffusion_y1 ux.c:271-297 */
  XLOOP {
    const double hy = HYXF(i);
    const double jd = JDXF(i);
    const double lyx_ym = LYX1(i, j, k), lyx_yp = LYX1(i, j + 1, k);
    const double lxy_ym = LXY(i, j, k), lxy_yp = LXY(i, j + 1, k);
    srca[CNTX(i, j, k)] += diffusivity / jd * (- jd / hy * (lyx_ym + lxy_ym) + jd / hy * (lyx_yp + lxy_yp));
  }
  /* diffusion_z0 ux.c:299-318 */
  {
    double * src = o->c.implicit[2] ? srcg : srca;
    const lap_t lp = o->lapz;
    XLOOP {
      src[CNTX(i, j, k)] += diffusivity * (+ lp.l * UX(i, j, k - 1) + lp.c * UX(i, j, k) + lp.u * UX(i, j, k + 1));
    }
  }
  /* diffusion_z1 ux.c:320-342 */
  XLOOP {
    const double jd = JDXF(i);
    const double lxz_zm = LXZ(i, j, k), lxz_zp = LXZ(i, j, k + 1);
    srca[CNTX(i, j, k)] += diffusivity / jd * (- jd / hz * lxz_zm + jd / hz * lxz_zp);
  }
  /* diffusion_a ux.c:344-375 */
  XLOOP {
    const double hx_xm = HXXF(i - 1), hx_x0 = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i - 1), jd_x0 = JDXF(i), jd_xp = JDXF(i + 1);
    const double djdhx_xm = - jd_xm / hx_xm + jd_x0 / hx_x0;
    const double djdhx_xp = - jd_x0 / hx_x0 + jd_xp / hx_xp;
    const double lyy_xm = LYY0(i - 1, j, k) + LYY1(i - 1, j, k);
    const double lyy_xp = LYY0(i, j, k) + LYY1(i, j, k);
    srca[CNTX(i, j, k)] -= 2. * diffusivity / jd_x0 * (+ 0.5 * djdhx_xm * lyy_xm + 0.5 * djdhx_xp * lyy_xp);
  }
  /* pressure ux.c:377-394 */
  XLOOP {
    srcg[CNTX(i, j, k)] -= 1. / HXXF(i) * (- P(i - 1, j, k) + P(i, j, k));
  }
}

static void rhs_uy(orc_t * o){
  const double hz = o->hz;
  const double diffusivity = nu_of(o);
  double * srca = o->srcuy[0];
  double * srcg = o->srcuy[2];
  /* advection_x uy.c:93-126 */
  CLOOP {
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double ux_xm = + 0.5 * jd_xm / hx_xm * UX(i, j - 1, k) + 0.5 * jd_xm / hx_xm * UX(i, j, k);
    const double ux_xp = + 0.5 * jd_xp / hx_xp * UX(i + 1, j - 1, k) + 0.5 * jd_xp / hx_xp * UX(i + 1, j, k);
    const double l = - 0.5 * ux_xm;
    const double u = + 0.5 * ux_xp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd_x0 * (+ l * UY(i - 1, j, k) + c * UY(i, j, k) + u * UY(i + 1, j, k));
  }
  /* advection_y uy.c:128-156 */
  CLOOP {
    const double jd = JDXC(i);
    const double hy = HYXC(i);
    const double uy_ym = + 0.5 * jd / hy * UY(i, j - 1, k) + 0.5 * jd / hy * UY(i, j, k);
    const double uy_yp = + 0.5 * jd / hy * UY(i, j, k) + 0.5 * jd / hy * UY(i, j + 1, k);
    const double l = - 0.5 * uy_ym;
    const double u = + 0.5 * uy_yp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd * (+ l * UY(i, j - 1, k) + c * UY(i, j, k) + u * UY(i, j + 1, k));
  }
  /* advection_z uy.c:158-186 */
  CLOOP {
    const double jd = JDXC(i);
    const double uz_zm = + 0.5 * jd / hz * UZ(i, j - 1, k) + 0.5 * jd / hz * UZ(i, j, k);
    const double uz_zp = + 0.5 * jd / hz * UZ(i, j - 1, k + 1) + 0.5 * jd / hz * UZ(i, j, k + 1);
    const double l = - 0.5 * uz_zm;
    const double u = + 0.5 * uz_zp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd * (+ l * UY(i, j, k - 1) + c * UY(i, j, k) + u * UY(i, j, k + 1));
  }
  /* advection_1 uy.c:188-223 */
  CLOOP {
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double djdhx = - jd_xm / hx_xm + jd_xp / hx_xp;
    const double uy_ym = + 0.5 * UY(i, j - 1, k) + 0.5 * UY(i, j, k);
    const double uy_yp = + 0.5 * UY(i, j, k) + 0.5 * UY(i, j + 1, k);
    const double ux_ym = + 0.5 * UX(i, j - 1, k) + 0.5 * UX(i + 1, j - 1, k);
    const double ux_yp = + 0.5 * UX(i, j, k) + 0.5 * UX(i + 1, j, k);
    srca[CNTC(i, j, k)] -= 1. / jd_x0 * (+ 0.5 * djdhx * uy_ym * ux_ym + 0.5 * djdhx * uy_yp * ux_yp);
  }
  /* diffusion_x0 uy.c:225-244 */
  {
    double * src = o->c.implicit[0] ? srcg : srca;
    CLOOP {
      const lap_t lp = o->lapx_c[i - 1];
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * UY(i - 1, j, k) + lp.c * UY(i, j, k) + lp.u * UY(i + 1, j, k));
    }
  }
  /* diffusion_x1 uy.c:246-274 */
  CLOOP {
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double lyx_xm = LYX0(i, j, k) + LYX1(i, j, k);
    const double lyx_xp = LYX0(i + 1, j, k) + LYX1(i + 1, j, k);
    srca[CNTC(i, j, k)] += diffusivity / jd_x0 * (- jd_xm / hx_xm * lyx_xm + jd_xp / hx_xp * lyx_xp);
  }
  /* diffusion_y0 uy.c:276-296 */
  {
    double * src = o->c.implicit[1] ? srcg : srca;
    CLOOP {
      const lap_t lp = o->lapy_c[i - 1];
      src[CNTC(i, j, k)] += 2. * diffusivity * (+ lp.l * UY(i, j - 1, k) + lp.c * UY(i, j, k) + lp.u * UY(i, j + 1, k));
    }
  }
  /* diffusion_y1 uy.c:298-320 */
  CLOOP {
    const double hy = HYXC(i);
    const double jd = JDXC(i);
    srca[CNTC(i, j, k)] += 2. * diffusivity / jd * (- jd / hy * LYY1(i, j - 1, k) + jd / hy * LYY1(i, j, k));
  }
  /* diffusion_z0 uy.c:322-341 */
  {
    double * src = o->c.implicit[2] ? srcg : srca;
    const lap_t lp = o->lapz;
    CLOOP {
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * UY(i, j, k - 1) + lp.c * UY(i, j, k) + lp.u * UY(i, j, k + 1));
    }
  }
  /* diffusion_z1 uy.c:343-365 */
  CLOOP {
    const double jd = JDXC(i);
    const double lyz_zm = LYZ(i, j, k), lyz_zp = LYZ(i, j, k + 1);
    srca[CNTC(i, j, k)] += diffusivity / jd * (- jd / hz * lyz_zm + jd / hz * lyz_zp);
  }
  /* diffusion_a uy.c:367-400 */
  CLOOP {
    const double hx_xm = HXXC(i - 1), hx_x0 = HXXC(i), hx_xp = HXXC(i + 1);
    const double jd_xm = JDXC(i - 1), jd_x0 = JDXC(i), jd_xp = JDXC(i + 1);
    const double djdhx_xm = - jd_xm / hx_xm + jd_x0 / hx_x0;
    const double djdhx_xp = - jd_x0 / hx_x0 + jd_xp / hx_xp;
    const double lyx_xm = LYX0(i, j, k) + LYX1(i, j, k);
    const double lyx_xp = LYX0(i + 1, j, k) + LYX1(i + 1, j, k);
    const double lxy_xm = LXY(i, j, k);
    const double lxy_xp = LXY(i + 1, j, k);
    srca[CNTC(i, j, k)] += diffusivity / jd_x0 * (+ 0.5 * djdhx_xm * (lyx_xm + lxy_xm) + 0.5 * djdhx_xp * (lyx_xp + lxy_xp));
  }
  /* pressure uy.c:402-419 */
  CLOOP {
    srcg[CNTC(i, j, k)] -= 1. / HYXC(i) * (- P(i, j - 1, k) + P(i, j, k));
  }
}

static void rhs_uz(orc_t * o){
  const double hz = o->hz;
  const double diffusivity = nu_of(o);
  double * srca = o->srcuz[0];
  double * srcg = o->srcuz[2];
  /* advection_x uz.c:88-121 */
  CLOOP {
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double ux_xm = + 0.5 * jd_xm / hx_xm * UX(i, j, k - 1) + 0.5 * jd_xm / hx_xm * UX(i, j, k);
    const double ux_xp = + 0.5 * jd_xp / hx_xp * UX(i + 1, j, k - 1) + 0.5 * jd_xp / hx_xp * UX(i + 1, j, k);
    const double l = - 0.5 * ux_xm;
    const double u = + 0.5 * ux_xp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd_x0 * (+ l * UZ(i - 1, j, k) + c * UZ(i, j, k) + u * UZ(i + 1, j, k));
  }
  /* advection_y uz.c:123-152 */
  CLOOP {
    const double hy = HYXC(i);
    const double jd = JDXC(i);
    const double uy_ym = + 0.5 * jd / hy * UY(i, j, k - 1) + 0.5 * jd / hy * UY(i, j, k);
    const double uy_yp = + 0.5 * jd / hy * UY(i, j + 1, k - 1) + 0.5 * jd / hy * UY(i, j + 1, k);
    const double l = - 0.5 * uy_ym;
    const double u = + 0.5 * uy_yp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd * (+ l * UZ(i, j - 1, k) + c * UZ(i, j, k) + u * UZ(i, j + 1, k));
  }
  /* advection_z uz.c:154-181 */
  CLOOP {
    const double jd = JDXC(i);
    const double uz_zm = + 0.5 * jd / hz * UZ(i, j, k - 1) + 0.5 * jd / hz * UZ(i, j, k);
    const double uz_zp = + 0.5 * jd / hz * UZ(i, j, k) + 0.5 * jd / hz * UZ(i, j, k + 1);
    const double l = - 0.5 * uz_zm;
    const double u = + 0.5 * uz_zp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd * (+ l * UZ(i, j, k - 1) + c * UZ(i, j, k) + u * UZ(i, j, k + 1));
  }
  /* diffusion_x0 uz.c:183-202 */
  {
    double * src = o->c.implicit[0] ? srcg : srca;
    CLOOP {
      const lap_t lp = o->lapx_c[i - 1];
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * UZ(i - 1, j, k) + lp.c * UZ(i, j, k) + lp.u * UZ(i + 1, j, k));
    }
  }
  /* diffusion_x1 uz.c:204-231 */
  CLOOP {
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double lzx_xm = LZX(i, j, k), lzx_xp = LZX(i + 1, j, k);
    srca[CNTC(i, j, k)] += diffusivity / jd_x0 * (- jd_xm / hx_xm * lzx_xm + jd_xp / hx_xp * lzx_xp);
  }
  /* diffusion_y0 uz.c:233-252 */
  {
    double * src = o->c.implicit[1] ? srcg : srca;
    CLOOP {
      const lap_t lp = o->lapy_c[i - 1];
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * UZ(i, j - 1, k) + lp.c * UZ(i, j, k) + lp.u * UZ(i, j + 1, k));
    }
  }
  /* diffusion_y1 uz.c:254-277 */
  CLOOP {
    const double hy = HYXC(i);
    const double jd = JDXC(i);
    const double lzy_ym = LZY(i, j, k), lzy_yp = LZY(i, j + 1, k);
    srca[CNTC(i, j, k)] += diffusivity / jd * (- jd / hy * lzy_ym + jd / hy * lzy_yp);
  }
  /* diffusion_z uz.c:279-299 */
  {
    double * src = o->c.implicit[2] ? srcg : srca;
    const lap_t lp = o->lapz;
    CLOOP {
      src[CNTC(i, j, k)] += 2. * diffusivity * (+ lp.l * UZ(i, j, k - 1) + lp.c * UZ(i, j, k) + lp.u * UZ(i, j, k + 1));
    }
  }
  /* pressure uz.c:301-318 */
  CLOOP {
    srcg[CNTC(i, j, k)] -= 1. / hz * (- P(i, j, k - 1) + P(i, j, k));
  }
}

static void rhs_t(orc_t * o){
  const double hz = o->hz;
  const double diffusivity = kappa_of(o);
  double * srca = o->srct[0];
  double * srcg = o->srct[2];
  /* advection_x t.c:88-119 */
  CLOOP {
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double ux_xm = jd_xm / hx_xm * UX(i, j, k);
    const double ux_xp = jd_xp / hx_xp * UX(i + 1, j, k);
    const double l = - 0.5 * ux_xm;
    const double u = + 0.5 * ux_xp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd_x0 * (+ l * T(i - 1, j, k) + c * T(i, j, k) + u * T(i + 1, j, k));
  }
  /* advection_y t.c:121-148 */
  CLOOP {
    const double jd = JDXC(i);
    const double hy = HYXC(i);
    const double uy_ym = jd / hy * UY(i, j, k);
    const double uy_yp = jd / hy * UY(i, j + 1, k);
    const double l = - 0.5 * uy_ym;
    const double u = + 0.5 * uy_yp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd * (+ l * T(i, j - 1, k) + c * T(i, j, k) + u * T(i, j + 1, k));
  }
  /* advection_z t.c:150-176 */
  CLOOP {
    const double jd = JDXC(i);
    const double uz_zm = jd / hz * UZ(i, j, k);
    const double uz_zp = jd / hz * UZ(i, j, k + 1);
    const double l = - 0.5 * uz_zm;
    const double u = + 0.5 * uz_zp;
    const double c = - l - u;
    srca[CNTC(i, j, k)] -= 1. / jd * (+ l * T(i, j, k - 1) + c * T(i, j, k) + u * T(i, j, k + 1));
  }
  /* diffusion_x/y/z t.c:178-239 */
  {
    double * src = o->c.implicit[0] ? srcg : srca;
    CLOOP {
      const lap_t lp = o->lapx_c[i - 1];
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * T(i - 1, j, k) + lp.c * T(i, j, k) + lp.u * T(i + 1, j, k));
    }
  }
  {
    double * src = o->c.implicit[1] ? srcg : srca;
    CLOOP {
      const lap_t lp = o->lapy_c[i - 1];
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * T(i, j - 1, k) + lp.c * T(i, j, k) + lp.u * T(i, j + 1, k));
    }
  }
  {
    double * src = o->c.implicit[2] ? srcg : srca;
    const lap_t lp = o->lapz;
    CLOOP {
      src[CNTC(i, j, k)] += diffusivity * (+ lp.l * T(i, j, k - 1) + lp.c * T(i, j, k) + lp.u * T(i, j, k + 1));
    }
  }
}

/* ---------------------------------------------------------------------------------------
 * implicit line solves, predict/linear_system.c:6-103, on the increment stored with the
 * RK-buffer layout (i fastest, no halos); isz = nx-1 (ux) or nx
 * ------------------------------------------------------------------------------------- */
static void solve_x(orc_t * o, double pf, const lap_t * lap, double * d, int isz){
  const size_t nlines = (size_t)NY * NZ;
  #pragma omp parallel
  {
    double * l = zalloc((size_t)isz, sizeof(double));
    double * c = zalloc((size_t)isz, sizeof(double));
    double * u = zalloc((size_t)isz, sizeof(double));
    double * v = zalloc((size_t)isz, sizeof(double));
    #pragma omp for schedule(static)
    for(size_t n = 0; n < nlines; n++){
      for(int i = 0; i < isz; i++){
        l[i] = - pf * lap[i].l;
        c[i] = 1. - pf * lap[i].c;
        u[i] = - pf * lap[i].u;
      }
      gtsv_r(isz, l, c, u, v, d + n * (size_t)isz, 1);
    }
    free(l); free(c); free(u); free(v);
  }
}

static void solve_y(orc_t * o, double pf, const lap_t * lap, double * d, int isz){
  const int n = NY;
  #pragma omp parallel
  {
    double * l = zalloc((size_t)n, sizeof(double));
    double * c = zalloc((size_t)n, sizeof(double));
    double * u = zalloc((size_t)n, sizeof(double));
    double * v = zalloc((size_t)n, sizeof(double));
    double * q1 = zalloc((size_t)n, sizeof(double));
    #pragma omp for schedule(static)
    for(int k = 0; k < NZ; k++){
      for(int i = 0; i < isz; i++){
        for(int j = 0; j < n; j++){
          l[j] = - pf * lap[i].l;
          c[j] = 1. - pf * lap[i].c;
          u[j] = - pf * lap[i].u;
        }
        tdm_periodic_r(n, l, c, u, v, q1, d + (size_t)i + (size_t)isz * (size_t)n * k, (size_t)isz);
      }
    }
    free(l); free(c); free(u); free(v); free(q1);
  }
}

static void solve_z(orc_t * o, double pf, const lap_t * lap, double * d, int isz){
  const int n = NZ;
  const size_t stride = (size_t)isz * NY;
  #pragma omp parallel
  {
    double * l = zalloc((size_t)n, sizeof(double));
    double * c = zalloc((size_t)n, sizeof(double));
    double * u = zalloc((size_t)n, sizeof(double));
    double * v = zalloc((size_t)n, sizeof(double));
    double * q1 = zalloc((size_t)n, sizeof(double));
    #pragma omp for schedule(static)
    for(int j = 0; j < NY; j++){
      for(int i = 0; i < isz; i++){
        for(int k = 0; k < n; k++){
          l[k] = - pf * (*lap).l;
          c[k] = 1. - pf * (*lap).c;
          u[k] = - pf * (*lap).u;
        }
        tdm_periodic_r(n, l, c, u, v, q1, d + (size_t)i + (size_t)isz * j, stride);
      }
    }
    free(l); free(c); free(u); free(v); free(q1);
  }
}

/* one field update, shared by update_{ux,uy,uz,t} (ux.c:449-558, uy.c:475-584,
 * uz.c:367-476, t.c:282-391); fx/fy/fz multiply the prefactor per direction */
static void update_field(orc_t * o, int rkstep, double dt, double diffusivity,
    double * src[3], const lap_t * lapx, const lap_t * lapy, int isz, double fx, double fy, double fz){
  const double coef_a = RKA[rkstep];
  const double coef_b = RKB[rkstep];
  const double coef_g = rk_g(rkstep);
  const size_t nitems = (size_t)isz * NY * NZ;
  double * d = o->delta;
  const double * sa = src[0];
  const double * sb = src[1];
  const double * sg = src[2];
  const double * wa = mut == ORC_MUT_AB_SWAP ? sb : sa;
  const double * wb = mut == ORC_MUT_AB_SWAP ? sa : sb;
  #pragma omp parallel for schedule(static)
  for(size_t n = 0; n < nitems; n++){
    d[n] = + coef_a * dt * wa[n] + coef_b * dt * wb[n] + coef_g * dt * sg[n];
  }
  const double prefactor = 0.5 * (mut == ORC_MUT_PF_ALPHA ? coef_a : rk_g(rkstep)) * dt * diffusivity;
  if(o->c.implicit[0]) solve_x(o, fx * prefactor, lapx, d, isz);
  if(o->c.implicit[1]) solve_y(o, fy * prefactor, lapy, d, isz);
  if(o->c.implicit[2]) solve_z(o, fz * prefactor, &o->lapz, d, isz);
}

/* reset_srcs predict/main.c:8-25 (pointer swap a <-> b, zero a and g) */
static void reset_srcs(double * s[3], size_t nitems){
  if(mut != ORC_MUT_NO_ROTATE){
    double * tmp = s[0];
    s[0] = s[1];
    s[1] = tmp;
  }
  memset(s[0], 0, sizeof(double) * nitems);
  memset(s[2], 0, sizeof(double) * nitems);
}

/* fluid_predict_field, predict/main.c:35-110 */
int orc_predict(orc_t * o, int rkstep, double dt){
  const size_t nx_ = (size_t)(NX - 1) * NY * NZ;
  const size_t nc_ = (size_t)NX * NY * NZ;
  reset_srcs(o->srcux, nx_);
  reset_srcs(o->srcuy, nc_);
  reset_srcs(o->srcuz, nc_);
  reset_srcs(o->srct, nc_);
  orc_compute_tensors(o);
  rhs_ux(o);
  rhs_uy(o);
  rhs_uz(o);
  if(o->c.scalar) rhs_t(o);
  const double nu = nu_of(o);
  /* update_ux: 2 f in x (ux.c:499-506) */
  update_field(o, rkstep, dt, nu, o->srcux, o->lapx_ux, o->lapy_ux, NX - 1,
      mut == ORC_MUT_UX_X_F ? 1. : 2., mut == ORC_MUT_UX_Y_2F ? 2. : 1., 1.);
  XLOOP { UX(i, j, k) += o->delta[CNTX(i, j, k)]; }
  bc_ux(o);
  /* update_uy: 2 f in y (uy.c:539-544) */
  update_field(o, rkstep, dt, nu, o->srcuy, o->lapx_c, o->lapy_c, NX, 1., mut == ORC_MUT_UY_Y_F ? 1. : 2., 1.);
  CLOOP { UY(i, j, k) += o->delta[CNTC(i, j, k)]; }
  bc_uy(o);
  /* update_uz: 2 f in z (uz.c:449-454) */
  update_field(o, rkstep, dt, nu, o->srcuz, o->lapx_c, o->lapy_c, NX,
      mut == ORC_MUT_UZ_X_2F ? 2. : 1., 1., mut == ORC_MUT_UZ_Z_F ? 1. : 2.);
  CLOOP { UZ(i, j, k) += o->delta[CNTC(i, j, k)]; }
  bc_uz(o);
  if(o->c.scalar){
    /* update_t: 2 f in y is a reference quirk (t.c:346-349), kept */
    update_field(o, rkstep, dt, kappa_of(o), o->srct, o->lapx_c, o->lapy_c, NX, 1., mut == ORC_MUT_T_Y_F ? 1. : 2., 1.);
    CLOOP { T(i, j, k) += o->delta[CNTC(i, j, k)]; }
    bc_t(o);
  }
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * Poisson solve, fluid/compute_potential.c:382-583
 * ------------------------------------------------------------------------------------- */
int orc_potential(orc_t * o, int rkstep, double dt){
  const int nx = NX, ny = NY, nz = NZ, nky = ny / 2 + 1;
  const double hz = o->hz;
  double * spec = o->spec;
  /* assign_input :382-425 with the FFT normalisation folded in; y r2c per (i, k) */
  const double norm = 1. * ny * nz;
  const double prefactor = 1. / (rk_g(rkstep) * dt) / norm;
  #pragma omp parallel
  {
    double * line = zalloc((size_t)ny, sizeof(double));
    double * xw = zalloc((size_t)2 * ny, sizeof(double));
    double * out = zalloc((size_t)2 * ny, sizeof(double));
    fftplan_t P;
    plan_init(&P, ny);
    #pragma omp for schedule(static)
    for(int k = 1; k <= nz; k++){
      for(int i = 1; i <= nx; i++){
        for(int j = 1; j <= ny; j++){
          const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
          const double hy = HYXC(i);
          const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
          const double ux_xm = UX(i, j, k), ux_xp = UX(i + 1, j, k);
          const double uy_ym = UY(i, j, k), uy_yp = UY(i, j + 1, k);
          const double uz_zm = UZ(i, j, k), uz_zp = UZ(i, j, k + 1);
          const double div = 1. / jd_x0 * (
              - jd_xm / hx_xm * ux_xm + jd_xp / hx_xp * ux_xp
              - jd_x0 / hy * uy_ym + jd_x0 / hy * uy_yp
              - jd_x0 / hz * uz_zm + jd_x0 / hz * uz_zp);
          line[j - 1] = prefactor * div;
        }
        plan_r2c(&P, line, xw, out);
        for(int f = 0; f < nky; f++){
          const size_t idx = (size_t)(i - 1) + (size_t)nx * ((size_t)f + (size_t)nky * (k - 1));
          spec[2 * idx + 0] = out[2 * f + 0];
          spec[2 * idx + 1] = out[2 * f + 1];
        }
      }
    }
    plan_free(&P);
    free(line);
    free(xw);
    free(out);
  }
  /* z c2c forward per (i, ky) */
  #pragma omp parallel
  {
    double * in = zalloc((size_t)2 * nz, sizeof(double));
    double * out = zalloc((size_t)2 * nz, sizeof(double));
    fftplan_t P;
    plan_init(&P, nz);
    #pragma omp for schedule(static)
    for(int f = 0; f < nky; f++){
      for(int i = 0; i < nx; i++){
        for(int k = 0; k < nz; k++){
          const size_t idx = (size_t)i + (size_t)nx * ((size_t)f + (size_t)nky * k);
          in[2 * k] = spec[2 * idx];
          in[2 * k + 1] = spec[2 * idx + 1];
        }
        plan_c2c(&P, -1, in, out);
        for(int k = 0; k < nz; k++){
          const size_t idx = (size_t)i + (size_t)nx * ((size_t)f + (size_t)nky * k);
          spec[2 * idx] = out[2 * k];
          spec[2 * idx + 1] = out[2 * k + 1];
        }
      }
    }
    plan_free(&P);
    free(in);
    free(out);
  }
  /* solve_linear_systems :445-488 */
  #pragma omp parallel
  {
    double * c = zalloc((size_t)nx, sizeof(double));
    double * v = zalloc((size_t)nx, sizeof(double));
    #pragma omp for schedule(static)
    for(int k = 0; k < nz; k++){
      const double eval_z = o->eval_z[k];
      for(int j = 0; j < nky; j++){
        const double eval_y = o->eval_y[j];
        for(int i = 1; i <= nx; i++){
          c[i - 1] = - o->ptdm_l[i - 1] - o->ptdm_u[i - 1] + eval_y / HYXC(i) / HYXC(i) + eval_z / hz / hz;
        }
        c[0] += o->ptdm_l[0];
        c[nx - 1] += o->ptdm_u[nx - 1];
        gtsv_c(nx, o->ptdm_l, c, o->ptdm_u, v, spec + 2 * (size_t)nx * ((size_t)j + (size_t)nky * k));
      }
    }
    free(c);
    free(v);
  }
  /* z c2c backward */
  #pragma omp parallel
  {
    double * in = zalloc((size_t)2 * nz, sizeof(double));
    double * out = zalloc((size_t)2 * nz, sizeof(double));
    fftplan_t P;
    plan_init(&P, nz);
    #pragma omp for schedule(static)
    for(int f = 0; f < nky; f++){
      for(int i = 0; i < nx; i++){
        for(int k = 0; k < nz; k++){
          const size_t idx = (size_t)i + (size_t)nx * ((size_t)f + (size_t)nky * k);
          in[2 * k] = spec[2 * idx];
          in[2 * k + 1] = spec[2 * idx + 1];
        }
        plan_c2c(&P, +1, in, out);
        for(int k = 0; k < nz; k++){
          const size_t idx = (size_t)i + (size_t)nx * ((size_t)f + (size_t)nky * k);
          spec[2 * idx] = out[2 * k];
          spec[2 * idx + 1] = out[2 * k + 1];
        }
      }
    }
    plan_free(&P);
    free(in);
    free(out);
  }
  /* y c2r and extract_output :427-443 */
  #pragma omp parallel
  {
    double * in = zalloc((size_t)2 * nky, sizeof(double));
    double * line = zalloc((size_t)ny, sizeof(double));
    double * xw = zalloc((size_t)2 * ny, sizeof(double));
    double * yw = zalloc((size_t)2 * ny, sizeof(double));
    fftplan_t P;
    plan_init(&P, ny);
    #pragma omp for schedule(static)
    for(int k = 1; k <= nz; k++){
      for(int i = 1; i <= nx; i++){
        for(int f = 0; f < nky; f++){
          const size_t idx = (size_t)(i - 1) + (size_t)nx * ((size_t)f + (size_t)nky * (k - 1));
          in[2 * f] = spec[2 * idx];
          in[2 * f + 1] = spec[2 * idx + 1];
        }
        plan_c2r(&P, in, xw, yw, line);
        for(int j = 1; j <= ny; j++) PSI(i, j, k) = line[j - 1];
      }
    }
    plan_free(&P);
    free(in);
    free(line);
    free(xw);
    free(yw);
  }
  bc_neumann(o, o->psi);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * correction correct/*.c, pressure update update_pressure.c:23-173
 * ------------------------------------------------------------------------------------- */
int orc_correct(orc_t * o, int rkstep, double dt){
  const double prefactor = rk_g(rkstep) * dt;
  const double hz = o->hz;
  XLOOP {
    const double hx = HXXF(i);
    UX(i, j, k) -= prefactor / hx * (- PSI(i - 1, j, k) + PSI(i, j, k));
  }
  bc_ux(o);
  CLOOP {
    const double hy = HYXC(i);
    UY(i, j, k) -= prefactor / hy * (- PSI(i, j - 1, k) + PSI(i, j, k));
  }
  bc_uy(o);
  CLOOP {
    UZ(i, j, k) -= prefactor / hz * (- PSI(i, j, k - 1) + PSI(i, j, k));
  }
  bc_uz(o);
  return 0;
}

int orc_pressure(orc_t * o, int rkstep, double dt){
  const double hz = o->hz;
  CLOOP { P(i, j, k) += PSI(i, j, k); }
  const double prefactor = 0.5 * rk_g(rkstep) * dt * nu_of(o);
  const double two = mut == ORC_MUT_P_F ? 1. : 2.;
  if(o->c.implicit[0]){
    CLOOP {
      const double hx_xm = HXXF(mut == ORC_MUT_P_X_INDEX ? i + 1 : i), hx_xp = HXXF(i + 1);
      const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
      const double l = jd_xm / hx_xm / hx_xm;
      const double u = jd_xp / hx_xp / hx_xp;
      const double c = - l - u;
      P(i, j, k) -= two * prefactor / jd_x0 * (+ l * PSI(i - 1, j, k) + c * PSI(i, j, k) + u * PSI(i + 1, j, k));
    }
  }
  if(o->c.implicit[1]){
    CLOOP {
      const double hy = HYXC(i);
      const double jd = JDXC(i);
      const double l = jd / hy / hy;
      const double u = jd / hy / hy;
      const double c = - l - u;
      P(i, j, k) -= two * prefactor / jd * (+ l * PSI(i, j - 1, k) + c * PSI(i, j, k) + u * PSI(i, j + 1, k));
    }
  }
  if(o->c.implicit[2] && mut != ORC_MUT_P_NO_Z){
    CLOOP {
      const double jd = JDXC(i);
      const double l = jd / hz / hz;
      const double u = jd / hz / hz;
      const double c = - l - u;
      P(i, j, k) -= two * prefactor / jd * (+ l * PSI(i, j, k - 1) + c * PSI(i, j, k) + u * PSI(i, j, k + 1));
    }
  }
  bc_neumann(o, o->p);
  return 0;
}

/* integrate.c:9-35 */
int orc_integrate(orc_t * o, double * dt){
  if(0 != orc_decide_dt(o, dt)) return 1;
  for(int s = 0; s < 3; s++){
    if(0 != orc_predict(o, s, *dt)) return 1;
    if(0 != orc_potential(o, s, *dt)) return 1;
    if(0 != orc_correct(o, s, *dt)) return 1;
    if(0 != orc_pressure(o, s, *dt)) return 1;
  }
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * diagnostics, src/logging/*.c (sums in the reference's loop order, one process)
 * ------------------------------------------------------------------------------------- */
double orc_divergence(const orc_t * o){
  const double hz = o->hz;
  double divmax = 0.;
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double hx_xm = HXXF(i), hx_xp = HXXF(i + 1);
    const double hy = HYXC(i);
    const double jd_xm = JDXF(i), jd_x0 = JDXC(i), jd_xp = JDXF(i + 1);
    const double div = 1. / jd_x0 * (
        - jd_xm / hx_xm * UX(i, j, k) + jd_xp / hx_xp * UX(i + 1, j, k)
        - jd_x0 / hy * UY(i, j, k) + jd_x0 / hy * UY(i, j + 1, k)
        - jd_x0 / hz * UZ(i, j, k) + jd_x0 / hz * UZ(i, j, k + 1));
    divmax = fmax(divmax, fabs(div));
  }
  return divmax;
}

void orc_injection(const orc_t * o, double out[2]){
  double values[2] = {0., 0.};
  const int nx = NX;
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++){
    const double hx_xm = HXXF(1), hx_xp = HXXF(nx + 1);
    const double jd_xm = JDXF(1), jd_xp = JDXF(nx + 1);
    const double tyx_xm = LYX0(1, j, k) + LYX1(1, j, k) + LXY(1, j, k);
    const double tyx_xp = LYX0(nx + 1, j, k) + LYX1(nx + 1, j, k) + LXY(nx + 1, j, k);
    values[0] += - jd_xm / hx_xm * UY(0, j, k) * tyx_xm + jd_xp / hx_xp * UY(nx + 1, j, k) * tyx_xp;
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++){
    const double hx_xm = HXXF(1), hx_xp = HXXF(nx + 1);
    const double jd_xm = JDXF(1), jd_xp = JDXF(nx + 1);
    const double t_xm = T(0, j, k), t_xp = T(nx + 1, j, k);
    const double dt_xm = - T(0, j, k) + T(1, j, k);
    const double dt_xp = - T(nx, j, k) + T(nx + 1, j, k);
    values[1] += - jd_xm / hx_xm * t_xm / hx_xm * dt_xm + jd_xp / hx_xp * t_xp / hx_xp * dt_xp;
  }
  out[0] = nu_of(o) * values[0];
  out[1] = kappa_of(o) * values[1];
}

void orc_dissipation(const orc_t * o, double out[2]){
  double v0 = 0., v1 = 0.;
  const double hz = o->hz;
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double lij = LXX(i, j, k);
    v0 += JDXC(i) * lij * (lij + lij);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double lij = LYX0(i, j, k) + LYX1(i, j, k);
    const double lji = LXY(i, j, k);
    v0 += JDXF(i) * lij * (lij + lji);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double lij = LZX(i, j, k);
    const double lji = LXZ(i, j, k);
    v0 += JDXF(i) * lij * (lij + lji);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double lij = LXY(i, j, k);
    const double lji = LYX0(i, j, k) + LYX1(i, j, k);
    v0 += JDXF(i) * lij * (lij + lji);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double lij = LYY0(i, j, k) + LYY1(i, j, k);
    v0 += JDXC(i) * lij * (lij + lij);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double lij = LZY(i, j, k);
    const double lji = LYZ(i, j, k);
    v0 += JDXC(i) * lij * (lij + lji);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double lij = LXZ(i, j, k);
    const double lji = LZX(i, j, k);
    v0 += JDXF(i) * lij * (lij + lji);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double lij = LYZ(i, j, k);
    const double lji = LZY(i, j, k);
    v0 += JDXC(i) * lij * (lij + lji);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double lij = LZZ(i, j, k);
    v0 += JDXC(i) * lij * (lij + lij);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double d = - T(i - 1, j, k) + T(i, j, k);
    v1 += JDXF(i) * pow(d / HXXF(i), 2.);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double d = - T(i, j - 1, k) + T(i, j, k);
    v1 += JDXC(i) * pow(d / HYXC(i), 2.);
  }
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++){
    const double d = - T(i, j, k - 1) + T(i, j, k);
    v1 += JDXC(i) * pow(d / hz, 2.);
  }
  out[0] = nu_of(o) * v0;
  out[1] = kappa_of(o) * v1;
}

void orc_total_energy(const orc_t * o, double out[4]){
  double q[4] = {0., 0., 0., 0.};
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 2; i <= NX; i++)
    q[0] += JDXF(i) * 0.5 * pow(UX(i, j, k), 2.);
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++)
    q[1] += JDXC(i) * 0.5 * pow(UY(i, j, k), 2.);
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++)
    q[2] += JDXC(i) * 0.5 * pow(UZ(i, j, k), 2.);
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX; i++)
    q[3] += JDXC(i) * 0.5 * pow(T(i, j, k), 2.);
  for(int n = 0; n < 4; n++) out[n] = q[n];
}

/* ---------------------------------------------------------------------------------------
 * statistics, src/statistics.c:111-255: one sample of the current field; the dif term reads
 * the tensor arrays as the last predict left them
 * ------------------------------------------------------------------------------------- */
#define SF(i, j, k) ((size_t)((i) - 1) + (size_t)(NX + 1) * ((size_t)((j) - 1) + (size_t)NY * (size_t)((k) - 1)))
#define SC(i, j, k) ((size_t)(i) + (size_t)(NX + 2) * ((size_t)((j) - 1) + (size_t)NY * (size_t)((k) - 1)))

size_t orc_stat_collect(orc_t * o){
  if(!o->st[0]){
    const size_t nf = (size_t)(NX + 1) * NY * NZ, nc = (size_t)(NX + 2) * NY * NZ;
    const size_t n[8] = {nf, nf, nc, nc, nc, nc, nf, nf};
    for(int q = 0; q < 8; q++) o->st[q] = calloc(n[q], sizeof(double));
    o->st_num = 0;
  }
  double * ux1 = o->st[0], * ux2 = o->st[1], * uy1 = o->st[2], * uy2 = o->st[3];
  double * uz1 = o->st[4], * uz2 = o->st[5], * adv = o->st[6], * dif = o->st[7];
  /* collect_mean_ux, :117-136 */
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    ux1[SF(i, j, k)] += pow(UX(i, j, k), 1.);
    ux2[SF(i, j, k)] += pow(UX(i, j, k), 2.);
  }
  /* collect_mean_uy, :143-162 */
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 0; i <= NX + 1; i++){
    uy1[SC(i, j, k)] += pow(UY(i, j, k), 1.);
    uy2[SC(i, j, k)] += pow(UY(i, j, k), 2.);
  }
  /* collect_mean_uz, :169-188 */
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 0; i <= NX + 1; i++){
    uz1[SC(i, j, k)] += pow(UZ(i, j, k), 1.);
    uz2[SC(i, j, k)] += pow(UZ(i, j, k), 2.);
  }
  /* collect_adv, :190-207 */
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    const double x = 0.5 * UX(i, j - 1, k) + 0.5 * UX(i, j, k);
    const double y = 0.5 * UY(i - 1, j, k) + 0.5 * UY(i, j, k);
    adv[SF(i, j, k)] += x * y;
  }
  /* collect_dif, :209-226 */
  for(int k = 1; k <= NZ; k++) for(int j = 1; j <= NY; j++) for(int i = 1; i <= NX + 1; i++){
    dif[SF(i, j, k)] += LYX0(i, j, k) + LYX1(i, j, k) + LXY(i, j, k);
  }
  o->st_num += 1;
  return o->st_num;
}

