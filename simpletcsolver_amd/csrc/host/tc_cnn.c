/*
 * tc_cnn.c -- host side of the learned subgrid closure (kernels in tck_cnn.hip).
 *
 * A tc_cnn_t holds the network (FP32 master weights, BF16 copies, Adam moments) and the
 * activation buffers of one batch.  Training data arrive from the host as FP32 channels-last
 * arrays (the caller normalises them); the library pads the inputs to 16 channels and rounds
 * them to BF16.  tc_closure_attach couples a trained network to a solver: at the start of
 * every step the solver evaluates the network on its current fields and the predictor adds
 * the resulting corrections to the explicit right-hand sides of ux, uy, uz and T.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <hip/hip_runtime_api.h>
#include "tcstep.h"
#include "tc_internal.h"

/* tck_cnn.hip */
int tck_cnn_launch_conv(int nx, int ny, int nz, int nb, const unsigned short * in, int ci,
    const unsigned short * w, int co, const float * bias, int co_store, int relu,
    unsigned short * out_bf, float * out_f, hipStream_t stream);
int tck_cnn_launch_wgrad(int nx, int ny, int nz, int nb, const unsigned short * in, int ci_n,
    const unsigned short * g, int co_n, float * part, int nblk, float * out, hipStream_t stream);
int tck_cnn_launch_bgrad(long long nv, const float * g, int co_n, float * part, int nblk, float * out,
    hipStream_t stream);
int tck_cnn_launch_mse(const float * pred, const float * target, long long n, float scale,
    float * grad, float * part, int nblk, hipStream_t stream);
int tck_cnn_launch_relu_bwd(const float * gpost, const unsigned short * act, long long n,
    unsigned short * gpre, float * gf, hipStream_t stream);
int tck_cnn_launch_to_bf16(const float * x, long long n, unsigned short * y, hipStream_t stream);
int tck_cnn_launch_flip(const unsigned short * w, int co_n, int ci_n, int ci_p, int co_p,
    unsigned short * wt, hipStream_t stream);
int tck_cnn_launch_adam(float * w, float * m, float * v, const float * g, int n, float lr,
    float b1, float b2, float eps, float c1, float c2, unsigned short * wbf, hipStream_t stream);
int tck_cnn_launch_gather(const tc_geom_t * g, const double * ux, const double * uy,
    const double * uz, const double * t, const double * xc, const float mean[5],
    const float inv_std[5], unsigned short * out, hipStream_t stream);
int tck_cnn_launch_scatter(const tc_geom_t * g, const float * pred, const float scale[4],
    float * cx, float * cy, float * cz, float * ct, hipStream_t stream);

#define CNN_C0 16      /* padded input channels */
#define CNN_C 32       /* hidden channels */
#define CNN_CO 16      /* padded output channels (4 used) */
#define CNN_NBLK 256   /* blocks of the weight-gradient reductions */

/* layer l: input channels ci, output channels co */
static const int g_ci[3] = {CNN_C0, CNN_C, CNN_C};
static const int g_co[3] = {CNN_C, CNN_C, CNN_CO};

struct tc_cnn_s {
  int nx, ny, nz, nb;
  tc_solver_t * owner;          /* the solver it is coupled to (its activations are that
                                 * solver's scratch), or null */
  long long nv;
  hipStream_t stream;
  int device;
  /* parameters: FP32 masters, BF16 copies, Adam moments, gradients */
  float * w[3]; float * b[3];
  unsigned short * wbf[3];
  float * mw[3]; float * vw[3]; float * mb[3]; float * vb[3];
  float * gw[3]; float * gb[3];
  unsigned short * wt[3];       /* flipped kernels of the backward data pass */
  long long step;
  /* activations of one batch */
  unsigned short * x0;          /* [nv][16] */
  unsigned short * a1; unsigned short * a2;   /* [nv][32] */
  float * out;                  /* [nv][16] */
  float * target;               /* [nv][16] */
  float * gout;                 /* d loss / d out */
  unsigned short * g3;          /* BF16 of gout */
  float * d2; float * d1;       /* d loss / d a2, d a1 (FP32) */
  unsigned short * g2; unsigned short * g1;   /* through the ReLUs, BF16 */
  float * g2f; float * g1f;
  float * part;                 /* reduction scratch */
  unsigned short * h_x;         /* host staging */
  float * h_y;
};

static long long wsize(int l){ return 27LL * g_co[l] * g_ci[l]; }

static int cnn_alloc(void ** p, size_t bytes){
  if(hipMalloc(p, bytes ? bytes : 8) != hipSuccess) return tc_fail("tc_cnn: hipMalloc of %zu bytes failed", bytes);
  if(hipMemset(*p, 0, bytes ? bytes : 8) != hipSuccess) return tc_fail("tc_cnn: hipMemset failed");
  return 0;
}

/* host FP32 -> BF16 (round to nearest even) */
static unsigned short bf16(float x){
  unsigned int u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

int tc_cnn_destroy(tc_cnn_t * c){
  if(!c) return 0;
  hipSetDevice(c->device);
  if(c->owner){
    /* the coupled solver's kernels read the network's buffers on the solver's stream (and
     * an ensemble's replays are ordered before that stream's later work): let them finish,
     * then uncouple; the generation bump makes recorded graphs re-record without it */
    if(c->owner->stream) hipStreamSynchronize(c->owner->stream);
    c->owner->closure = NULL;
    c->owner->gen++;
    c->owner = NULL;
  }
  if(c->stream) hipStreamSynchronize(c->stream);
  for(int l = 0; l < 3; l++){
    void * p[] = {c->w[l], c->b[l], c->wbf[l], c->mw[l], c->vw[l], c->mb[l], c->vb[l], c->gw[l], c->gb[l], c->wt[l]};
    for(size_t q = 0; q < sizeof(p) / sizeof(p[0]); q++) if(p[q]) hipFree(p[q]);
  }
  void * a[] = {c->x0, c->a1, c->a2, c->out, c->target, c->gout, c->g3, c->d2, c->d1, c->g2, c->g1,
    c->g2f, c->g1f, c->part};
  for(size_t q = 0; q < sizeof(a) / sizeof(a[0]); q++) if(a[q]) hipFree(a[q]);
  if(c->stream) hipStreamDestroy(c->stream);
  free(c->h_x);
  free(c->h_y);
  free(c);
  return 0;
}

/* weights: He-normal from a seeded generator (Box-Muller over splitmix64), biases zero;
 * output rows beyond the 4 used channels stay zero */
static unsigned long long sm64(unsigned long long * s){
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int tc_cnn_create(int nx, int ny, int nz, int batch, unsigned long long seed, int device, tc_cnn_t ** out){
  *out = NULL;
  if(nx < 1 || ny < 1 || nz < 1 || batch < 1) return tc_fail("tc_cnn_create: bad sizes");
  tc_cnn_t * c = calloc(1, sizeof(*c));
  if(!c) return tc_fail("host allocation failed");
  c->nx = nx; c->ny = ny; c->nz = nz; c->nb = batch; c->device = device;
  c->nv = (long long)nx * ny * nz * batch;
  if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess){
    free(c);
    return tc_fail("tc_cnn_create: stream creation failed");
  }
  int rc = 0;
  for(int l = 0; l < 3 && !rc; l++){
    const size_t nw = (size_t)wsize(l), nbias = (size_t)g_co[l];
    rc |= cnn_alloc((void **)&c->w[l], 4 * nw) || cnn_alloc((void **)&c->b[l], 4 * nbias);
    rc |= cnn_alloc((void **)&c->wbf[l], 2 * nw) || cnn_alloc((void **)&c->wt[l], 2 * nw);
    rc |= cnn_alloc((void **)&c->mw[l], 4 * nw) || cnn_alloc((void **)&c->vw[l], 4 * nw);
    rc |= cnn_alloc((void **)&c->mb[l], 4 * nbias) || cnn_alloc((void **)&c->vb[l], 4 * nbias);
    rc |= cnn_alloc((void **)&c->gw[l], 4 * nw) || cnn_alloc((void **)&c->gb[l], 4 * nbias);
  }
  const size_t nv = (size_t)c->nv;
  if(!rc){
    rc |= cnn_alloc((void **)&c->x0, 2 * nv * CNN_C0);
    rc |= cnn_alloc((void **)&c->a1, 2 * nv * CNN_C) || cnn_alloc((void **)&c->a2, 2 * nv * CNN_C);
    rc |= cnn_alloc((void **)&c->out, 4 * nv * CNN_CO) || cnn_alloc((void **)&c->target, 4 * nv * CNN_CO);
    rc |= cnn_alloc((void **)&c->gout, 4 * nv * CNN_CO) || cnn_alloc((void **)&c->g3, 2 * nv * CNN_CO);
    rc |= cnn_alloc((void **)&c->d2, 4 * nv * CNN_C) || cnn_alloc((void **)&c->d1, 4 * nv * CNN_C);
    rc |= cnn_alloc((void **)&c->g2, 2 * nv * CNN_C) || cnn_alloc((void **)&c->g1, 2 * nv * CNN_C);
    rc |= cnn_alloc((void **)&c->g2f, 4 * nv * CNN_C) || cnn_alloc((void **)&c->g1f, 4 * nv * CNN_C);
    rc |= cnn_alloc((void **)&c->part, 4 * (size_t)CNN_NBLK * 27 * CNN_C * CNN_C);
  }
  c->h_x = calloc(nv * CNN_C0, 2);
  c->h_y = calloc(nv * CNN_CO, 4);
  if(!c->h_x || !c->h_y) rc = tc_fail("host allocation failed");
  /* initial weights */
  unsigned long long st = seed * 0x9E3779B97F4A7C15ULL + 12345;
  for(int l = 0; l < 3 && !rc; l++){
    const long long nw = wsize(l);
    float * h = calloc((size_t)nw, 4);
    if(!h){ rc = tc_fail("host allocation failed"); break; }
    const int fan_in = 27 * (l == 0 ? 5 : CNN_C);
    const double sd = sqrt(2.0 / fan_in) * (l == 2 ? 0.1 : 1.0);
    for(int tap = 0; tap < 27; tap++) for(int co = 0; co < g_co[l]; co++) for(int ci = 0; ci < g_ci[l]; ci++){
      const double u1 = ((sm64(&st) >> 11) + 0.5) * 0x1.0p-53, u2 = (sm64(&st) >> 11) * 0x1.0p-53;
      const double gv = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      const int live = (l == 0 ? ci < 5 : 1) && (l == 2 ? co < 4 : 1);
      h[((size_t)tap * g_co[l] + co) * g_ci[l] + ci] = live ? (float)(sd * gv) : 0.f;
    }
    rc = tc_cnn_set_layer(c, l, h, NULL);
    free(h);
  }
  if(rc){
    tc_cnn_destroy(c);
    return 1;
  }
  *out = c;
  return 0;
}

/* layer weights [27][co][ci] and biases [co] (either may be null) from host FP32 */
int tc_cnn_set_layer(tc_cnn_t * c, int l, const float * w, const float * b){
  if(l < 0 || l > 2) return tc_fail("tc_cnn_set_layer: layer %d", l);
  HIPOK(hipSetDevice(c->device));
  const long long nw = wsize(l);
  if(w){
    HIPOK(hipMemcpyAsync(c->w[l], w, 4 * (size_t)nw, hipMemcpyHostToDevice, c->stream));
    KOK(tck_cnn_launch_to_bf16(c->w[l], nw, c->wbf[l], c->stream));
  }
  if(b) HIPOK(hipMemcpyAsync(c->b[l], b, 4 * (size_t)g_co[l], hipMemcpyHostToDevice, c->stream));
  HIPOK(hipStreamSynchronize(c->stream));
  return 0;
}

int tc_cnn_get_layer(tc_cnn_t * c, int l, float * w, float * b, float * gw, float * gb){
  if(l < 0 || l > 2) return tc_fail("tc_cnn_get_layer: layer %d", l);
  HIPOK(hipSetDevice(c->device));
  const long long nw = wsize(l);
  if(w) HIPOK(hipMemcpyAsync(w, c->w[l], 4 * (size_t)nw, hipMemcpyDeviceToHost, c->stream));
  if(b) HIPOK(hipMemcpyAsync(b, c->b[l], 4 * (size_t)g_co[l], hipMemcpyDeviceToHost, c->stream));
  if(gw) HIPOK(hipMemcpyAsync(gw, c->gw[l], 4 * (size_t)nw, hipMemcpyDeviceToHost, c->stream));
  if(gb) HIPOK(hipMemcpyAsync(gb, c->gb[l], 4 * (size_t)g_co[l], hipMemcpyDeviceToHost, c->stream));
  HIPOK(hipStreamSynchronize(c->stream));
  return 0;
}

/* inputs [nb][nz][ny][nx][5] FP32 (normalised) -> BF16 [nv][16] on the device */
static int upload_x(tc_cnn_t * c, const float * x){
  const long long nv = c->nv;
  for(long long v = 0; v < nv; v++){
    for(int q = 0; q < CNN_C0; q++) c->h_x[v * CNN_C0 + q] = q < 5 ? bf16(x[v * 5 + q]) : 0;
  }
  HIPOK(hipMemcpyAsync(c->x0, c->h_x, 2 * (size_t)nv * CNN_C0, hipMemcpyHostToDevice, c->stream));
  return 0;
}

static int forward(tc_cnn_t * c, const unsigned short * x0){
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, c->nb, x0, CNN_C0, c->wbf[0], CNN_C, c->b[0], CNN_C, 1, c->a1, NULL, c->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, c->nb, c->a1, CNN_C, c->wbf[1], CNN_C, c->b[1], CNN_C, 1, c->a2, NULL, c->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, c->nb, c->a2, CNN_C, c->wbf[2], CNN_CO, c->b[2], CNN_CO, 0, NULL, c->out, c->stream));
  return 0;
}

/* loss and all gradients for the uploaded batch; targets [nb][nz][ny][nx][4] */
static int backward(tc_cnn_t * c, const float * y, double * loss){
  const long long nv = c->nv;
  for(long long v = 0; v < nv; v++){
    for(int q = 0; q < CNN_CO; q++) c->h_y[v * CNN_CO + q] = q < 4 ? y[v * 4 + q] : 0.f;
  }
  HIPOK(hipMemcpyAsync(c->target, c->h_y, 4 * (size_t)nv * CNN_CO, hipMemcpyHostToDevice, c->stream));
  const int nblk = 256;
  KOK(tck_cnn_launch_mse(c->out, c->target, nv * CNN_CO, (float)(1.0 / (double)(nv * 4)), c->gout, c->part, nblk, c->stream));
  float hp[256];
  HIPOK(hipMemcpyAsync(hp, c->part, sizeof(hp), hipMemcpyDeviceToHost, c->stream));
  /* layer 3 */
  KOK(tck_cnn_launch_to_bf16(c->gout, nv * CNN_CO, c->g3, c->stream));
  KOK(tck_cnn_launch_wgrad(c->nx, c->ny, c->nz, c->nb, c->a2, CNN_C, c->g3, CNN_CO, c->part, CNN_NBLK, c->gw[2], c->stream));
  KOK(tck_cnn_launch_bgrad(nv, c->gout, CNN_CO, c->part, CNN_NBLK, c->gb[2], c->stream));
  KOK(tck_cnn_launch_flip(c->wbf[2], CNN_CO, CNN_C, CNN_C, CNN_CO, c->wt[2], c->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, c->nb, c->g3, CNN_CO, c->wt[2], CNN_C, NULL, CNN_C, 0, NULL, c->d2, c->stream));
  /* layer 2 */
  KOK(tck_cnn_launch_relu_bwd(c->d2, c->a2, nv * CNN_C, c->g2, c->g2f, c->stream));
  KOK(tck_cnn_launch_wgrad(c->nx, c->ny, c->nz, c->nb, c->a1, CNN_C, c->g2, CNN_C, c->part, CNN_NBLK, c->gw[1], c->stream));
  KOK(tck_cnn_launch_bgrad(nv, c->g2f, CNN_C, c->part, CNN_NBLK, c->gb[1], c->stream));
  KOK(tck_cnn_launch_flip(c->wbf[1], CNN_C, CNN_C, CNN_C, CNN_C, c->wt[1], c->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, c->nb, c->g2, CNN_C, c->wt[1], CNN_C, NULL, CNN_C, 0, NULL, c->d1, c->stream));
  /* layer 1 */
  KOK(tck_cnn_launch_relu_bwd(c->d1, c->a1, nv * CNN_C, c->g1, c->g1f, c->stream));
  KOK(tck_cnn_launch_wgrad(c->nx, c->ny, c->nz, c->nb, c->x0, CNN_C0, c->g1, CNN_C, c->part, CNN_NBLK, c->gw[0], c->stream));
  KOK(tck_cnn_launch_bgrad(nv, c->g1f, CNN_C, c->part, CNN_NBLK, c->gb[0], c->stream));
  HIPOK(hipStreamSynchronize(c->stream));
  if(loss){
    double s = 0.;
    for(int q = 0; q < nblk; q++) s += hp[q];
    *loss = s / (double)(nv * 4);
  }
  return 0;
}

int tc_cnn_predict(tc_cnn_t * c, const float * x, float * y){
  HIPOK(hipSetDevice(c->device));
  if(upload_x(c, x) || forward(c, c->x0)) return 1;
  float * h = c->h_y;
  HIPOK(hipMemcpyAsync(h, c->out, 4 * (size_t)c->nv * CNN_CO, hipMemcpyDeviceToHost, c->stream));
  HIPOK(hipStreamSynchronize(c->stream));
  for(long long v = 0; v < c->nv; v++) for(int q = 0; q < 4; q++) y[v * 4 + q] = h[v * CNN_CO + q];
  return 0;
}

int tc_cnn_gradients(tc_cnn_t * c, const float * x, const float * y, double * loss){
  HIPOK(hipSetDevice(c->device));
  if(upload_x(c, x) || forward(c, c->x0)) return 1;
  return backward(c, y, loss);
}

int tc_cnn_train_step(tc_cnn_t * c, const float * x, const float * y, float lr, double * loss){
  if(tc_cnn_gradients(c, x, y, loss)) return 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  c->step += 1;
  const float c1 = (float)(1.0 - pow(b1, (double)c->step)), c2 = (float)(1.0 - pow(b2, (double)c->step));
  for(int l = 0; l < 3; l++){
    KOK(tck_cnn_launch_adam(c->w[l], c->mw[l], c->vw[l], c->gw[l], (int)wsize(l), lr, b1, b2, eps, c1, c2, c->wbf[l], c->stream));
    KOK(tck_cnn_launch_adam(c->b[l], c->mb[l], c->vb[l], c->gb[l], g_co[l], lr, b1, b2, eps, c1, c2, NULL, c->stream));
  }
  HIPOK(hipStreamSynchronize(c->stream));
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * coupling with a solver (batch 1 network of the solver's grid)
 * ------------------------------------------------------------------------------------- */
int tc_closure_attach(tc_solver_t * s, tc_cnn_t * c, const float in_mean[5], const float in_std[5],
    const float out_scale[4]){
  if(!c){
    if(s->closure) s->closure->owner = NULL;
    s->closure = NULL;
    s->gen++;
    return 0;
  }
  if(c->owner && c->owner != s)
    return tc_fail("tc_closure_attach: the network is coupled to another solver (one network per solver)");
  if(c->nx != s->p.nx || c->ny != s->p.ny || c->nz != s->p.nz || c->nb != 1)
    return tc_fail("tc_closure_attach: network grid %d x %d x %d (batch %d) does not match the solver's",
        c->nx, c->ny, c->nz, c->nb);
  if(!s->d_corr[0]){
    for(int f = 0; f < TC_NFIELD; f++){
      void * p = NULL;
      if(cnn_alloc(&p, sizeof(float) * (size_t)s->g.n3)) return 1;
      s->d_corr[f] = p;
    }
    double * xc = NULL;
    if(cnn_alloc((void **)&xc, sizeof(double) * (size_t)(s->p.nx + 2))) return 1;
    HIPOK(hipMemcpy(xc, s->h_xc, sizeof(double) * (size_t)(s->p.nx + 2), hipMemcpyHostToDevice));
    s->d_cnn_xc = xc;
  }
  for(int q = 0; q < 5; q++){
    s->cnn_mean[q] = in_mean[q];
    s->cnn_inv_std[q] = in_std[q] > 0.f ? 1.f / in_std[q] : 1.f;
  }
  for(int q = 0; q < 4; q++) s->cnn_scale[q] = out_scale[q];
  if(s->closure && s->closure != c) s->closure->owner = NULL;
  s->closure = c;
  c->owner = s;
  s->gen++;
  return 0;
}

/* the network on the solver's current fields -> corrections (on the solver's stream) */
int tc_closure_eval(tc_solver_t * s){
  tc_cnn_t * c = s->closure;
  if(!c) return 0;
  KOK(tck_cnn_launch_gather(&s->g, tc_field_ptr(s, 0), tc_field_ptr(s, 1), tc_field_ptr(s, 2),
      s->p.scalar ? tc_field_ptr(s, 3) : NULL, s->d_cnn_xc, s->cnn_mean, s->cnn_inv_std, c->x0, s->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, 1, c->x0, CNN_C0, c->wbf[0], CNN_C, c->b[0], CNN_C, 1, c->a1, NULL, s->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, 1, c->a1, CNN_C, c->wbf[1], CNN_C, c->b[1], CNN_C, 1, c->a2, NULL, s->stream));
  KOK(tck_cnn_launch_conv(c->nx, c->ny, c->nz, 1, c->a2, CNN_C, c->wbf[2], CNN_CO, c->b[2], CNN_CO, 0, NULL, c->out, s->stream));
  KOK(tck_cnn_launch_scatter(&s->g, c->out, s->cnn_scale, (float *)s->d_corr[0], (float *)s->d_corr[1],
      (float *)s->d_corr[2], s->p.scalar ? (float *)s->d_corr[3] : NULL, s->stream));
  return 0;
}

/* a solver is destroyed: its network is free to be coupled elsewhere */
void tc_closure_release(tc_solver_t * s){
  if(s->closure) s->closure->owner = NULL;
  s->closure = NULL;
  s->gen++;
}
