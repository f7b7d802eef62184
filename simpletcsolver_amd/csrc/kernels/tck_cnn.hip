/*
 * tck_cnn.hip -- the learned subgrid closure: a small 3-D CNN on the coarse grid (gfx950).
 *
 * No reference counterpart: the reference solves the flow on one grid only.  The network
 * maps (u, v, w, T, r) at the coarse cell centres to a 4-component correction of the
 * right-hand sides of ux, uy, uz and T.  Layers: 3x3x3 convolutions, 16 (padded inputs) ->
 * 32 -> 32 -> 4 channels, ReLU between; the x direction (walls) is zero-padded, y and z are
 * periodic like the flow.
 *
 * Activations are channels-last BF16 rows act[v][C], v = x + NX (y + NY (z + NZ b)) for
 * sample b of a batch.  The convolution
 * is an implicit GEMM on the matrix cores: for 16 consecutive cells of a row and 16 output
 * channels a wavefront accumulates, over the 27 taps and the input channels in groups of 16,
 * v_mfma_f32_16x16x16_bf16 products in FP32 (lane l holds A row l % 16 = cell, B column
 * l % 16 = output channel, both over input channels 4 (l / 16) .. +3; the accumulator holds
 * output rows 4 (l / 16) .. +3 of column l % 16).  Weights are W[tap][co][ci] BF16 with
 * FP32 master copies; biases FP32.
 *
 * Training: the backward data pass is the same convolution with the kernel flipped and its
 * channel roles exchanged (the adjoint of zero padding is zero padding, of periodic wrapping
 * periodic wrapping), so tck_cnn_conv serves both; the weight gradient is an FP32 reduction
 * over cells (per-block partial sums, then a fixed-order sum, so gradients are reproducible);
 * Adam updates the FP32 masters and re-rounds the BF16 copies.
 */
#include "tck_common.h"
#include <stdint.h>

typedef short cnn_s4 __attribute__((ext_vector_type(4)));
typedef float cnn_f4 __attribute__((ext_vector_type(4)));

/* FP32 -> BF16, round to nearest even (inputs here are finite) */
__device__ __forceinline__ unsigned short cnn_bf16(float x){
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float cnn_f32(unsigned short h){
  return __uint_as_float(((uint32_t)h) << 16);
}

/* nb samples of nx x ny x nz cells, v = x + nx (y + ny (z + nz b)) */
struct CnnGrid { int nx, ny, nz, nb; long long nv; };

/* cell of a tap within sample b: x zero-padded (returns -1 outside), y and z periodic */
__device__ __forceinline__ long long cnn_tap(const CnnGrid & G, int x, int y, int z, int b, int tap){
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int xx = x + dx;
  if(xx < 0 || xx >= G.nx) return -1;
  int yy = y + dy; yy = yy < 0 ? yy + G.ny : (yy >= G.ny ? yy - G.ny : yy);
  int zz = z + dz; zz = zz < 0 ? zz + G.nz : (zz >= G.nz ? zz - G.nz : zz);
  return (long long)xx + (long long)G.nx * ((long long)yy + (long long)G.ny * ((long long)zz + (long long)G.nz * b));
}

/* ---------------------------------------------------------------------------------------
 * convolution: out[v][co] = bias[co] + sum_tap sum_ci W[tap][co][ci] in[v + tap][ci]
 *   CI input channels (multiple of 16), CO output channels (multiple of 16) of which the
 *   first co_store are written; relu: out = max(0, .) stored BF16 (next layer's input),
 *   otherwise FP32.  One wavefront per 16 cells x all CO; 4 waves per block.
 * ------------------------------------------------------------------------------------- */
struct ConvArgs {
  CnnGrid G;
  const unsigned short * in;     /* [nv][CI] bf16 */
  const unsigned short * w;      /* [27][CO][CI] bf16 */
  const float * bias;            /* [CO] or null */
  unsigned short * out_bf;       /* [nv][co_store] bf16 (relu) */
  float * out_f;                 /* [nv][co_store] fp32 (no relu) */
  int co_store;
  int relu;
};

#define CNN_CONV_THREADS 512

/* The block stages the whole kernel W (27 CO CI BF16, at most 54 KiB) in LDS once and its
 * waves then loop over 16-cell segments, so the B operands come from LDS and only the
 * activations are read from memory. */
template <int CI, int CO>
__global__ void __launch_bounds__(CNN_CONV_THREADS) tck_cnn_conv(const ConvArgs A){
  constexpr int NW = 27 * CO * CI;
  __shared__ __align__(16) unsigned short wl[NW];
  for(int e = 8 * threadIdx.x; e < NW; e += 8 * CNN_CONV_THREADS) *(uint4 *)(wl + e) = *(const uint4 *)(A.w + e);
  __syncthreads();
  const int lane = threadIdx.x % TCK_WAVE;
  const int wpb = CNN_CONV_THREADS / TCK_WAVE;
  const int rows = (A.G.nx + 15) / 16;                 /* 16-cell segments per x row */
  const long long nseg = (long long)rows * A.G.ny * A.G.nz * A.G.nb;
  const int kq = 4 * (lane / 16);                      /* this lane's 4 K entries */
  for(long long wave = (long long)blockIdx.x * wpb + threadIdx.x / TCK_WAVE; wave < nseg; wave += (long long)gridDim.x * wpb){
    const int seg = (int)(wave % rows);
    const long long yzb = wave / rows;
    const int y = (int)(yzb % A.G.ny);
    const int z = (int)((yzb / A.G.ny) % A.G.nz);
    const int b = (int)(yzb / ((long long)A.G.ny * A.G.nz));
    const int x = seg * 16 + lane % 16;                /* A row: this lane's cell */
    cnn_f4 acc[CO / 16];
    #pragma unroll
    for(int t = 0; t < CO / 16; t++) acc[t] = cnn_f4{0.f, 0.f, 0.f, 0.f};
    #pragma unroll 3
    for(int tap = 0; tap < 27; tap++){
      const long long src = x < A.G.nx ? cnn_tap(A.G, x, y, z, b, tap) : -1;
      #pragma unroll
      for(int c0 = 0; c0 < CI; c0 += 16){
        /* 4 consecutive BF16 channels: one 8-byte load (rows are 32-byte aligned) */
        cnn_s4 a = {0, 0, 0, 0};
        if(src >= 0) a = *(const cnn_s4 *)(A.in + src * CI + c0 + kq);
        #pragma unroll
        for(int t = 0; t < CO / 16; t++){
          const cnn_s4 wq = *(const cnn_s4 *)(wl + (tap * CO + t * 16 + lane % 16) * CI + c0 + kq);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wq, acc[t], 0, 0, 0);
        }
      }
    }
    /* accumulator: rows (cells) kq .. kq+3 of column (output channel) lane % 16 */
    #pragma unroll
    for(int t = 0; t < CO / 16; t++){
      const int co = t * 16 + lane % 16;
      if(co >= A.co_store) continue;
      const float bv = A.bias ? A.bias[co] : 0.f;
      #pragma unroll
      for(int q = 0; q < 4; q++){
        const int xc = seg * 16 + kq + q;
        if(xc >= A.G.nx) continue;
        const long long v = (long long)xc + (long long)A.G.nx * ((long long)y + (long long)A.G.ny * ((long long)z + (long long)A.G.nz * b));
        const float val = acc[t][q] + bv;
        if(A.relu) A.out_bf[v * A.co_store + co] = cnn_bf16(val > 0.f ? val : 0.f);
        else A.out_f[v * A.co_store + co] = val;
      }
    }
  }
}

/* ---------------------------------------------------------------------------------------
 * weight gradient on the matrix cores: for each tap a GEMM over cells,
 *   part[blk][tap][co][ci] = sum over the block's cells v of g[v][co] * in[v + tap][ci]
 * (FP32), then a fixed-order sum over blocks.  Block (blk, tap): 4 waves, each taking
 * 16-cell groups of the block's range in turn; lane l holds A = g^T row co = l % 16 over
 * cells 4 (l / 16) .. +3 of the group, B = the shifted inputs, column ci = l % 16 over the
 * same cells; the waves' accumulators are summed in LDS in wave order.
 * ------------------------------------------------------------------------------------- */
struct WgradArgs {
  CnnGrid G;
  const unsigned short * in;   /* [nv][CI] bf16 */
  const unsigned short * g;    /* [nv][CO] bf16 */
  long long cells_per_block;   /* multiple of 64 */
  float * part;                /* [nblk][27][CO][CI] */
};

template <int CI, int CO>
__global__ void __launch_bounds__(256) tck_cnn_wgrad(const WgradArgs A){
  constexpr int NT = CO / 16, NS = CI / 16;
  __shared__ float red[4][CO * CI];
  const int lane = threadIdx.x % TCK_WAVE, wv = threadIdx.x / TCK_WAVE;
  const int tap = blockIdx.y;
  const long long v0 = (long long)blockIdx.x * A.cells_per_block;
  const long long v1 = min(v0 + A.cells_per_block, A.G.nv);
  cnn_f4 acc[NT][NS];
  #pragma unroll
  for(int t = 0; t < NT; t++)
    #pragma unroll
    for(int u = 0; u < NS; u++) acc[t][u] = cnn_f4{0.f, 0.f, 0.f, 0.f};
  const int kq = 4 * (lane / 16), col = lane % 16;
  for(long long c = v0 + 16 * wv; c < v1; c += 64){
    cnn_s4 a[NT], b[NS];
    #pragma unroll
    for(int t = 0; t < NT; t++) a[t] = cnn_s4{0, 0, 0, 0};
    #pragma unroll
    for(int u = 0; u < NS; u++) b[u] = cnn_s4{0, 0, 0, 0};
    #pragma unroll
    for(int q = 0; q < 4; q++){
      const long long v = c + kq + q;
      if(v >= v1) break;
      const int x = (int)(v % A.G.nx);
      const long long r = v / A.G.nx;
      const int y = (int)(r % A.G.ny);
      const int z = (int)((r / A.G.ny) % A.G.nz);
      const int bb = (int)(r / ((long long)A.G.ny * A.G.nz));
      const long long src = cnn_tap(A.G, x, y, z, bb, tap);
      if(src < 0) continue;
      #pragma unroll
      for(int t = 0; t < NT; t++) a[t][q] = (short)A.g[v * CO + 16 * t + col];
      #pragma unroll
      for(int u = 0; u < NS; u++) b[u][q] = (short)A.in[src * CI + 16 * u + col];
    }
    #pragma unroll
    for(int t = 0; t < NT; t++)
      #pragma unroll
      for(int u = 0; u < NS; u++) acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], b[u], acc[t][u], 0, 0, 0);
  }
  /* accumulator [t][u][q]: co = 16 t + kq + q, ci = 16 u + col */
  #pragma unroll
  for(int t = 0; t < NT; t++)
    #pragma unroll
    for(int u = 0; u < NS; u++)
      #pragma unroll
      for(int q = 0; q < 4; q++) red[wv][(16 * t + kq + q) * CI + 16 * u + col] = acc[t][u][q];
  __syncthreads();
  float * out = A.part + ((long long)blockIdx.x * 27 + tap) * (CO * CI);
  for(int e = threadIdx.x; e < CO * CI; e += 256) out[e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
}

/* out[e] = sum_b part[b][e] (b in order) */
__global__ void tck_cnn_reduce(const float * part, int nblk, int n, float * out){
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= n) return;
  float s = 0.f;
  for(int b = 0; b < nblk; b++) s += part[(long long)b * n + e];
  out[e] = s;
}

/* bias gradient: db[co] partial sums of g over cells (fp32 from the un-rounded gradient) */
__global__ void tck_cnn_bgrad(const float * g, long long nv, int co_n, long long cells_per_block, float * part){
  const long long v0 = (long long)blockIdx.x * cells_per_block;
  const long long v1 = min(v0 + cells_per_block, nv);
  for(int co = threadIdx.x; co < co_n; co += blockDim.x){
    float s = 0.f;
    for(long long v = v0; v < v1; v++) s += g[v * co_n + co];
    part[(long long)blockIdx.x * co_n + co] = s;
  }
}

/* ---------------------------------------------------------------------------------------
 * elementwise pieces
 * ------------------------------------------------------------------------------------- */
/* loss = mean over cells and channels of (pred - target)^2; grad = d loss / d pred.
 * part[blk] = partial sum of squares */
__global__ void __launch_bounds__(256) tck_cnn_mse(const float * pred, const float * target, long long n,
    float scale, float * grad, float * part){
  __shared__ float red[256];
  float s = 0.f;
  for(long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x){
    const float d = pred[e] - target[e];
    grad[e] = 2.f * d * scale;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for(int h = 128; h > 0; h >>= 1){
    if((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* gradient through ReLU: g_pre = g_post where the stored activation is positive, as BF16
 * (the operand of the next products); gf: the same in FP32 for the bias gradient */
__global__ void tck_cnn_relu_bwd(const float * gpost, const unsigned short * act, long long n,
    unsigned short * gpre, float * gf){
  for(long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x){
    const float v = cnn_f32(act[e]) > 0.f ? gpost[e] : 0.f;
    gpre[e] = cnn_bf16(v);
    if(gf) gf[e] = v;
  }
}

/* FP32 -> BF16 copy */
__global__ void tck_cnn_to_bf16(const float * x, long long n, unsigned short * y){
  for(long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    y[e] = cnn_bf16(x[e]);
}

/* adjoint kernel for the backward data pass: Wt[26 - tap][ci][co] = W[tap][co][ci]
 * (co, ci padded in the transposed tensor to co_p, ci_p with zeros) */
__global__ void tck_cnn_flip(const unsigned short * w, int co_n, int ci_n, int ci_p, int co_p, unsigned short * wt){
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = 27 * ci_p * co_p;
  if(e >= n) return;
  const int co = e % co_p;                /* wt's input channel */
  const int ci = (e / co_p) % ci_p;       /* wt's output channel */
  const int tap = e / (co_p * ci_p);
  unsigned short v = 0;
  if(co < co_n && ci < ci_n) v = w[((long long)(26 - tap) * co_n + co) * ci_n + ci];
  wt[e] = v;
}

/* Adam on FP32 masters, BF16 copies re-rounded */
__global__ void tck_cnn_adam(float * w, float * m, float * v, const float * g, int n, float lr,
    float b1, float b2, float eps, float c1, float c2, unsigned short * wbf){
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gg;
  const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  const float upd = lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w[e] -= upd;
  if(wbf) wbf[e] = cnn_bf16(w[e]);
}

/* ---------------------------------------------------------------------------------------
 * coupling with the flow (device layout of tc_shared.h)
 * ------------------------------------------------------------------------------------- */
struct GatherArgs {
  tc_geom_t g;
  const double * ux; const double * uy; const double * uz; const double * t;
  const double * xc;        /* XC(i) at [i] */
  float mean[5], inv_std[5];
  unsigned short * out;     /* [nv][16] bf16: u v w T r, channels 5..15 zero */
};

/* network input at cell centres: u = (UX(i) + UX(i+1)) / 2, v = UY, w = UZ, T, r = XC(i) */
__global__ void tck_cnn_gather(const GatherArgs A){
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if(i > A.g.nx) return;
  const long long o = tck_idx(A.g, i, j, k);
  float c[5];
  c[0] = (float)(0.5 * A.ux[o] + 0.5 * A.ux[o + 1]);
  c[1] = (float)A.uy[o];
  c[2] = (float)A.uz[o];
  c[3] = A.t ? (float)A.t[o] : 0.f;
  c[4] = (float)A.xc[i];
  const long long v = (long long)(i - 1) + (long long)A.g.nx * ((long long)(j - 1) + (long long)A.g.ny * (k - 1));
  unsigned short * p = A.out + v * 16;
  #pragma unroll
  for(int q = 0; q < 5; q++) p[q] = cnn_bf16((c[q] - A.mean[q]) * A.inv_std[q]);
  #pragma unroll
  for(int q = 5; q < 16; q++) p[q] = 0;
}

struct ScatterArgs {
  tc_geom_t g;
  const float * pred;       /* [nv][16] network output, channels 0..3 used */
  float scale[4];           /* output de-normalisation */
  float * cx; float * cy; float * cz; float * ct;   /* corrections, field layout */
};

/* corrections at the variables' own positions: ux faces i = 2..nx average the two centres,
 * uy, uz, T at centres */
__global__ void tck_cnn_scatter(const ScatterArgs A){
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if(i > A.g.nx) return;
  const long long o = tck_idx(A.g, i, j, k);
  const long long v = (long long)(i - 1) + (long long)A.g.nx * ((long long)(j - 1) + (long long)A.g.ny * (k - 1));
  const float * p = A.pred + v * 16;
  A.cy[o] = A.scale[1] * p[1];
  A.cz[o] = A.scale[2] * p[2];
  if(A.ct) A.ct[o] = A.scale[3] * p[3];
  if(i >= 2) A.cx[o] = A.scale[0] * (0.5f * A.pred[(v - 1) * 16] + 0.5f * p[0]);
  else A.cx[o] = 0.f;
}

/* ------------------------------------------------------------------------------------- */
static int g_cnn_ncu = 0;

static int cnn_ncu(void){
  if(g_cnn_ncu == 0){
    int dev = 0, v = 0;
    if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    g_cnn_ncu = v;
  }
  return g_cnn_ncu;
}

extern "C" {

int tck_cnn_launch_conv(int nx, int ny, int nz, int nb, const unsigned short * in, int ci,
    const unsigned short * w, int co, const float * bias, int co_store, int relu,
    unsigned short * out_bf, float * out_f, hipStream_t stream){
  ConvArgs A;
  A.G.nx = nx; A.G.ny = ny; A.G.nz = nz; A.G.nb = nb; A.G.nv = (long long)nx * ny * nz * nb;
  A.in = in; A.w = w; A.bias = bias; A.out_bf = out_bf; A.out_f = out_f;
  A.co_store = co_store; A.relu = relu;
  /* as many blocks as the LDS holds per CU (the staged kernel takes 27 co ci 2 bytes), at
   * most four, each looping over segments */
  const long long nseg = (long long)((nx + 15) / 16) * ny * nz * nb;
  const int wpb = CNN_CONV_THREADS / TCK_WAVE;
  long long per_cu = (160LL * 1024) / (27LL * co * ci * 2);
  per_cu = per_cu < 1 ? 1 : (per_cu > 4 ? 4 : per_cu);
  long long nblk = (nseg + wpb - 1) / wpb;
  if(nblk > per_cu * cnn_ncu()) nblk = per_cu * cnn_ncu();
  const dim3 grd((unsigned)nblk), blk(CNN_CONV_THREADS);
  if(ci == 16 && co == 32) hipLaunchKernelGGL((tck_cnn_conv<16, 32>), grd, blk, 0, stream, A);
  else if(ci == 32 && co == 32) hipLaunchKernelGGL((tck_cnn_conv<32, 32>), grd, blk, 0, stream, A);
  else if(ci == 32 && co == 16) hipLaunchKernelGGL((tck_cnn_conv<32, 16>), grd, blk, 0, stream, A);
  else if(ci == 16 && co == 16) hipLaunchKernelGGL((tck_cnn_conv<16, 16>), grd, blk, 0, stream, A);
  else return (int)hipErrorInvalidValue;
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_wgrad(int nx, int ny, int nz, int nb, const unsigned short * in, int ci_n,
    const unsigned short * g, int co_n, float * part, int nblk, float * out, hipStream_t stream){
  WgradArgs A;
  A.G.nx = nx; A.G.ny = ny; A.G.nz = nz; A.G.nb = nb; A.G.nv = (long long)nx * ny * nz * nb;
  A.in = in; A.g = g; A.part = part;
  A.cells_per_block = ((A.G.nv + nblk - 1) / nblk + 63) / 64 * 64;
  const dim3 grd((unsigned)nblk, 27);
  if(ci_n == 16 && co_n == 32) hipLaunchKernelGGL((tck_cnn_wgrad<16, 32>), grd, dim3(256), 0, stream, A);
  else if(ci_n == 32 && co_n == 32) hipLaunchKernelGGL((tck_cnn_wgrad<32, 32>), grd, dim3(256), 0, stream, A);
  else if(ci_n == 32 && co_n == 16) hipLaunchKernelGGL((tck_cnn_wgrad<32, 16>), grd, dim3(256), 0, stream, A);
  else return (int)hipErrorInvalidValue;
  TCK_CHECK_LAUNCH();
  const int n = 27 * co_n * ci_n;
  hipLaunchKernelGGL(tck_cnn_reduce, dim3((n + 255) / 256), dim3(256), 0, stream, (const float *)part, nblk, n, out);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_bgrad(long long nv, const float * g, int co_n, float * part, int nblk, float * out,
    hipStream_t stream){
  hipLaunchKernelGGL(tck_cnn_bgrad, dim3(nblk), dim3(64), 0, stream, g, nv, co_n, (nv + nblk - 1) / nblk, part);
  TCK_CHECK_LAUNCH();
  hipLaunchKernelGGL(tck_cnn_reduce, dim3((co_n + 255) / 256), dim3(256), 0, stream, (const float *)part, nblk, co_n, out);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_mse(const float * pred, const float * target, long long n, float scale,
    float * grad, float * part, int nblk, hipStream_t stream){
  hipLaunchKernelGGL(tck_cnn_mse, dim3(nblk), dim3(256), 0, stream, pred, target, n, scale, grad, part);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_relu_bwd(const float * gpost, const unsigned short * act, long long n,
    unsigned short * gpre, float * gf, hipStream_t stream){
  hipLaunchKernelGGL(tck_cnn_relu_bwd, dim3(2048), dim3(256), 0, stream, gpost, act, n, gpre, gf);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_to_bf16(const float * x, long long n, unsigned short * y, hipStream_t stream){
  hipLaunchKernelGGL(tck_cnn_to_bf16, dim3(2048), dim3(256), 0, stream, x, n, y);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_flip(const unsigned short * w, int co_n, int ci_n, int ci_p, int co_p,
    unsigned short * wt, hipStream_t stream){
  const int n = 27 * ci_p * co_p;
  hipLaunchKernelGGL(tck_cnn_flip, dim3((n + 255) / 256), dim3(256), 0, stream, w, co_n, ci_n, ci_p, co_p, wt);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_adam(float * w, float * m, float * v, const float * g, int n, float lr,
    float b1, float b2, float eps, float c1, float c2, unsigned short * wbf, hipStream_t stream){
  hipLaunchKernelGGL(tck_cnn_adam, dim3((n + 255) / 256), dim3(256), 0, stream, w, m, v, g, n, lr,
      b1, b2, eps, c1, c2, wbf);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_gather(const tc_geom_t * g, const double * ux, const double * uy,
    const double * uz, const double * t, const double * xc, const float mean[5],
    const float inv_std[5], unsigned short * out, hipStream_t stream){
  GatherArgs A;
  A.g = *g; A.ux = ux; A.uy = uy; A.uz = uz; A.t = t; A.xc = xc; A.out = out;
  for(int q = 0; q < 5; q++){ A.mean[q] = mean[q]; A.inv_std[q] = inv_std[q]; }
  hipLaunchKernelGGL(tck_cnn_gather, dim3((g->nx + 63) / 64, g->ny, g->nz), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_cnn_launch_scatter(const tc_geom_t * g, const float * pred, const float scale[4],
    float * cx, float * cy, float * cz, float * ct, hipStream_t stream){
  ScatterArgs A;
  A.g = *g; A.pred = pred; A.cx = cx; A.cy = cy; A.cz = cz; A.ct = ct;
  for(int q = 0; q < 4; q++) A.scale[q] = scale[q];
  hipLaunchKernelGGL(tck_cnn_scatter, dim3((g->nx + 63) / 64, g->ny, g->nz), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

}
