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
 * sample b of a batch.  The convolution is an implicit GEMM on the matrix cores
 * (v_mfma_f32_16x16x32_bf16, FP32 accumulation) over activation tiles staged in LDS with
 * their halo; the weight gradient uses v_mfma_f32_16x16x16_bf16.  Weights are W[tap][co][ci]
 * BF16 with FP32 master copies; biases FP32.
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
 *   CI input channels (16 or 32), CO output channels (16 or 32) of which the first co_store
 *   are written; relu: out = max(0, .) stored BF16 (next layer's input), otherwise FP32.
 *
 * Implicit GEMM on v_mfma_f32_16x16x32_bf16: A = 16 cells of an x row by 32 K entries, B = 32
 * K entries by 16 output channels.  K is the 32 input channels of one tap (CI = 32), or the
 * 16 input channels of two taps (CI = 16, the 27th tap paired with zeros).  Lane l holds the
 * eight K entries 8 (l / 16) .. +7 of A row l % 16 and of B column l % 16 (the same K order on
 * both sides, which is all the product needs); the accumulator holds rows (cells)
 * 4 (l / 16) .. +3 of column l % 16.
 *
 * A block stages the kernel once (rows of CI BF16 padded by 16 bytes, so that the 16 lanes
 * reading one K group hit distinct banks) and loops over tiles of 16 x CNN_TY x CNN_TZ cells
 * of one sample: each tile's activations with a one-cell halo (zero outside the walls in x,
 * periodic images in y and z) are staged in LDS with the same padded rows, and every tap of
 * every cell is then read from LDS.  Each wave takes CNN_RPW rows of the tile.
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
#define CNN_TY 4
#define CNN_TZ 4
#define CNN_HX 18                             /* halo'd tile: 16 + 2 cells in x */
#define CNN_HCELLS (CNN_HX * (CNN_TY + 2) * (CNN_TZ + 2))
#define CNN_RPW ((CNN_TY * CNN_TZ) / (CNN_CONV_THREADS / TCK_WAVE))   /* rows per wave */

typedef __bf16 cnn_b8 __attribute__((ext_vector_type(8)));

/* LDS bytes of the convolution of CI -> CO channels: the kernel and one halo'd tile */
__host__ __device__ constexpr int cnn_conv_lds(int ci, int co){
  return (27 * co + CNN_HCELLS) * (2 * ci + 16);
}

__device__ __forceinline__ int cnn_wrap(int a, int n){
  a %= n;
  return a < 0 ? a + n : a;
}

template <int CI, int CO>
__global__ void __launch_bounds__(CNN_CONV_THREADS) tck_cnn_conv(const ConvArgs A){
  constexpr int RS = 2 * CI + 16;                 /* padded row: bytes per cell or per co */
  constexpr int NT = CO / 16;                     /* 16-channel output blocks */
  constexpr int NK = CI == 32 ? 27 : 14;          /* MFMA steps over the taps */
  extern __shared__ __align__(16) unsigned char cnn_smem[];
  unsigned char * wl = cnn_smem;                  /* [27][CO] rows */
  unsigned char * al = cnn_smem + 27 * CO * RS;   /* [HZ][HY][HX] rows */
  /* the kernel: 16-byte pieces, CI / 8 per row */
  for(int e = threadIdx.x; e < 27 * CO * (CI / 8); e += CNN_CONV_THREADS){
    const int row = e / (CI / 8), pc = e % (CI / 8);
    *(uint4 *)(wl + row * RS + 16 * pc) = *(const uint4 *)(A.w + (long long)row * CI + 8 * pc);
  }
  const int lane = threadIdx.x % TCK_WAVE;
  const int wv = threadIdx.x / TCK_WAVE;
  const int kg = lane / 16;                       /* this lane's K group */
  const int cl = lane % 16;                       /* A row (cell in x) / B column (channel) */
  const CnnGrid & G = A.G;
  const int ntx = (G.nx + 15) / 16, nty = (G.ny + CNN_TY - 1) / CNN_TY, ntz = (G.nz + CNN_TZ - 1) / CNN_TZ;
  const long long ntiles = (long long)ntx * nty * ntz * G.nb;
  /* CI = 16: K group kg reads tap 2 s + (kg >> 1), channels 8 (kg & 1) .. +7 */
  const int ktap = CI == 32 ? 0 : (kg >> 1);
  const int kbyte = CI == 32 ? 16 * kg : 16 * (kg & 1);
  for(long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x){
    const int tx = (int)(tile % ntx);
    long long r = tile / ntx;
    const int ty = (int)(r % nty); r /= nty;
    const int tz = (int)(r % ntz);
    const int b = (int)(r / ntz);
    const int x0 = 16 * tx, y0 = CNN_TY * ty, z0 = CNN_TZ * tz;
    __syncthreads();                              /* the previous tile's reads are done */
    for(int e = threadIdx.x; e < CNN_HCELLS * (CI / 8); e += CNN_CONV_THREADS){
      const int c = e / (CI / 8), pc = e % (CI / 8);
      const int hx = c % CNN_HX, hy = (c / CNN_HX) % (CNN_TY + 2), hz = c / (CNN_HX * (CNN_TY + 2));
      const int xx = x0 + hx - 1;
      uint4 v = {0u, 0u, 0u, 0u};
      if(xx >= 0 && xx < G.nx){
        const int yy = cnn_wrap(y0 + hy - 1, G.ny), zz = cnn_wrap(z0 + hz - 1, G.nz);
        const long long src = (long long)xx + (long long)G.nx * ((long long)yy + (long long)G.ny * ((long long)zz + (long long)G.nz * b));
        v = *(const uint4 *)(A.in + src * CI + 8 * pc);
      }
      *(uint4 *)(al + c * RS + 16 * pc) = v;
    }
    __syncthreads();
    cnn_f4 acc[CNN_RPW][NT];
    #pragma unroll
    for(int q = 0; q < CNN_RPW; q++)
      #pragma unroll
      for(int t = 0; t < NT; t++) acc[q][t] = cnn_f4{0.f, 0.f, 0.f, 0.f};
    for(int s = 0; s < NK; s++){
      const int tap = CI == 32 ? s : 2 * s + ktap;
      const bool live = tap < 27;
      const int tp = live ? tap : 26;
      const int dx = tp % 3, dy = (tp / 3) % 3, dz = tp / 9;   /* halo offsets 0..2 */
      cnn_b8 bw[NT];
      #pragma unroll
      for(int t = 0; t < NT; t++){
        bw[t] = *(const cnn_b8 *)(wl + (tp * CO + 16 * t + cl) * RS + kbyte);
        if(!live) bw[t] = cnn_b8{};
      }
      #pragma unroll
      for(int q = 0; q < CNN_RPW; q++){
        const int row = wv * CNN_RPW + q;
        const int ry = row % CNN_TY, rz = row / CNN_TY;
        const int c = (cl + dx) + CNN_HX * ((ry + dy) + (CNN_TY + 2) * (rz + dz));
        const cnn_b8 av = *(const cnn_b8 *)(al + c * RS + kbyte);
        #pragma unroll
        for(int t = 0; t < NT; t++) acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[t], acc[q][t], 0, 0, 0);
      }
    }
    /* accumulator: cells 4 kg .. +3 of the row, output channel 16 t + cl */
    #pragma unroll
    for(int q = 0; q < CNN_RPW; q++){
      const int row = wv * CNN_RPW + q;
      const int y = y0 + row % CNN_TY, z = z0 + row / CNN_TY;
      if(y >= G.ny || z >= G.nz) continue;
      #pragma unroll
      for(int t = 0; t < NT; t++){
        const int co = 16 * t + cl;
        if(co >= A.co_store) continue;
        const float bv = A.bias ? A.bias[co] : 0.f;
        #pragma unroll
        for(int u = 0; u < 4; u++){
          const int xc = x0 + 4 * kg + u;
          if(xc >= G.nx) continue;
          const long long v = (long long)xc + (long long)G.nx * ((long long)y + (long long)G.ny * ((long long)z + (long long)G.nz * b));
          const float val = acc[q][t][u] + bv;
          if(A.relu) A.out_bf[v * A.co_store + co] = cnn_bf16(val > 0.f ? val : 0.f);
          else A.out_f[v * A.co_store + co] = val;
        }
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
/* a convolution launch; the first one of an instantiation raises its dynamic LDS limit */
template <int CI, int CO>
static int cnn_conv_launch(const ConvArgs & A, dim3 grd, int lds, hipStream_t stream){
  static int ready = 0;
  if(!ready){
    const hipError_t e = hipFuncSetAttribute((const void *)tck_cnn_conv<CI, CO>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if(e != hipSuccess) return (int)e;
    ready = 1;
  }
  hipLaunchKernelGGL((tck_cnn_conv<CI, CO>), grd, dim3(CNN_CONV_THREADS), lds, stream, A);
  return 0;
}

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
  /* persistent blocks, as many as the LDS holds per CU (the kernel and a halo'd tile) */
  const long long ntiles = (long long)((nx + 15) / 16) * ((ny + CNN_TY - 1) / CNN_TY) * ((nz + CNN_TZ - 1) / CNN_TZ) * nb;
  const int lds = cnn_conv_lds(ci, co);
  long long per_cu = (160LL * 1024) / lds;
  per_cu = per_cu < 1 ? 1 : (per_cu > 4 ? 4 : per_cu);
  long long nblk = ntiles;
  if(nblk > per_cu * cnn_ncu()) nblk = per_cu * cnn_ncu();
  const dim3 grd((unsigned)nblk);
  int rc;
  if(ci == 16 && co == 32) rc = cnn_conv_launch<16, 32>(A, grd, lds, stream);
  else if(ci == 32 && co == 32) rc = cnn_conv_launch<32, 32>(A, grd, lds, stream);
  else if(ci == 32 && co == 16) rc = cnn_conv_launch<32, 16>(A, grd, lds, stream);
  else if(ci == 16 && co == 16) rc = cnn_conv_launch<16, 16>(A, grd, lds, stream);
  else return (int)hipErrorInvalidValue;
  if(rc) return rc;
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
