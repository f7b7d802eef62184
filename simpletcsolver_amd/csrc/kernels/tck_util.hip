/*
 * tck_util.hip -- device initial-condition fill and the bandwidth probe (gfx950).
 *
 *  tck_ic_fill   seeded field of the IC tool (host/tc_ic.c states the definition): per-i
 *                profile + amp (2u - 1) splitmix64 noise + mode_amp X(i) Y(j) Z(k), evaluated
 *                on the device with the same integer hash and the same operand order as
 *                the host generator, so both give identical values.
 *  tck_copy      dst = src over n doubles: each lane moves U independent 16 B chunks of one
 *                block tile (U loads in flight before the first store), loaded and stored
 *                non-temporally; one tile per block, no grid-stride loop.  The measured HBM
 *                copy ceiling (the best of U = 1, 2, 4 and 8: on MI355X many small blocks,
 *                U = 1, copy fastest) and the reference kernel for scaling byte counters.
 */
#include "tck_common.h"
#include <mutex>

/* Staging of the allocation-free initial-condition fill (tc_ic_fields_noalloc): the four
 * per-index tables of one field (profile, x, y and z mode factors), TC_IC_NOALLOC_N doubles
 * each, in one file-scope device array.  Every device the library runs on holds its own copy
 * (128 KiB) until the process ends.  Host code never passes the array's name to a kernel or a
 * copy: tck_ic_scratch looks up its device address once per device (hipGetSymbolAddress with
 * that device current) and hands that out. */
static __device__ __align__(256) double scratch[4 * TC_IC_NOALLOC_N];

#define TCK_MAXDEV 64
static std::mutex g_scratch_mu;
static double * g_scratch_ptr[TCK_MAXDEV];

extern "C" int tck_ic_scratch(double ** ptr){
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if(e != hipSuccess) return (int)e;
  if(dev < 0 || dev >= TCK_MAXDEV) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if(!g_scratch_ptr[dev]){
    void * p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = (double *)p;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}

__device__ __forceinline__ unsigned long long tck_splitmix64(unsigned long long z){
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct IcArgs {
  tc_geom_t g;
  double * dst;
  const double * prof;    /* per NPY x index q */
  const double * mx;      /* per q, or null */
  const double * my;      /* per j */
  const double * mz;      /* per k */
  double amp, mode_amp;
  unsigned long long seed;
  unsigned field;
  int nxs, i0;
};

__global__ void tck_ic_fill(const IcArgs A){
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;        /* 0..ny-1 */
  const int k = blockIdx.z;        /* 0..nz-1 */
  if(q >= A.nxs) return;
  double v = A.prof[q];
  if(A.amp != 0.){
    const unsigned long long lin = ((unsigned long long)k * A.g.ny + j) * A.nxs + q;
    const unsigned long long key = A.seed * 0x9E3779B97F4A7C15ULL + ((unsigned long long)A.field << 56) + lin;
    const double u = (double)(tck_splitmix64(key) >> 11) * 0x1.0p-53;
    v = v + A.amp * (2. * u - 1.);
  }
  if(A.mx) v = v + A.mode_amp * A.mx[q] * A.my[j] * A.mz[k];
  A.dst[tck_idx(A.g, A.i0 + q, j + 1, k + 1)] = v;
}

template <int U>
__global__ __launch_bounds__(TCK_BLOCK) void tck_copy(const double2 * __restrict__ src,
    double2 * __restrict__ dst, long long n2){
  const long long base = (long long)blockIdx.x * (TCK_BLOCK * U) + threadIdx.x;
  double2 v[U];
  #pragma unroll
  for(int u = 0; u < U; u++){
    const long long t = base + (long long)u * TCK_BLOCK;
    if(t < n2) v[u] = tck_ld(&src[t]);
  }
  #pragma unroll
  for(int u = 0; u < U; u++){
    const long long t = base + (long long)u * TCK_BLOCK;
    if(t < n2) tck_st(&dst[t], v[u]);
  }
}

template <int U>
static void launch_copy(const double * src, double * dst, long long n2, hipStream_t stream){
  const long long nb = (n2 + TCK_BLOCK * U - 1) / (TCK_BLOCK * U);
  hipLaunchKernelGGL(tck_copy<U>, dim3((unsigned)nb), dim3(TCK_BLOCK), 0, stream,
      (const double2 *)src, (double2 *)dst, n2);
}

extern "C" {

int tck_launch_ic_fill(const tc_geom_t * g, double * dst, const double * prof, const double * mx,
    const double * my, const double * mz, double amp, double mode_amp, unsigned long long seed,
    unsigned field, int nxs, int i0, hipStream_t stream){
  IcArgs A;
  A.g = *g; A.dst = dst; A.prof = prof; A.mx = mx; A.my = my; A.mz = mz; A.amp = amp;
  A.mode_amp = mode_amp; A.seed = seed; A.field = field; A.nxs = nxs; A.i0 = i0;
  hipLaunchKernelGGL(tck_ic_fill, dim3((nxs + 63) / 64, g->ny, g->nz), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

/* n doubles (even); chunks: 16 B chunks per lane, 1, 2, 4 or 8 */
int tck_launch_copy(const double * src, double * dst, long long n, int chunks, hipStream_t stream){
  const long long n2 = n / 2;
  switch(chunks){
    case 1: launch_copy<1>(src, dst, n2, stream); break;
    case 2: launch_copy<2>(src, dst, n2, stream); break;
    case 4: launch_copy<4>(src, dst, n2, stream); break;
    case 8: launch_copy<8>(src, dst, n2, stream); break;
    default: return (int)hipErrorInvalidValue;
  }
  TCK_CHECK_LAUNCH();
  return 0;
}

}
