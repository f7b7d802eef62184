/*
 * tck_fft.hip -- batched FFTs of the Poisson solve (gfx950), replacing the FFTW plans of
 * src/fluid/compute_potential.c:247-309 and the pencil transposes around them (:517-581).
 *
 * Spectrum layout in HBM: S[kz][ky][i] (complex, i fastest, row pitch nx), ky = 0..ny/2.
 * With i innermost, the theta and z transforms read and write rows of a few consecutive i
 * (coalesced), and the radial solve reads each (ky, kz) line contiguously: no transpose
 * pass is needed anywhere.
 *
 *  tck_div_yfft   assign_input (:382-425: divergence / (gamma dt ny nz)) fused with the
 *                 real-to-complex theta transform; a real length-N transform is done as a
 *                 complex length-N/2 transform of (x_2m + i x_2m+1) plus the usual split.
 *  tck_zfft       complex axial transforms, forward or backward, in place.
 *  tck_yifft_psi  complex-to-real theta transform fused with extract_output (:427-443): psi
 *                 with its periodic halos and Neumann wall copies (boundary/psi.c:13-45).
 * Each has a persistent form (suffix _p) for FP64 tiles of long transforms: a block loops
 * over tiles and issues part of the next tile's loads before the current tile's LDS passes.
 * Tiles whose LDS leaves few blocks per CU run with half the elements per thread (twice the
 * threads), so that a CU still holds 16 waves.  Every form computes the same bits.
 * Transforms are unnormalised, as FFTW's; the imaginary parts of the DC and Nyquist bins are
 * ignored by the complex-to-real transform, as FFTW's c2r does with a Hermitian input.
 *
 * The in-LDS transform is a Stockham autosort FFT (Govindaraju et al., "High performance
 * discrete Fourier transforms on graphics processors", SC 2008) with radix 8, 4, 3 and 2
 * passes; each pass reads all its butterfly inputs into registers, synchronises, and
 * writes in place.
 */
#include "tck_common.h"

#define FFT_MAXST 16
#define FFT_EPT 16   /* complex elements per thread per pass */
#define FFT_MAXTHREADS 1024
#define FFT_G 4      /* one-tile transforms: divergence elements whose loads are issued together */
#define FFT_GZ 8     /* one-tile transforms: spectrum elements whose loads are issued together */

/* complex pair of the transform's arithmetic type T (double, or float in the mixed step) */
template <typename T> struct cvec;
template <> struct cvec<double> { typedef double2 type; };
template <> struct cvec<float> { typedef float2 type; };
template <typename T> __device__ __forceinline__ typename cvec<T>::type mk2(T x, T y);
template <> __device__ __forceinline__ double2 mk2<double>(double x, double y){ return make_double2(x, y); }
template <> __device__ __forceinline__ float2 mk2<float>(float x, float y){ return make_float2(x, y); }

struct FftPlan {
  int M;
  int nst;
  int rad[FFT_MAXST];
  /* per stage: multipliers that divide by M / rad and by the stage's span Ns with one high
   * multiply (floor(2^32 / d) + 1, exact for dividends below 2^16; 0 for d = 1) */
  unsigned mper[FFT_MAXST], mns[FFT_MAXST];
  /* threads per block of the launch.  The kernels read the block size from here (a kernel
   * argument, in scalar registers) rather than from blockDim.x, whose 16-bit read from the
   * dispatch packet is a vector memory load that the compiler may repeat inside the passes:
   * a wait for it there would also wait for every global load and store issued before it. */
  int nt;
};

/* log2 of a tile width (tile widths are powers of two, at most 64) */
__device__ __forceinline__ int fft_lg(int x){ return __builtin_ctz((unsigned)x); }

/* x / d for the multiplier m of d (see FftPlan) */
__device__ __forceinline__ unsigned fft_div(unsigned x, unsigned m){
  return m ? __umulhi(x, m) : x;
}

template <typename V> __device__ __forceinline__ V cmul(V a, V b){
  V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b){ V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __device__ __forceinline__ V csub(V a, V b){ V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
/* s i a */
template <typename V, typename T> __device__ __forceinline__ V crot(V a, T s){ V r; r.x = - s * a.y; r.y = s * a.x; return r; }
template <typename V> __device__ __forceinline__ V cconj(V a){ V r; r.x = a.x; r.y = - a.y; return r; }

/* R-point DFTs, y_k = sum_n x_n exp(s 2 pi i n k / R) */
template <int R, typename T> struct dftr;

template <typename T> struct dftr<2, T> {
  typedef typename cvec<T>::type V;
  static __device__ __forceinline__ void run(V * x, T s){
    (void)s;
    const V a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <typename T> struct dftr<4, T> {
  typedef typename cvec<T>::type V;
  static __device__ __forceinline__ void run(V * x, T s){
    const V a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    const V b0 = cadd(x[1], x[3]), b1 = crot(csub(x[1], x[3]), s);
    x[0] = cadd(a0, b0);
    x[2] = csub(a0, b0);
    x[1] = cadd(a1, b1);
    x[3] = csub(a1, b1);
  }
};

template <typename T> struct dftr<3, T> {
  typedef typename cvec<T>::type V;
  static __device__ __forceinline__ void run(V * x, T s){
    const T h = T(0.86602540378443864676);   /* sqrt(3) / 2 */
    const V t1 = cadd(x[1], x[2]);
    const V t2 = crot(csub(x[1], x[2]), s * h);
    const V m = mk2<T>(x[0].x - T(0.5) * t1.x, x[0].y - T(0.5) * t1.y);
    x[0] = cadd(x[0], t1);
    x[1] = cadd(m, t2);
    x[2] = csub(m, t2);
  }
};

/* odd R-point DFT by conjugate pairs (R = 5, 7): with a_n = x_n + x_{R-n} and
 * b_n = x_n - x_{R-n}, n = 1..h = (R-1)/2,
 *   y_0 = x_0 + sum a_n,  y_k, y_{R-k} = x_0 + sum_n a_n cos(2 pi n k / R)
 *                                     +- s i sum_n b_n sin(2 pi n k / R),  k = 1..h */
template <int R> struct oddtab;
template <> struct oddtab<5> {
  static __device__ __forceinline__ double c(int m){
    return m == 1 ? 0.3090169943749474241023 : -0.8090169943749474241023;
  }
  static __device__ __forceinline__ double sn(int m){
    return m == 1 ? 0.9510565162951535721164 : 0.5877852522924731291687;
  }
};
template <> struct oddtab<7> {
  static __device__ __forceinline__ double c(int m){
    return m == 1 ? 0.6234898018587335305250 : m == 2 ? -0.2225209339563144042889 : -0.9009688679024191262361;
  }
  static __device__ __forceinline__ double sn(int m){
    return m == 1 ? 0.7818314824680298087084 : m == 2 ? 0.9749279121818236070181 : 0.4338837391175581204758;
  }
};

template <int R, typename T> struct dftodd {
  typedef typename cvec<T>::type V;
  static __device__ __forceinline__ void run(V * x, T s){
    constexpr int h = (R - 1) / 2;
    V a[h], b[h];
    #pragma unroll
    for(int n = 1; n <= h; n++){ a[n - 1] = cadd(x[n], x[R - n]); b[n - 1] = csub(x[n], x[R - n]); }
    V y0 = x[0];
    #pragma unroll
    for(int n = 0; n < h; n++) y0 = cadd(y0, a[n]);
    V re[h], im[h];
    #pragma unroll
    for(int k = 1; k <= h; k++){
      V r = x[0];
      V q = mk2<T>(T(0.), T(0.));
      #pragma unroll
      for(int n = 1; n <= h; n++){
        /* m = n k mod R, folded into 1..h with the sign of the sine */
        const int m0 = (n * k) % R;
        const int m = m0 <= h ? m0 : R - m0;
        const T cv = T(oddtab<R>::c(m));
        const T sv = T(m0 <= h ? oddtab<R>::sn(m) : - oddtab<R>::sn(m));
        r = mk2<T>(r.x + cv * a[n - 1].x, r.y + cv * a[n - 1].y);
        q = mk2<T>(q.x + sv * b[n - 1].x, q.y + sv * b[n - 1].y);
      }
      re[k - 1] = r;
      im[k - 1] = crot(q, s);
    }
    x[0] = y0;
    #pragma unroll
    for(int k = 1; k <= h; k++){
      x[k] = cadd(re[k - 1], im[k - 1]);
      x[R - k] = csub(re[k - 1], im[k - 1]);
    }
  }
};
template <typename T> struct dftr<5, T> : dftodd<5, T> {};
template <typename T> struct dftr<7, T> : dftodd<7, T> {};

template <typename T> struct dftr<8, T> {
  typedef typename cvec<T>::type V;
  static __device__ __forceinline__ void run(V * x, T s){
    const T r = T(0.70710678118654752440);   /* sqrt(1/2) */
    V e[4] = {x[0], x[2], x[4], x[6]};
    V o[4] = {x[1], x[3], x[5], x[7]};
    dftr<4, T>::run(e, s);
    dftr<4, T>::run(o, s);
    /* o_k *= w8^k, w8 = r (1 + s i) */
    o[1] = mk2<T>(r * (o[1].x - s * o[1].y), r * (o[1].y + s * o[1].x));
    o[2] = crot(o[2], s);
    o[3] = mk2<T>(r * (- o[3].x - s * o[3].y), r * (- o[3].y + s * o[3].x));
    #pragma unroll
    for(int k = 0; k < 4; k++){
      x[k] = cadd(e[k], o[k]);
      x[k + 4] = csub(e[k], o[k]);
    }
  }
};

/* block barrier that orders the LDS accesses only: global loads issued before it stay in
 * flight and pending global stores are not waited for */
__device__ __forceinline__ void fft_lds_barrier(){
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

/* the same, and no instruction is scheduled across it: the loads of the pass after it (its
 * twiddle factors) are not hoisted into the pass before it, which would hold more registers */
__device__ __forceinline__ void fft_pass_barrier(){
  __builtin_amdgcn_sched_barrier(0);
  fft_lds_barrier();
  __builtin_amdgcn_sched_barrier(0);
}

/* padded position of element n of a line in LDS: one spare slot after every 8 elements, so
 * that the first Stockham pass, which writes elements R apart from consecutive threads, and
 * the row-wise fills spread over all banks */
__host__ __device__ __forceinline__ int fpad(int n){ return n + (n >> 3); }

template <int R, typename T, int EPT = FFT_EPT>
__device__ __forceinline__ void fft_stage(typename cvec<T>::type * buf, int M, int ld, int nb, int Ns,
    const typename cvec<T>::type * __restrict__ tw, T s, unsigned mper, unsigned mns, int nt){
  typedef typename cvec<T>::type V;
  constexpr int BPT = (EPT + R - 1) / R;
  const int tid = threadIdx.x;
  const int per = M / R;
  const int nbf = nb * per;
  const int twstep = M / (Ns * R);
  V v[BPT][R];
  #pragma unroll
  for(int q = 0; q < BPT; q++){
    const int bf = tid + q * nt;
    if(bf < nbf){
      const int b = (int)fft_div((unsigned)bf, mper);
      const int jj = bf - b * per;
      const int jm = jj - (int)fft_div((unsigned)jj, mns) * Ns;
      const V * src = buf + b * ld;
      #pragma unroll
      for(int r = 0; r < R; r++){
        V x = src[fpad(jj + r * per)];
        if(r > 0){
          V w = tw[jm * r * twstep];
          if(s > T(0.)) w.y = - w.y;
          x = cmul(x, w);
        }
        v[q][r] = x;
      }
      dftr<R, T>::run(v[q], s);
    }
  }
  fft_pass_barrier();
  #pragma unroll
  for(int q = 0; q < BPT; q++){
    const int bf = tid + q * nt;
    if(bf < nbf){
      const int b = (int)fft_div((unsigned)bf, mper);
      const int jj = bf - b * per;
      const int jm = jj - (int)fft_div((unsigned)jj, mns) * Ns;
      V * dst = buf + b * ld;
      const int d0 = (jj - jm) * R + jm;
      #pragma unroll
      for(int r = 0; r < R; r++) dst[fpad(d0 + r * Ns)] = v[q][r];
    }
  }
  fft_pass_barrier();
}

/* nb transforms of length P.M stored at buf + b * ld; tw[t] = exp(-2 pi i t / M); each pass
 * keeps at most EPT elements per thread in registers, so blockDim.x * EPT >= nb * P.M.  W:
 * the plan may hold radix 5 and 7 passes (compiled only into the kernels of such lengths, so
 * that the 2^a 3^b kernels keep their register budget) */
template <typename T, int EPT = FFT_EPT, bool W = false>
__device__ void lds_fft(typename cvec<T>::type * buf, int ld, int nb, const FftPlan & P,
    const typename cvec<T>::type * __restrict__ tw, T s){
  int Ns = 1;
  for(int st = 0; st < P.nst; st++){
    const int R = P.rad[st];
    const unsigned mp = P.mper[st], mn = P.mns[st];
    if(R == 8) fft_stage<8, T, EPT>(buf, P.M, ld, nb, Ns, tw, s, mp, mn, P.nt);
    else if(R == 4) fft_stage<4, T, EPT>(buf, P.M, ld, nb, Ns, tw, s, mp, mn, P.nt);
    else if(R == 3) fft_stage<3, T, EPT>(buf, P.M, ld, nb, Ns, tw, s, mp, mn, P.nt);
    else if(W && R == 5) fft_stage<5, T, EPT>(buf, P.M, ld, nb, Ns, tw, s, mp, mn, P.nt);
    else if(W && R == 7) fft_stage<7, T, EPT>(buf, P.M, ld, nb, Ns, tw, s, mp, mn, P.nt);
    else if(R == 2) fft_stage<2, T, EPT>(buf, P.M, ld, nb, Ns, tw, s, mp, mn, P.nt);
    Ns *= R;
  }
}

__constant__ double c_fft_rka[3] = {32. / 60., 25. / 60., 45. / 60.};
__constant__ double c_fft_rkb[3] = { 0. / 60., -17. / 60., -25. / 60.};

/* tile of this block among ntiles: blocks that run on the same XCD (consecutive slots of the
 * round-robin dispatch) take consecutive tiles, whose rows share 128 B lines in its L2 */
__device__ __forceinline__ long long fft_tile_of(long long ntiles){
  long long t = blockIdx.x;
  if(ntiles % TCK_NXCD == 0) t = (t % TCK_NXCD) * (ntiles / TCK_NXCD) + t / TCK_NXCD;
  return t;
}

/* ------------------------------------------------------------------------------------- */
struct YFwdArgs {
  tc_geom_t g;
  tc_coef_t c;
  const double * ux; const double * uy; const double * uz;
  const double * dt;
  int rkstep;
  void * S;                      /* spectrum of the transform type */
  int nky, ti, ld;               /* ld: LDS row pitch of a padded line (fft_ld) */
  FftPlan P;                     /* length M = ny / 2 */
  const void * twM;              /* exp(-2 pi i t / M) */
  const void * twN;              /* exp(-2 pi i k / ny), k = 0..M */
};

/* the divergence is formed in FP64 (assign_input :404-423) and rounded to the transform
 * type T only when it enters the transform */
template <typename T, bool W>
__global__ void __launch_bounds__(FFT_MAXTHREADS) tck_div_yfft(const YFwdArgs A){
  typedef typename cvec<T>::type V;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  V * sm = (V *)fft_smem;
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  const int ti = A.ti;
  const int M = A.P.M;
  const int ntx = (g.nx + ti - 1) / ti;
  const long long tile = fft_tile_of((long long)ntx * g.nz);
  const int i0 = 1 + (int)(tile % ntx) * ti;
  const int k = 1 + (int)(tile / ntx);
  const double dt = *A.dt;
  const double gamma = c_fft_rka[A.rkstep] + c_fft_rkb[A.rkstep];
  const double norm = 1. * g.ny * g.nz;
  const double prefactor = 1. / (gamma * dt) / norm;
  const long long sy = g.px;
  const long long sz = g.plane;
  const V * twM = (const V *)A.twM;
  const V * twN = (const V *)A.twN;
  /* divergence packed as z_m = x_2m + i x_2m+1; the velocities of FFT_G elements are loaded
   * together (out-of-field elements read cell (1, 1) of the plane and are put as 0, so that no
   * load is conditional) */
  const int nel = ti * g.ny;
  for(int e0 = threadIdx.x; e0 < nel; e0 += FFT_G * A.P.nt){
    double P[FFT_G][6];
    #pragma unroll
    for(int u = 0; u < FFT_G; u++){
      const int e = e0 + u * A.P.nt;
      const int t = (e & (ti - 1)), j0 = (e >> fft_lg(ti)), i = i0 + t;
      const bool ok = e < nel && i <= g.nx;
      const long long o = tck_idx(g, ok ? i : 1, ok ? j0 + 1 : 1, k);
      P[u][0] = A.ux[o]; P[u][1] = A.ux[o + 1]; P[u][2] = A.uy[o]; P[u][3] = A.uy[o + sy];
      P[u][4] = A.uz[o]; P[u][5] = A.uz[o + sz];
    }
    #pragma unroll
    for(int u = 0; u < FFT_G; u++){
      const int e = e0 + u * A.P.nt;
      if(e >= nel) break;
      const int t = (e & (ti - 1)), j0 = (e >> fft_lg(ti)), i = i0 + t;
      double val = 0.;
      if(i <= g.nx){
        const double div = c.ijdxc[i] * (
            - c.jhxf[i] * P[u][0] + c.jhxf[i + 1] * P[u][1]
            - c.jhyc[i] * P[u][2] + c.jhyc[i] * P[u][3]
            - c.jhzc[i] * P[u][4] + c.jhzc[i] * P[u][5]);
        val = prefactor * div;
      }
      ((T *)(sm + t * A.ld))[2 * fpad(j0 >> 1) + (j0 & 1)] = T(val);
    }
  }
  __syncthreads();
  lds_fft<T, FFT_EPT, W>(sm, A.ld, ti, A.P, twM, T(-1.));
  /* split: X_k = (Z_k + conj Z_{M-k}) / 2 - i W^k (Z_k - conj Z_{M-k}) / 2, k = 0..M */
  V * S = (V *)A.S;
  for(int e = threadIdx.x; e < ti * (M + 1); e += A.P.nt){
    const int t = (e & (ti - 1));
    const int kk = (e >> fft_lg(ti));
    const int i = i0 + t;
    if(i > g.nx) continue;
    const V zk = sm[t * A.ld + fpad(kk == M ? 0 : kk)];
    const V zc = cconj(sm[t * A.ld + fpad(kk == 0 ? 0 : M - kk)]);
    const V w = twN[kk];
    const V ev = mk2<T>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
    const V od = mk2<T>(T(0.5) * (zk.x - zc.x), T(0.5) * (zk.y - zc.y));
    /* - i w od */
    const V wo = cmul(w, od);
    const V X = mk2<T>(ev.x + wo.y, ev.y - wo.x);
    tck_st(&S[(long long)(i - 1) + (long long)g.nx * ((long long)kk + (long long)A.nky * (k - 1))], X);
  }
}

/* Persistent form (compiled for EPTK elements per thread per pass): each block loops over
 * tiles and loads the velocities of the first PF of its divergence elements of the next tile
 * before the current tile's passes. */
#ifndef TCK_DIVY_PF
#define TCK_DIVY_PF 1
#endif
#ifndef TCK_DIVY_KRUN
#define TCK_DIVY_KRUN 1
#endif

template <typename T, int PF, int EPTK, bool W, bool TL>
__global__ void __launch_bounds__(FFT_MAXTHREADS) tck_div_yfft_p(const YFwdArgs A){
  typedef typename cvec<T>::type V;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  V * sm = (V *)fft_smem;
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  const int ti = A.ti;
  const int M = A.P.M;
  const int nt = A.P.nt;
  const int ntx = (g.nx + ti - 1) / ti;
  const int ntiles = ntx * g.nz;
  const double dt = *A.dt;
  const double gamma = c_fft_rka[A.rkstep] + c_fft_rkb[A.rkstep];
  const double norm = 1. * g.ny * g.nz;
  const double prefactor = 1. / (gamma * dt) / norm;
  const long long sy = g.px;
  const long long sz = g.plane;
  const V * twM = (const V *)A.twM;
  const V * twN = (const V *)A.twN;
  if constexpr (TL){
    /* TL: both twiddle tables copied to LDS after the tile's lines, so that neither the passes
     * nor the split below wait for a global load (a wait for a load also waits for the stores
     * issued before it: in the split, each element's stores would wait for the last's) */
    V * twl = sm + ti * A.ld;
    for(int t = threadIdx.x; t < M; t += nt) twl[t] = twM[t];
    for(int t = threadIdx.x; t <= M; t += nt) twl[M + t] = twN[t];
    twM = twl;
    twN = twl + M;
  }
  V * S = (V *)A.S;
  /* element q of this thread in a tile: line t0, row j = r0 + q dr (the block size is a
   * multiple of the tile width, a power of two no larger than 64) */
  const int t0i = (threadIdx.x & (ti - 1)), r0i = (threadIdx.x >> fft_lg(ti)), dr = (nt >> fft_lg(ti));
  int t0 = t0i, r0 = r0i;
  double pre[PF][6];
  bool pok[PF];
  /* the six velocities of element q of tile tl: ok when inside the field; otherwise the loads
   * read an in-bounds stand-in (cell i = 1, j = 1 of the tile's plane) and the element is put
   * as 0, so that no load is conditional */
  #define DV_LOAD(tl, q, P, ok) { \
    const int j_ = r0 + (q) * dr; \
    const int i_ = 1 + ((tl) % ntx) * ti + t0; \
    ok = j_ < g.ny && i_ <= g.nx; \
    const long long o_ = tck_idx(g, ok ? i_ : 1, ok ? j_ + 1 : 1, 1 + (tl) / ntx); \
    P[0] = A.ux[o_]; P[1] = A.ux[o_ + 1]; P[2] = A.uy[o_]; P[3] = A.uy[o_ + sy]; \
    P[4] = A.uz[o_]; P[5] = A.uz[o_ + sz]; }
  /* elements per thread of a tile (the launch guarantees that they cover it), and how many
   * are loaded together */
  constexpr int EPT2 = 2 * EPTK;
  constexpr int G = 2;
#if TCK_DIVY_KRUN
  /* The same with uz's lower face taken from uzc when cz: in a block's run of planes the next
   * tile is the same column one plane up, whose lower faces are this tile's upper faces, held
   * in registers (the same values, so the same bits) instead of read again */
  double uzc[EPT2];
  #define DV_LOADC(tl, q, P, ok, cz) { \
    const int j_ = r0 + (q) * dr; \
    const int i_ = 1 + ((tl) % ntx) * ti + t0; \
    ok = j_ < g.ny && i_ <= g.nx; \
    const long long o_ = tck_idx(g, ok ? i_ : 1, ok ? j_ + 1 : 1, 1 + (tl) / ntx); \
    P[0] = A.ux[o_]; P[1] = A.ux[o_ + 1]; P[2] = A.uy[o_]; P[3] = A.uy[o_ + sy]; \
    if(cz) P[4] = uzc[q]; else P[4] = A.uz[o_]; \
    P[5] = A.uz[o_ + sz]; }
  /* each block walks a contiguous run of planes of one column of tiles, so that the uz
   * plane a tile reads as its upper face is read again as the next tile's lower face soon
   * after, from cache; tile index tile = tx + ntx (k - 1) as in the plain order */
  const int per = (ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
  const int c0 = blockIdx.x * per, c1 = min(c0 + per, ntiles);
  #define DV_TILE(c) (((c) / g.nz) + ntx * ((c) % g.nz))
  int cc = c0;
  if(cc < c1){
    #pragma unroll
    for(int q = 0; q < PF; q++){ DV_LOAD(DV_TILE(cc), q, pre[q], pok[q]); }
  }
  for(; cc < c1; cc++){
    const int tile = DV_TILE(cc);
    const int tnext = cc + 1 < c1 ? DV_TILE(cc + 1) : ntiles;
    /* this tile's lower faces in uzc (the previous tile was the plane below), and the next
     * tile's (it is the plane above); not with radix 5 and 7 plans, whose passes need the
     * registers */
    const bool cz = !W && cc > c0 && tile == DV_TILE(cc - 1) + ntx;
    const bool czn = !W && tnext == tile + ntx;
#else
  int tile = blockIdx.x;
  if(tile < ntiles){
    #pragma unroll
    for(int q = 0; q < PF; q++){ DV_LOAD(tile, q, pre[q], pok[q]); }
  }
  for(; tile < ntiles; tile += gridDim.x){
    const int tnext = tile + (int)gridDim.x;
#endif
    /* the element offsets are recomputed in every tile rather than held across the passes */
    t0 = t0i; r0 = r0i;
    asm volatile("" : "+v"(t0), "+v"(r0));
    const int i0 = 1 + (tile % ntx) * ti;
    const int k = 1 + tile / ntx;
    /* divergence (assign_input :404-423, same operand order as tck_div_yfft) packed as
     * z_m = x_2m + i x_2m+1 */
    const int i = i0 + t0;
    auto put = [&](const int q, const double * P, const bool ok){
      const int j0 = r0 + q * dr;
      if(j0 >= g.ny) return;
      double val = 0.;
      if(ok){
        const double div = c.ijdxc[i] * (
            - c.jhxf[i] * P[0] + c.jhxf[i + 1] * P[1]
            - c.jhyc[i] * P[2] + c.jhyc[i] * P[3]
            - c.jhzc[i] * P[4] + c.jhzc[i] * P[5]);
        val = prefactor * div;
      }
      ((T *)(sm + t0 * A.ld))[2 * fpad(j0 >> 1) + (j0 & 1)] = T(val);
    };
    /* the rest of the tile, G elements (6 G loads) issued before the first is waited for */
    #pragma unroll
    for(int q0 = PF; q0 < EPT2; q0 += G){
      double P[G][6];
      bool ok[G];
#if TCK_DIVY_KRUN
      #pragma unroll
      for(int u = 0; u < G; u++) if(q0 + u < EPT2){ DV_LOADC(tile, q0 + u, P[u], ok[u], cz); }
#else
      #pragma unroll
      for(int u = 0; u < G; u++) if(q0 + u < EPT2){ DV_LOAD(tile, q0 + u, P[u], ok[u]); }
#endif
      if(q0 == PF){
        #pragma unroll
        for(int q = 0; q < PF; q++){
          put(q, pre[q], pok[q]);
#if TCK_DIVY_KRUN
          uzc[q] = pre[q][5];
#endif
        }
      }
      #pragma unroll
      for(int u = 0; u < G; u++) if(q0 + u < EPT2){
        put(q0 + u, P[u], ok[u]);
#if TCK_DIVY_KRUN
        uzc[q0 + u] = P[u][5];
#endif
      }
    }
    fft_lds_barrier();
    if(tnext < ntiles){
#if TCK_DIVY_KRUN
      #pragma unroll
      for(int q = 0; q < PF; q++){ DV_LOADC(tnext, q, pre[q], pok[q], czn); }
#else
      #pragma unroll
      for(int q = 0; q < PF; q++){ DV_LOAD(tnext, q, pre[q], pok[q]); }
#endif
    }
    lds_fft<T, EPTK, W>(sm, A.ld, ti, A.P, twM, T(-1.));
    for(int e = threadIdx.x; e < ti * (M + 1); e += nt){
      const int t = (e & (ti - 1));
      const int kk = (e >> fft_lg(ti));
      const int ii = i0 + t;
      if(ii > g.nx) continue;
      const V zk = sm[t * A.ld + fpad(kk == M ? 0 : kk)];
      const V zc = cconj(sm[t * A.ld + fpad(kk == 0 ? 0 : M - kk)]);
      const V w = twN[kk];
      const V ev = mk2<T>(T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y));
      const V od = mk2<T>(T(0.5) * (zk.x - zc.x), T(0.5) * (zk.y - zc.y));
      const V wo = cmul(w, od);
      const V Xk = mk2<T>(ev.x + wo.y, ev.y - wo.x);
      tck_st(&S[(long long)(ii - 1) + (long long)g.nx * ((long long)kk + (long long)A.nky * (k - 1))], Xk);
    }
    fft_lds_barrier();
  }
  #undef DV_LOAD
#if TCK_DIVY_KRUN
  #undef DV_LOADC
#endif
#if TCK_DIVY_KRUN
  #undef DV_TILE
#endif
}

/* ------------------------------------------------------------------------------------- */
struct ZArgs {
  int nx, nky, nz, tz, ld;
  void * S;
  FftPlan P;               /* length nz */
  const void * tw;
  double sign;
};

/* one block per tile of tz consecutive i at one ky */
template <typename T, bool W>
__global__ void __launch_bounds__(FFT_MAXTHREADS) tck_zfft(const ZArgs A){
  typedef typename cvec<T>::type V;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  V * sm = (V *)fft_smem;
  const int tz = A.tz;
  const int nz = A.nz;
  const int ntx = (A.nx + tz - 1) / tz;
  const long long tile = fft_tile_of((long long)ntx * A.nky);
  const int i0 = (int)(tile % ntx) * tz;
  const int ky = (int)(tile / ntx);
  const long long kstride = (long long)A.nx * A.nky;
  V * base = (V *)A.S + (long long)A.nx * ky;
  /* FFT_GZ elements loaded together; out-of-range ones read element 0 and are put as 0 */
  const int nel = tz * nz;
  for(int e0 = threadIdx.x; e0 < nel; e0 += FFT_GZ * A.P.nt){
    V v[FFT_GZ];
    #pragma unroll
    for(int u = 0; u < FFT_GZ; u++){
      const int e = e0 + u * A.P.nt;
      const int t = (e & (tz - 1)), kz = (e >> fft_lg(tz)), ii = i0 + t;
      v[u] = tck_ld((e < nel && ii < A.nx) ? &base[ii + kstride * kz] : (const V *)A.S);
    }
    #pragma unroll
    for(int u = 0; u < FFT_GZ; u++){
      const int e = e0 + u * A.P.nt;
      if(e >= nel) break;
      const int t = (e & (tz - 1)), kz = (e >> fft_lg(tz));
      sm[t * A.ld + fpad(kz)] = i0 + t < A.nx ? v[u] : mk2<T>(T(0.), T(0.));
    }
  }
  __syncthreads();
  lds_fft<T, FFT_EPT, W>(sm, A.ld, tz, A.P, (const V *)A.tw, T(A.sign));
  for(int e = threadIdx.x; e < tz * nz; e += A.P.nt){
    const int t = (e & (tz - 1));
    const int kz = (e >> fft_lg(tz));
    const int ii = i0 + t;
    if(ii < A.nx) tck_st(&base[ii + kstride * kz], sm[t * A.ld + fpad(kz)]);
  }
}

/* Persistent form: each block loops over tiles and issues the loads of the first PF of its
 * elements of the next tile before the current tile's transform, so that part of the HBM
 * traffic overlaps the LDS passes; the remaining elements are loaded when the tile starts, all
 * issued before the first is waited for.  Compiled for EPTK elements per thread per pass: the
 * launch guarantees that blockDim.x EPTK covers the tile and that blockDim.x is a multiple of
 * the tile width. */
#ifndef TCK_ZFFT_PF
#define TCK_ZFFT_PF 2            /* elements per thread loaded one tile ahead */
#endif

template <typename T, int PF, int EPTK, bool W, bool X>
__global__ void __launch_bounds__(FFT_MAXTHREADS) tck_zfft_p(const ZArgs A){
  typedef typename cvec<T>::type V;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  V * sm = (V *)fft_smem;
  const int tz = A.tz, nz = A.nz, nt = A.P.nt;
  const int ntx = (A.nx + tz - 1) / tz;
  const int ntiles = ntx * A.nky;
  const long long kstride = (long long)A.nx * A.nky;
  /* element q of this thread in a tile: line t0, row kz = r0 + q dr (the block size is a
   * multiple of the tile width, a power of two no larger than 64) */
  const int t0i = (threadIdx.x & (tz - 1)), r0i = (threadIdx.x >> fft_lg(tz)), dr = (nt >> fft_lg(tz));
  if constexpr (X){
    /* Exact form: the tiles cover the spectrum exactly (nx a multiple of tz, blockDim.x EPTK =
     * tz nz), so no access is conditional; the twiddle factors are read from LDS (after the
     * tile's lines), so that the passes issue no global loads; and every element of the next
     * tile is loaded before the current tile's passes and stores.  Outstanding global loads and
     * stores complete in issue order as far as the wait counter is concerned, so a wait for a
     * load issued after a store also waits for the store: here the loads of tile n+1 are
     * issued before the stores of tile n, and the put of tile n+1 waits only for its loads.
     * The stores into a zero-size buffer before the loop (they write nothing) give the loop's
     * entry the same outstanding loads-then-stores sequence as the end of an iteration, so that
     * the compiler's wait at the put is the same counted wait on both paths into it. */
    V * twl = sm + tz * A.ld;
    for(int t = threadIdx.x; t < A.P.M; t += nt) twl[t] = ((const V *)A.tw)[t];
    int tile = blockIdx.x;             /* the launch has at least four tiles per block */
    V pre[EPTK];
    #define ZX_OFF(tl, q) ((long long)A.nx * ((tl) / ntx) + (long long)(((tl) % ntx) * tz + t0i) + kstride * (long long)(r0i + (q) * dr))
    /* FP32: the spectrum spans under 4 GB (checked on the host), so a tile's accesses are a
     * buffer resource at the tile's first column plus 32-bit offsets: one register of address
     * per thread instead of one 64-bit address per element */
    const unsigned vz = (unsigned)sizeof(V) * (unsigned)(t0i + kstride * r0i);
    const unsigned sq = (unsigned)sizeof(V) * (unsigned)(kstride * dr);
    #define ZX_RS(tl) tck_rsrc_u((const V *)A.S + (long long)A.nx * ((tl) / ntx) + (long long)(((tl) % ntx) * tz), \
      (unsigned)sizeof(V) * (unsigned)(kstride * nz))
    if constexpr (sizeof(T) == 4){
      const __amdgpu_buffer_rsrc_t rs = ZX_RS(tile);
      #pragma unroll
      for(int q = 0; q < EPTK; q++) pre[q] = tck_bldf2(rs, vz, sq * q);
    }else{
      #pragma unroll
      for(int q = 0; q < EPTK; q++) pre[q] = tck_ld((const V *)A.S + ZX_OFF(tile, q));
    }
    {
      const __amdgpu_buffer_rsrc_t nul = tck_rsrc(A.S, 0u);
      #pragma unroll
      for(int q = 0; q < EPTK; q++) tck_bst0(nul, 16u * q);
    }
    for(;;){
      #pragma unroll
      for(int q = 0; q < EPTK; q++) sm[t0i * A.ld + fpad(r0i + q * dr)] = pre[q];
      fft_lds_barrier();
      const int nxt = tile + (int)gridDim.x < ntiles ? tile + (int)gridDim.x : tile;
      if constexpr (sizeof(T) == 4){
        const __amdgpu_buffer_rsrc_t rs = ZX_RS(nxt);
        #pragma unroll
        for(int q = 0; q < EPTK; q++) pre[q] = tck_bldf2(rs, vz, sq * q);
      }else{
        #pragma unroll
        for(int q = 0; q < EPTK; q++) pre[q] = tck_ld((const V *)A.S + ZX_OFF(nxt, q));
      }
      lds_fft<T, EPTK, W>(sm, A.ld, tz, A.P, twl, T(A.sign));
      if constexpr (sizeof(T) == 4){
        const __amdgpu_buffer_rsrc_t rs = ZX_RS(tile);
        #pragma unroll
        for(int q = 0; q < EPTK; q++) tck_bstf2(rs, vz, sq * q, sm[t0i * A.ld + fpad(r0i + q * dr)]);
      }else{
        V * base = (V *)A.S + (long long)A.nx * (tile / ntx) + (long long)((tile % ntx) * tz + t0i);
        #pragma unroll
        for(int q = 0; q < EPTK; q++){
          const int kz = r0i + q * dr;
          tck_st(&base[kstride * kz], sm[t0i * A.ld + fpad(kz)]);
        }
      }
      fft_lds_barrier();
      tile += (int)gridDim.x;
      if(tile >= ntiles) break;
    }
    #undef ZX_OFF
    #undef ZX_RS
    return;
  }
  int t0 = t0i, r0 = r0i;
  V pre[PF];
  bool pok[PF];
  int tile = blockIdx.x;
  /* element q of tile tl: ok when inside the spectrum; otherwise the load reads element 0
   * (in bounds) and the value is replaced by 0 when put, so that no load is conditional */
  #define ZF_ELEM(tl, q, out, ok) { \
    const int kz_ = r0 + (q) * dr; \
    const int tx_ = (tl) % ntx; \
    const int ii_ = tx_ * tz + t0; \
    ok = kz_ < nz && ii_ < A.nx; \
    const V * b_ = (const V *)A.S + (long long)A.nx * ((tl) / ntx); \
    out = tck_ld(ok ? &b_[ii_ + kstride * kz_] : (const V *)A.S); }
  #define ZF_PUT(q, v, ok) { const int kz_ = r0 + (q) * dr; \
    if(kz_ < nz) sm[t0 * A.ld + fpad(kz_)] = (ok) ? (v) : mk2<T>(T(0.), T(0.)); }
  if(tile < ntiles){
    #pragma unroll
    for(int q = 0; q < PF; q++){ ZF_ELEM(tile, q, pre[q], pok[q]); }
  }
  for(; tile < ntiles; tile += gridDim.x){
    /* the element offsets are recomputed in every tile rather than held across the passes */
    t0 = t0i; r0 = r0i;
    asm volatile("" : "+v"(t0), "+v"(r0));
    /* the rest of the tile: all loads issued before the first is waited for */
    V cur[EPTK - PF > 0 ? EPTK - PF : 1];
    bool cok[EPTK - PF > 0 ? EPTK - PF : 1];
    #pragma unroll
    for(int q = PF; q < EPTK; q++){ ZF_ELEM(tile, q, cur[q - PF], cok[q - PF]); }
    #pragma unroll
    for(int q = 0; q < PF; q++){ ZF_PUT(q, pre[q], pok[q]); }
    #pragma unroll
    for(int q = PF; q < EPTK; q++){ ZF_PUT(q, cur[q - PF], cok[q - PF]); }
    fft_lds_barrier();
    if(tile + (int)gridDim.x < ntiles){
      #pragma unroll
      for(int q = 0; q < PF; q++){ ZF_ELEM(tile + (int)gridDim.x, q, pre[q], pok[q]); }
    }
    lds_fft<T, EPTK, W>(sm, A.ld, tz, A.P, (const V *)A.tw, T(A.sign));
    const int ii = (tile % ntx) * tz + t0;
    V * base = (V *)A.S + (long long)A.nx * (tile / ntx);
    #pragma unroll
    for(int q = 0; q < EPTK; q++){
      const int kz = r0 + q * dr;
      if(kz < nz && ii < A.nx) tck_st(&base[ii + kstride * kz], sm[t0 * A.ld + fpad(kz)]);
    }
    fft_lds_barrier();
  }
  #undef ZF_ELEM
  #undef ZF_PUT
}

/* ------------------------------------------------------------------------------------- */
struct YInvArgs {
  tc_geom_t g;
  const void * S;
  void * psi;              /* of the transform type */
  int nky, ti, ld;
  FftPlan P;
  const void * twM;
  const void * twN;
};

/* Z'_k = (X_k + conj X_{M-k}) + i W^{-k} (X_k - conj X_{M-k}) so that
 * IFFT_M(Z')_m = x_2m + i x_2m+1 (unnormalised, as FFTW's c2r) */
template <typename V>
__device__ __forceinline__ V c2r_pre(V xk, V xmk, V w){
  const V xc = cconj(xmk);
  const V a = cadd(xk, xc);
  const V d = csub(xk, xc);
  const V wd = cmul(cconj(w), d);
  V r;
  r.x = a.x - wd.y;
  r.y = a.y + wd.x;
  return r;
}

template <typename T, bool W>
__global__ void __launch_bounds__(FFT_MAXTHREADS) tck_yifft_psi(const YInvArgs A){
  typedef typename cvec<T>::type V;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  V * sm = (V *)fft_smem;
  const tc_geom_t & g = A.g;
  const int ti = A.ti;
  const int M = A.P.M;
  const int ntx = (g.nx + ti - 1) / ti;
  const long long tile = fft_tile_of((long long)ntx * g.nz);
  const int i0 = 1 + (int)(tile % ntx) * ti;
  const int k = 1 + (int)(tile / ntx);
  const long long rowbase = (long long)g.nx * (long long)A.nky * (k - 1);
  const V * S = (const V *)A.S;
  const V * twN = (const V *)A.twN;
  /* load X_1..X_{M-1}; Z'_0 needs X_0 and X_M only and is formed here.  FFT_GZ elements are
   * loaded together (with X_M for a kk = 0 element); out-of-field elements read element 0 and
   * are put as 0 */
  const int nel = ti * M;
  for(int e0 = threadIdx.x; e0 < nel; e0 += FFT_GZ * A.P.nt){
    V v[FFT_GZ], xm[FFT_GZ];
    #pragma unroll
    for(int u = 0; u < FFT_GZ; u++){
      const int e = e0 + u * A.P.nt;
      const int t = (e & (ti - 1)), kk = (e >> fft_lg(ti)), i = i0 + t;
      const bool ok = e < nel && i <= g.nx;
      const long long o = (long long)(i - 1) + rowbase;
      v[u] = tck_ld(ok ? &S[o + (long long)g.nx * kk] : S);
      xm[u] = tck_ld((ok && kk == 0) ? &S[o + (long long)g.nx * M] : S);
    }
    #pragma unroll
    for(int u = 0; u < FFT_GZ; u++){
      const int e = e0 + u * A.P.nt;
      if(e >= nel) break;
      const int t = (e & (ti - 1)), kk = (e >> fft_lg(ti)), i = i0 + t;
      V w = mk2<T>(T(0.), T(0.));
      if(i <= g.nx){
        w = v[u];
        if(kk == 0){
          V x0 = v[u], xmm = xm[u];
          x0.y = T(0.);
          xmm.y = T(0.);
          w = c2r_pre(x0, xmm, twN[0]);
        }
      }
      sm[t * A.ld + fpad(kk)] = w;
    }
  }
  __syncthreads();
  /* pairs (k, M-k), k = 1..M/2, in place */
  const int half = M / 2;
  for(int e = threadIdx.x; e < ti * half; e += A.P.nt){
    const int t = (e & (ti - 1));
    const int k1 = 1 + (e >> fft_lg(ti));
    const int k2 = M - k1;
    const V xa = sm[t * A.ld + fpad(k1)];
    const V xb = sm[t * A.ld + fpad(k2)];
    const V z1 = c2r_pre(xa, xb, twN[k1]);
    if(k2 != k1){
      const V z2 = c2r_pre(xb, xa, twN[k2]);
      sm[t * A.ld + fpad(k2)] = z2;
    }
    sm[t * A.ld + fpad(k1)] = z1;
  }
  __syncthreads();
  lds_fft<T, FFT_EPT, W>(sm, A.ld, ti, A.P, (const V *)A.twM, T(1.));
  T * psi = (T *)A.psi;
  for(int e = threadIdx.x; e < ti * g.ny; e += A.P.nt){
    const int t = (e & (ti - 1));
    const int jj = (e >> fft_lg(ti));
    const int i = i0 + t;
    if(i > g.nx) continue;
    const T v = ((const T *)(sm + t * A.ld))[2 * fpad(jj >> 1) + (jj & 1)];
    tck_store_halo(psi, g, i, jj + 1, k, v, true);
  }
}

/* Persistent form of tck_yifft_psi (compiled for EPTK elements per thread per pass): each
 * block loops over tiles and issues the spectrum loads of the first PF of its elements of the
 * next tile before the current tile's passes. */
#ifndef TCK_YIFFT_PF
#define TCK_YIFFT_PF 2
#endif

template <typename T, int PF, int EPTK, bool W, bool X>
__global__ void __launch_bounds__(FFT_MAXTHREADS) tck_yifft_psi_p(const YInvArgs A){
  typedef typename cvec<T>::type V;
  extern __shared__ __align__(16) unsigned char fft_smem[];
  V * sm = (V *)fft_smem;
  const tc_geom_t & g = A.g;
  const int ti = A.ti;
  const int M = A.P.M;
  const int nt = A.P.nt;
  const int ntx = (g.nx + ti - 1) / ti;
  const int ntiles = ntx * g.nz;
  const V * S = (const V *)A.S;
  const V * twN = (const V *)A.twN;
  T * psi = (T *)A.psi;
  /* element q of this thread in a tile: line t0, wavenumber kk = r0 + q dr (the block size is
   * a multiple of the tile width, a power of two no larger than 64); only q = 0 can hold
   * kk = 0 */
  const int t0i = (threadIdx.x & (ti - 1)), r0i = (threadIdx.x >> fft_lg(ti)), dr = (nt >> fft_lg(ti));
  if constexpr (X){
    /* Exact form: the tiles cover the field exactly (nx a multiple of ti, blockDim.x EPTK =
     * ti M), so no load is conditional and the loops have fixed counts; the twiddle factors
     * are read from LDS (after the tile's lines); every spectrum value of the next tile is
     * loaded before the current tile's passes and stores, and the put waits only for those
     * loads (the zero-size stores before the loop give its entry the outstanding
     * loads-then-stores sequence of an iteration's end, as in tck_zfft_p). */
    V * twl = sm + ti * A.ld;
    V * twnl = twl + M;
    const V * twM = (const V *)A.twM;
    for(int t = threadIdx.x; t < M; t += nt){ twl[t] = twM[t]; twnl[t] = twN[t]; }
    constexpr int NST = 2 * EPTK;        /* psi values stored per thread: ti ny = 2 blockDim.x EPTK */
    int tile = blockIdx.x;               /* the launch has at least four tiles per block */
    /* buffer addressing: a tile's spectrum values lie in one plane of S (nky nx complex values)
     * and its psi values in one plane of the field, so each access is a per-thread byte offset
     * that does not change between tiles plus a uniform one */
    constexpr unsigned CB = sizeof(V), RB = sizeof(T);                 /* complex, real bytes */
    const unsigned vs = CB * (unsigned)(t0i + g.nx * r0i);             /* X_kk, kk = r0 + q dr */
    const unsigned ss = CB * (unsigned)(g.nx * dr);                    /* per q */
    const unsigned vxm = CB * (unsigned)(t0i + g.nx * M);              /* X_M */
    const unsigned sbytes = CB * (unsigned)(g.nx * A.nky);
    const unsigned vp = RB * (unsigned)(t0i + g.px * r0i);             /* psi row r0 + 1 + q dr */
    const unsigned sp = RB * (unsigned)(g.px * dr);
    const unsigned pbytes = RB * (unsigned)(g.px * g.ny);
    #define YX_RS(tl) tck_rsrc(S + (long long)((tl) % ntx * ti) + (long long)g.nx * A.nky * ((tl) / ntx), sbytes)
    V pre[EPTK], prexm;
    {
      const __amdgpu_buffer_rsrc_t rs = YX_RS(tile);
      #pragma unroll
      for(int q = 0; q < EPTK; q++) pre[q] = tck_bldc<V>(rs, vs, ss * q);
      prexm = tck_bldc<V>(rs, vxm, 0u);
      const __amdgpu_buffer_rsrc_t nul = tck_rsrc(A.S, 0u);
      #pragma unroll
      for(int q = 0; q < NST; q++) tck_bst0(nul, 16u * q);
    }
    for(;;){
      /* the LDS offsets are recomputed in every tile rather than held across the passes */
      int t0 = t0i, r0 = r0i;
      asm volatile("" : "+v"(t0), "+v"(r0));
      const int i0 = 1 + (tile % ntx) * ti;
      const int k = 1 + tile / ntx;
      /* X_1..X_{M-1} as loaded; Z'_0 from X_0 and X_M (element 0 of the threads with r0 = 0) */
      #pragma unroll
      for(int q = 0; q < EPTK; q++){
        V v = pre[q];
        if(q == 0 && r0 == 0){
          V xm = prexm;
          v.y = T(0.);
          xm.y = T(0.);
          v = c2r_pre(v, xm, twnl[0]);
        }
        sm[t0 * A.ld + fpad(r0 + q * dr)] = v;
      }
      fft_lds_barrier();
      {
        const int nxt = tile + (int)gridDim.x < ntiles ? tile + (int)gridDim.x : tile;
        const __amdgpu_buffer_rsrc_t rs = YX_RS(nxt);
        #pragma unroll
        for(int q = 0; q < EPTK; q++) pre[q] = tck_bldc<V>(rs, vs, ss * q);
        prexm = tck_bldc<V>(rs, vxm, 0u);
      }
      /* Z'_k1, Z'_k2 for k1 = 1..M/2 (ti M / 2 = blockDim.x EPTK / 2 pairs) */
      #pragma unroll
      for(int q = 0; q < EPTK / 2; q++){
        const int e = threadIdx.x + q * nt;
        const int t = (e & (ti - 1));
        const int k1 = 1 + (e >> fft_lg(ti));
        const int k2 = M - k1;
        const V xa = sm[t * A.ld + fpad(k1)];
        const V xb = sm[t * A.ld + fpad(k2)];
        const V z1 = c2r_pre(xa, xb, twnl[k1]);
        if(k2 != k1){
          const V z2 = c2r_pre(xb, xa, twnl[k2]);
          sm[t * A.ld + fpad(k2)] = z2;
        }
        sm[t * A.ld + fpad(k1)] = z1;
      }
      fft_lds_barrier();
      lds_fft<T, EPTK, W>(sm, A.ld, ti, A.P, twl, T(1.));
      /* element q of this thread: line t0, row jj = r0 + q dr; the interior values first, all
       * unconditional, then the halo images and wall copies of the threads that have any */
      const int i = i0 + t0;
      const T * ln = (const T *)(sm + t0 * A.ld);
      {
        const __amdgpu_buffer_rsrc_t rp = tck_rsrc(psi + tck_idx(g, i0, 1, k), pbytes);
        #pragma unroll
        for(int q = 0; q < NST; q++){
          const int jj = r0 + q * dr;
          tck_bstn<T>(rp, vp, sp * q, ln[2 * fpad(jj >> 1) + (jj & 1)]);
        }
      }
      if(i == 1 || i == g.nx || k == 1 || k == g.nz || r0 == 0 || r0 == dr - 1){
        for(int q = 0; q < NST; q++){
          const int jj = r0 + q * dr;
          tck_store_halo<T, false>(psi, g, i, jj + 1, k, ln[2 * fpad(jj >> 1) + (jj & 1)], true);
        }
      }
      fft_lds_barrier();
      tile += (int)gridDim.x;
      if(tile >= ntiles) break;
    }
    #undef YX_RS
    return;
  }
  int t0 = t0i, r0 = r0i;
  V pre[PF];
  bool pok[PF];
  int tile = blockIdx.x;
  /* raw spectrum value X_kk of element q of tile tl: ok when inside the field; otherwise the
   * load reads element 0 (in bounds) and the value is replaced by 0 when put, so that no load
   * is conditional */
  #define YI_ELEM(tl, q, out, ok) { \
    const int kk_ = r0 + (q) * dr; \
    const int i_ = 1 + ((tl) % ntx) * ti + t0; \
    ok = kk_ < M && i_ <= g.nx; \
    const long long o_ = (long long)(i_ - 1) + (long long)g.nx * ((long long)A.nky * ((tl) / ntx) + kk_); \
    out = tck_ld(ok ? &S[o_] : S); }
  /* X_M of the line of element 0 (used when that element is kk = 0) */
  #define YI_XM(tl, out) { \
    const int i_ = 1 + ((tl) % ntx) * ti + t0; \
    const bool ok_ = r0 == 0 && i_ <= g.nx; \
    const long long o_ = (long long)(i_ - 1) + (long long)g.nx * ((long long)A.nky * ((tl) / ntx) + M); \
    out = tck_ld(ok_ ? &S[o_] : S); }
  V prexm = mk2<T>(T(0.), T(0.));
  if(tile < ntiles){
    #pragma unroll
    for(int q = 0; q < PF; q++){ YI_ELEM(tile, q, pre[q], pok[q]); }
    YI_XM(tile, prexm);
  }
  for(; tile < ntiles; tile += gridDim.x){
    /* the element offsets are recomputed in every tile rather than held across the passes */
    t0 = t0i; r0 = r0i;
    asm volatile("" : "+v"(t0), "+v"(r0));
    const int i0 = 1 + (tile % ntx) * ti;
    const int k = 1 + tile / ntx;
    /* X_1..X_{M-1} as loaded; Z'_0 from X_0 and X_M */
    auto put = [&](const int q, V v, const bool ok, V xm){
      const int kk = r0 + q * dr;
      if(kk >= M) return;
      if(!ok) v = mk2<T>(T(0.), T(0.));
      else if(kk == 0){
        v.y = T(0.);
        xm.y = T(0.);
        v = c2r_pre(v, xm, twN[0]);
      }
      sm[t0 * A.ld + fpad(kk)] = v;
    };
    /* the rest of the tile: all loads issued before the first is waited for */
    V cur[EPTK - PF > 0 ? EPTK - PF : 1];
    bool cok[EPTK - PF > 0 ? EPTK - PF : 1];
    #pragma unroll
    for(int q = PF; q < EPTK; q++){ YI_ELEM(tile, q, cur[q - PF], cok[q - PF]); }
    #pragma unroll
    for(int q = 0; q < PF; q++) put(q, pre[q], pok[q], prexm);
    #pragma unroll
    for(int q = PF; q < EPTK; q++) put(q, cur[q - PF], cok[q - PF], prexm);
    fft_lds_barrier();
    if(tile + (int)gridDim.x < ntiles){
      #pragma unroll
      for(int q = 0; q < PF; q++){ YI_ELEM(tile + (int)gridDim.x, q, pre[q], pok[q]); }
      YI_XM(tile + (int)gridDim.x, prexm);
    }
    const int half = M / 2;
    for(int e = threadIdx.x; e < ti * half; e += nt){
      const int t = (e & (ti - 1));
      const int k1 = 1 + (e >> fft_lg(ti));
      const int k2 = M - k1;
      const V xa = sm[t * A.ld + fpad(k1)];
      const V xb = sm[t * A.ld + fpad(k2)];
      const V z1 = c2r_pre(xa, xb, twN[k1]);
      if(k2 != k1){
        const V z2 = c2r_pre(xb, xa, twN[k2]);
        sm[t * A.ld + fpad(k2)] = z2;
      }
      sm[t * A.ld + fpad(k1)] = z1;
    }
    fft_lds_barrier();
    lds_fft<T, EPTK, W>(sm, A.ld, ti, A.P, (const V *)A.twM, T(1.));
    for(int e = threadIdx.x; e < ti * g.ny; e += nt){
      const int t = (e & (ti - 1));
      const int jj = (e >> fft_lg(ti));
      const int i = i0 + t;
      if(i > g.nx) continue;
      const T v = ((const T *)(sm + t * A.ld))[2 * fpad(jj >> 1) + (jj & 1)];
      tck_store_halo(psi, g, i, jj + 1, k, v, true);
    }
    fft_lds_barrier();
  }
  #undef YI_ELEM
  #undef YI_XM
}

/* ------------------------------------------------------------------------------------- */
extern "C" {

static int make_plan(int M, FftPlan * P){
  P->M = M;
  P->nst = 0;
  int n = M;
  while(n > 1){
    int r;
    if(n % 8 == 0) r = 8;
    else if(n % 4 == 0) r = 4;
    else if(n % 2 == 0) r = 2;
    else if(n % 3 == 0) r = 3;
    else if(n % 5 == 0) r = 5;
    else if(n % 7 == 0) r = 7;
    else return -1;
    if(P->nst >= FFT_MAXST) return -1;
    P->rad[P->nst++] = r;
    n /= r;
  }
  /* division multipliers of each stage (dividends stay below 16 * 4096 = 2^16) */
  int ns = 1;
  for(int st = 0; st < P->nst; st++){
    const unsigned per = (unsigned)(M / P->rad[st]);
    P->mper[st] = per > 1 ? (unsigned)(0x100000000ULL / per + 1) : 0u;
    P->mns[st] = ns > 1 ? (unsigned)(0x100000000ULL / (unsigned)ns + 1) : 0u;
    ns *= P->rad[st];
  }
  return 0;
}

int tck_fft_supported(int n){
  FftPlan P;
  return make_plan(n, &P) == 0;
}

/* the plan holds a radix 5 or 7 pass: the launch takes the kernels compiled with them */
static int plan_wide(const FftPlan * P){
  for(int q = 0; q < P->nst; q++) if(P->rad[q] == 5 || P->rad[q] == 7) return 1;
  return 0;
}

/* Tile of TB transforms of length M per block: TB = 16 (one 128 B line of doubles, or two
 * lines of complex values, per row access) when the LDS allows it, fewer for very long
 * transforms; threads = TB * M / FFT_EPT, or TB * M / (FFT_EPT / 2) for tiles so large that
 * the LDS holds too few blocks per CU (measured on the FP64 transforms of 512x1024x1024:
 * 7.2 -> 5.6 ms for the divergence + theta transform, 4.6 -> 3.7 ms for the inverse theta
 * transform; the FP32 tiles, already at 16 waves per CU, were slower with twice the
 * threads). */
#ifndef TCK_FFT_HALFEPT
#define TCK_FFT_HALFEPT 1    /* LDS-bound tiles: FFT_EPT / 2 elements per thread */
#endif
#ifndef TCK_FFT_YSHARE
#define TCK_FFT_YSHARE 1     /* theta transforms: blocks per CU the tiles are sized for */
#endif
#ifndef TCK_FFT_ZSHARE
#define TCK_FFT_ZSHARE 1     /* axial transforms */
#endif
static size_t g_lds_limit = 0;

static size_t lds_limit(void){
  if(g_lds_limit == 0){
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess) dev = 0;
    int v = 0;
    if(hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || v <= 0){
      if(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
    }
    g_lds_limit = (size_t)v;
  }
  return g_lds_limit;
}

static int g_fft_ncu = 0;

static int fft_ncu(void){
  if(g_fft_ncu == 0){
    int dev = 0, v = 0;
    if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    g_fft_ncu = v;
  }
  return g_fft_ncu;
}

/* share: tiles sized to leave room for `share` blocks per CU in the LDS (while at least 4
 * transforms, 64 B of contiguous FP64 complex values, remain per row access).  lines: the
 * transforms of the launch; small launches take narrower tiles, so that there are at least
 * two blocks per CU */
/* LDS pitch of one padded line of M elements, odd so that the rows of a tile start in
 * different banks */
static int fft_ld(int M){
  const int ld = fpad(M - 1) + 1;
  return ld | 1;
}

/* the tile width b is a power of two (16 halved or doubled, capped at a power of two): the
 * kernels divide by it with shifts */
static int fft_tile(int M, int cap, size_t csize, int * threads, int share, long long lines){
  int b = 16;
  const size_t lb = (size_t)fft_ld(M) * csize;
  while(b > 1 && ((size_t)b * lb > lds_limit() || b * M > FFT_EPT * FFT_MAXTHREADS)) b /= 2;
  while(share > 1 && b > 4 && (size_t)b * lb * share > lds_limit()) b /= 2;
  while(b * 2 <= cap && b * 2 * M <= FFT_EPT * 64) b *= 2;   /* short transforms: wider tiles */
  while(b > 1 && lines / b < 2LL * fft_ncu()) b /= 2;
  if(b > cap) b = cap < 1 ? 1 : cap;
  /* elements per thread: FFT_EPT, or half that where the LDS holds so few blocks per CU that
   * FFT_EPT would leave fewer than 16 waves on it (FP64 tiles of long transforms: one block
   * per CU).  The device passes loop over their butterflies, so any count up to FFT_EPT
   * computes the same bits. */
  int ept = FFT_EPT;
  const long long per = (long long)(lds_limit() / ((size_t)b * lb));
  const int w16 = ((b * M + FFT_EPT - 1) / FFT_EPT + 63) / 64;
  if(TCK_FFT_HALFEPT && per * w16 < 16 && (b * M + FFT_EPT / 2 - 1) / (FFT_EPT / 2) <= FFT_MAXTHREADS) ept = FFT_EPT / 2;
  int t = (b * M + ept - 1) / ept;
  t = (t + 63) / 64 * 64;
  if(t < 64) t = 64;
  *threads = t;
  return b;
}

/* raises a kernel's dynamic LDS limit once per (kernel, size): later launches, including
 * those recorded into a graph, make no attribute call */
#define FFT_LDS_CACHE 32
static const void * g_lds_fn[FFT_LDS_CACHE];
static size_t g_lds_bytes[FFT_LDS_CACHE];
static int g_lds_n = 0;

static int set_lds(const void * fn, size_t bytes){
  if(bytes <= 65536) return 0;
  for(int q = 0; q < g_lds_n; q++) if(g_lds_fn[q] == fn && g_lds_bytes[q] >= bytes) return 0;
  if(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) return -1;
  if(g_lds_n < FFT_LDS_CACHE){
    g_lds_fn[g_lds_n] = fn;
    g_lds_bytes[g_lds_n] = bytes;
    g_lds_n++;
  }
  return 0;
}

/* Persistent launches, FP64 only (the FP32 forms, at 16 elements per thread, spill with the
 * prefetch registers).  The persistent kernels are compiled for FFT_EPT / 2 elements per
 * thread per pass, the count the LDS-bound FP64 tiles take; a launch qualifies when its
 * threads cover the tile at that count and it has at least four tiles per resident block.
 * Returns 1 when launched (the caller checks the launch), 0 when the caller should launch the
 * one-tile form, < 0 when the LDS limit could not be raised.  Resident blocks per CU are
 * queried once per (kernel, threads, LDS), so launches recorded into a graph make no query. */
#define PERSIST_EPT (FFT_EPT / 2)
#define PERSIST_CACHE 16
static const void * g_pc_fn[PERSIST_CACHE];
static int g_pc_nt[PERSIST_CACHE], g_pc_per[PERSIST_CACHE];
static size_t g_pc_lds[PERSIST_CACHE];
static int g_pc_n = 0;

}   /* extern "C" */

/* sets the kernel's LDS limit, then launches (lambda) */
template <typename F>
static int fft_one(const void * fn, F launch, size_t lds){
  if(set_lds(fn, lds)) return 1;
  launch();
  return 0;
}

template <typename Args>
static int launch_persistent(void (*kern)(const Args), const Args & A, int nt, size_t lds,
    long long elems, long long ntiles, hipStream_t stream, int ept = PERSIST_EPT){
  if((long long)nt * ept < elems || nt > FFT_MAXTHREADS) return 0;
  const void * fn = (const void *)kern;
  int per = 0;
  for(int q = 0; q < g_pc_n; q++) if(g_pc_fn[q] == fn && g_pc_nt[q] == nt && g_pc_lds[q] == lds) per = g_pc_per[q];
  if(per == 0){
    if(set_lds(fn, lds)) return -1;
    if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, nt, lds) != hipSuccess || per < 1) per = 1;
    if(g_pc_n < PERSIST_CACHE){
      g_pc_fn[g_pc_n] = fn; g_pc_nt[g_pc_n] = nt; g_pc_lds[g_pc_n] = lds; g_pc_per[g_pc_n] = per;
      g_pc_n++;
    }
  }
  const long long nb = (long long)per * fft_ncu();
  if(ntiles < 4 * nb) return 0;
  if(set_lds(fn, lds)) return -1;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(nt), lds, stream, A);
  return 1;
}

extern "C" {

/* the one-tile form of kernel K for the transform type (float when mixed) and radix set;
 * nonzero when the LDS limit could not be raised */
#define FFT_ONE(K, mixed, wide) ( \
  (mixed) ? ((wide) ? fft_one((const void *)K<float, true>, [&]{ hipLaunchKernelGGL((K<float, true>), grd, dim3(nt), lds, stream, A); }, lds) \
                    : fft_one((const void *)K<float, false>, [&]{ hipLaunchKernelGGL((K<float, false>), grd, dim3(nt), lds, stream, A); }, lds)) \
          : ((wide) ? fft_one((const void *)K<double, true>, [&]{ hipLaunchKernelGGL((K<double, true>), grd, dim3(nt), lds, stream, A); }, lds) \
                    : fft_one((const void *)K<double, false>, [&]{ hipLaunchKernelGGL((K<double, false>), grd, dim3(nt), lds, stream, A); }, lds)))

/* mixed: spectrum, transforms and psi in FP32 (twiddles twM, twN, tw as float2) */
int tck_launch_div_yfft(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * uz, const double * dt, int rkstep, void * S,
    const void * twM, const void * twN, int mixed, int twl, hipStream_t stream){
  YFwdArgs A;
  A.g = *g; A.c = *c; A.ux = ux; A.uy = uy; A.uz = uz; A.dt = dt; A.rkstep = rkstep;
  A.S = S; A.nky = g->ny / 2 + 1;
  if(make_plan(g->ny / 2, &A.P) != 0) return (int)hipErrorInvalidValue;
  const size_t cs = mixed ? sizeof(float2) : sizeof(double2);
  int nt;
  A.ti = fft_tile(A.P.M, 64, cs, &nt, TCK_FFT_YSHARE, (long long)g->nx * g->nz);
  A.P.nt = nt;
  A.twM = twM; A.twN = twN;
  A.ld = fft_ld(A.P.M);
  const size_t lds = cs * (size_t)A.ti * A.ld;
  const long long ntiles = (long long)((g->nx + A.ti - 1) / A.ti) * g->nz;
#if TCK_DIVY_PF > 0
  {
    const long long el = (long long)A.ti * A.P.M;
    /* the twiddle tables in LDS where they fit beside the tile (radix 2, 3, 4, 8 plans) */
    const size_t ldsx = lds + cs * (size_t)(2 * A.P.M + 1);
    const bool tl = twl && !plan_wide(&A.P) && ldsx <= lds_limit();
    /* FP32 tiles: persistent only in this form, with FFT_EPT / 2 elements per thread (twice
     * the threads of the one-tile form), which keeps the carried faces and the prefetch in
     * registers */
    const int ntf = (int)(el / PERSIST_EPT);
    YFwdArgs AF = A;
    AF.P.nt = ntf;
    const bool tlf = tl && ntf <= FFT_MAXTHREADS && ntf % 64 == 0 && ntf % A.ti == 0 && (long long)ntf * PERSIST_EPT == el;
    const int r = mixed ? (tlf ? launch_persistent(tck_div_yfft_p<float, TCK_DIVY_PF, PERSIST_EPT, false, true>, AF, ntf, ldsx, el, ntiles, stream) : 0)
                : tl ? launch_persistent(tck_div_yfft_p<double, TCK_DIVY_PF, PERSIST_EPT, false, true>, A, nt, ldsx, el, ntiles, stream)
                    : (plan_wide(&A.P) ? launch_persistent(tck_div_yfft_p<double, TCK_DIVY_PF, PERSIST_EPT, true, false>, A, nt, lds, el, ntiles, stream)
                                       : launch_persistent(tck_div_yfft_p<double, TCK_DIVY_PF, PERSIST_EPT, false, false>, A, nt, lds, el, ntiles, stream));
    if(r < 0) return (int)hipErrorInvalidValue;
    if(r > 0){ TCK_CHECK_LAUNCH(); return 0; }
  }
#endif
  const dim3 grd((unsigned)ntiles);
  if(FFT_ONE(tck_div_yfft, mixed, plan_wide(&A.P))) return (int)hipErrorInvalidValue;
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_zfft(int nx, int nky, int nz, void * S, const void * tw, int sign, int mixed, int exact, hipStream_t stream){
  ZArgs A;
  A.nx = nx; A.nky = nky; A.nz = nz; A.S = S; A.tw = tw;
  A.sign = sign < 0 ? -1. : 1.;
  if(make_plan(nz, &A.P) != 0) return (int)hipErrorInvalidValue;
  const size_t cs = mixed ? sizeof(float2) : sizeof(double2);
  int nt;
  A.tz = fft_tile(nz, 64, cs, &nt, TCK_FFT_ZSHARE, (long long)nx * nky);
  A.P.nt = nt;
  A.ld = fft_ld(nz);
  const size_t lds = cs * (size_t)A.tz * A.ld;
  const long long ntiles = (long long)((nx + A.tz - 1) / A.tz) * nky;
#if TCK_ZFFT_PF > 0
  {
    const long long el = (long long)A.tz * nz;
    /* the exact form when the tiles cover the spectrum exactly, the plan has no radix 5 or 7
     * passes and its twiddle factors fit in the LDS beside the tile */
    const size_t ldsx = lds + cs * (size_t)nz;
    const bool x = exact && !plan_wide(&A.P) && nx % A.tz == 0 && (long long)nt * PERSIST_EPT == el && ldsx <= lds_limit();
    /* FP32 tiles (twice as wide, FFT_EPT elements per thread): only the exact form is
     * persistent, its registers holding the next tile in place of the prefetch */
    const bool xf = exact && !plan_wide(&A.P) && nx % A.tz == 0 && (long long)nt * FFT_EPT == el && ldsx <= lds_limit()
        && (long long)sizeof(float2) * nx * nky * nz < (1LL << 32);
    const int r = mixed ? (xf ? launch_persistent(tck_zfft_p<float, TCK_ZFFT_PF, FFT_EPT, false, true>, A, nt, ldsx, el, ntiles, stream, FFT_EPT) : 0)
                : x ? launch_persistent(tck_zfft_p<double, TCK_ZFFT_PF, PERSIST_EPT, false, true>, A, nt, ldsx, el, ntiles, stream)
                    : (plan_wide(&A.P) ? launch_persistent(tck_zfft_p<double, TCK_ZFFT_PF, PERSIST_EPT, true, false>, A, nt, lds, el, ntiles, stream)
                                       : launch_persistent(tck_zfft_p<double, TCK_ZFFT_PF, PERSIST_EPT, false, false>, A, nt, lds, el, ntiles, stream));
    if(r < 0) return (int)hipErrorInvalidValue;
    if(r > 0){ TCK_CHECK_LAUNCH(); return 0; }
  }
#endif
  const dim3 grd((unsigned)ntiles);
  if(FFT_ONE(tck_zfft, mixed, plan_wide(&A.P))) return (int)hipErrorInvalidValue;
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_yifft_psi(const tc_geom_t * g, const void * S, void * psi, const void * twM,
    const void * twN, int mixed, int exact, hipStream_t stream){
  YInvArgs A;
  A.g = *g; A.S = S; A.psi = psi; A.nky = g->ny / 2 + 1;
  if(make_plan(g->ny / 2, &A.P) != 0) return (int)hipErrorInvalidValue;
  const size_t cs = mixed ? sizeof(float2) : sizeof(double2);
  int nt;
  A.ti = fft_tile(A.P.M, 64, cs, &nt, TCK_FFT_YSHARE, (long long)g->nx * g->nz);
  A.P.nt = nt;
  A.twM = twM; A.twN = twN;
  A.ld = fft_ld(A.P.M);
  const size_t lds = cs * (size_t)A.ti * A.ld;
  const long long ntiles = (long long)((g->nx + A.ti - 1) / A.ti) * g->nz;
#if TCK_YIFFT_PF > 0
  {
    const long long el = (long long)A.ti * A.P.M;
    /* the exact form when the tiles cover the field exactly, the plan has no radix 5 or 7
     * passes, both twiddle tables fit in the LDS beside the tile and a plane's byte offsets fit
     * in 31 bits */
    const size_t ldsx = lds + cs * (size_t)(2 * A.P.M);
    const bool x = exact && !plan_wide(&A.P) && g->nx % A.ti == 0 && (long long)nt * PERSIST_EPT == el && ldsx <= lds_limit()
        && 16LL * g->plane < (1LL << 31) && 16LL * g->nx * A.nky < (1LL << 31);
    /* FP32 tiles: only the exact form is persistent, with FFT_EPT / 2 elements per thread
     * (twice the threads of the one-tile form), which keeps its registers below 128 */
    const int ntf = (int)(el / PERSIST_EPT);
    YInvArgs AF = A;
    AF.P.nt = ntf;
    const bool xf = exact && !plan_wide(&A.P) && g->nx % A.ti == 0 && ntf <= FFT_MAXTHREADS && ntf % 64 == 0
        && ntf % A.ti == 0 && (long long)ntf * PERSIST_EPT == el && ldsx <= lds_limit()
        && 16LL * g->plane < (1LL << 31) && 16LL * g->nx * A.nky < (1LL << 31);
    const int r = mixed ? (xf ? launch_persistent(tck_yifft_psi_p<float, TCK_YIFFT_PF, PERSIST_EPT, false, true>, AF, ntf, ldsx, el, ntiles, stream) : 0)
                : x ? launch_persistent(tck_yifft_psi_p<double, TCK_YIFFT_PF, PERSIST_EPT, false, true>, A, nt, ldsx, el, ntiles, stream)
                    : (plan_wide(&A.P) ? launch_persistent(tck_yifft_psi_p<double, TCK_YIFFT_PF, PERSIST_EPT, true, false>, A, nt, lds, el, ntiles, stream)
                                       : launch_persistent(tck_yifft_psi_p<double, TCK_YIFFT_PF, PERSIST_EPT, false, false>, A, nt, lds, el, ntiles, stream));
    if(r < 0) return (int)hipErrorInvalidValue;
    if(r > 0){ TCK_CHECK_LAUNCH(); return 0; }
  }
#endif
  const dim3 grd((unsigned)ntiles);
  if(FFT_ONE(tck_yifft_psi, mixed, plan_wide(&A.P))) return (int)hipErrorInvalidValue;
  TCK_CHECK_LAUNCH();
  return 0;
}

}
