/*
 * tck_linesolve.hip -- batched tridiagonal line solves (gfx950).
 *
 * Replaces src/tdm.c (gtsv :17-50, periodic Sherman-Morrison :66-106) as driven by
 * src/fluid/predict/linear_system.c:6-103 (implicit diffusion, real) and
 * src/fluid/compute_potential.c:445-488 (radial Poisson systems, complex).
 *
 *  tck_factor_*     per-substep LU factors of the implicit-diffusion systems.  The
 *                   coefficients of these systems depend on the line direction and on i
 *                   only (never on the other two indices), so one factor table serves every
 *                   line; it is computed on the device from dt, exactly as gtsv would.
 *  tck_xsolve       x lines (contiguous in memory): one wavefront per line, each lane owns a
 *                   contiguous segment; the forward and backward substitutions are affine
 *                   recurrences whose segment compositions are combined by a 6-step
 *                   wave-wide scan, after which each lane replays its segment with the
 *                   reference's own update formulas.
 *  tck_ysolve/zsolve  periodic y and z lines: one lane per line, lanes run over i so every
 *                   step of the recurrence is a coalesced row access.
 *  tck_pfactor/pxsolve  Poisson radial systems per (ky, kz): factors precomputed once (the
 *                   operator does not change in time), solved like tck_xsolve on complex
 *                   right-hand sides.
 */
#include <float.h>
#include "tck_common.h"

#define XS_MAXSEG 16   /* line length <= 64 * 16 = 1024 */

/* ---------------------------------------------------------------------------------------
 * factor tables.  For a system (l, c, u) of size n, gtsv (tdm.c:17-50) computes
 *   v0 = u0 / c0,  r_i = 1 / (c_i - l_i v_{i-1}),  v_i = r_i u_i  (0 < i < n-1)
 *   last pivot c_{n-1} - l_{n-1} v_{n-2}: r_{n-1} = 1 / pivot, or 0 when |pivot| <= eps.
 * Tables: r[0] holds c0 itself (row 0 divides), l[i] = l_i, v[i] = v_i.
 * ------------------------------------------------------------------------------------- */
__device__ void factor_sys(int n, const double * __restrict__ lc, const double * __restrict__ cc,
    const double * __restrict__ uc, double pf, long long stride, double * __restrict__ r,
    double * __restrict__ v, double * __restrict__ l){
  /* tdm coefficients as solve_in_x builds them (predict/linear_system.c:28-32) */
  #define TL(i) (- pf * lc[i])
  #define TC(i) (1. - pf * cc[i])
  #define TU(i) (- pf * uc[i])
  double vprev = TU(0) / TC(0);
  r[0] = TC(0);
  v[0] = vprev;
  l[0] = TL(0);
  /* eight rows' coefficients are loaded together before their recurrence steps */
  int i = 1;
  for(; i + 8 <= n - 1; i += 8){
    double tla[8], tca[8], tua[8];
    #pragma unroll
    for(int e = 0; e < 8; e++){ tla[e] = TL(i + e); tca[e] = TC(i + e); tua[e] = TU(i + e); }
    #pragma unroll
    for(int e = 0; e < 8; e++){
      const double val = 1. / (tca[e] - tla[e] * vprev);
      vprev = val * (tua[e]);
      r[stride * (i + e)] = val;
      v[stride * (i + e)] = vprev;
      l[stride * (i + e)] = tla[e];
    }
  }
  for(; i < n - 1; i++){
    const double val = 1. / (TC(i) - TL(i) * vprev);
    vprev = val * (TU(i));
    r[stride * i] = val;
    v[stride * i] = vprev;
    l[stride * i] = TL(i);
  }
  const double val = TC(n - 1) - TL(n - 1) * vprev;
  r[stride * (n - 1)] = fabs(val) > DBL_EPSILON ? 1. / val : 0.;
  v[stride * (n - 1)] = 0.;
  l[stride * (n - 1)] = TL(n - 1);
  #undef TL
  #undef TC
  #undef TU
}

/* constant-coefficient periodic system of size n (tdm.c:78-99): factors of the leading
 * (n-1) block, the correction vector q1 and the meta row {den, l, u, jc, rc, vc, rlast}.
 * The factor recurrence r_i = 1 / (c - l v_{i-1}), v_i = r_i u of a constant-coefficient
 * system is a deterministic map of v_{i-1}: once two consecutive rows are equal bit for bit,
 * every later row (up to the last, which has its own pivot rule) repeats them.  jc is the
 * first such row; solvers read rows below jc from the tables and use rc, vc above. */
#define TC_PMETA 7

__device__ void factor_periodic(int n, double tl, double tc, double tu, long long stride,
    double * __restrict__ r, double * __restrict__ v, double * __restrict__ q1,
    double * __restrict__ meta, long long mstride){
  const int m = n - 1;
  /* factors of the (n-1) leading block */
  double vprev = tu / tc;
  r[0] = tc;
  v[0] = vprev;
  int jc = m - 1;
  double rprev = 0.;
  for(int i = 1; i < m - 1; i++){
    const double val = 1. / (tc - tl * vprev);
    const double vi = val * (tu);
    if(jc == m - 1 && i >= 2 && val == rprev && vi == vprev) jc = i;
    r[stride * i] = val;
    v[stride * i] = vi;
    rprev = val;
    vprev = vi;
  }
  double rlast;
  {
    const double val = tc - tl * vprev;
    rlast = fabs(val) > DBL_EPSILON ? 1. / val : 0.;
    r[stride * (m - 1)] = rlast;
    v[stride * (m - 1)] = 0.;
  }
  /* q1 = A_{n-1}^{-1} (-l0, 0, ..., 0, -u_{n-2}) by the same gtsv arithmetic */
  /* the loads of r and v below read rows written by the loop above: eight rows' loads are
   * issued together; the backward pass carries q1_{i+1} in a register */
  double qprev = - tl / tc;   /* row 0 */
  q1[0] = qprev;
  int i = 1;
  for(; i + 8 <= m; i += 8){
    double ra[8];
    #pragma unroll
    for(int e = 0; e < 8; e++) ra[e] = r[stride * (i + e)];
    #pragma unroll
    for(int e = 0; e < 8; e++){
      const double rhs = i + e == m - 1 ? - tu : 0.;
      const double q = ra[e] * (rhs - tl * qprev);
      q1[stride * (i + e)] = q;
      qprev = q;
    }
  }
  for(; i < m; i++){
    const double rhs = i == m - 1 ? - tu : 0.;
    const double q = r[stride * i] * (rhs - tl * qprev);
    q1[stride * i] = q;
    qprev = q;
  }
  double qnext = q1[stride * (m - 1)];
  i = m - 2;
  /* eight rows' loads before any of their stores, so that the rows do not wait one by one
   * for a store to reach memory before the next load */
  for(; i >= 7; i -= 8){
    double qa[8], va[8];
    #pragma unroll
    for(int e = 0; e < 8; e++){ qa[e] = q1[stride * (i - e)]; va[e] = v[stride * (i - e)]; }
    #pragma unroll
    for(int e = 0; e < 8; e++){
      const double q = qa[e] - va[e] * qnext;
      q1[stride * (i - e)] = q;
      qnext = q;
    }
  }
  for(; i >= 0; i--){
    const double q = q1[stride * i] - v[stride * i] * qnext;
    q1[stride * i] = q;
    qnext = q;
  }
  meta[0] = tc + tu * q1[0] + tl * q1[stride * (m - 1)];
  meta[mstride] = tl;
  meta[2 * mstride] = tu;
  meta[3 * mstride] = (double)jc;
  meta[4 * mstride] = jc < m - 1 ? r[stride * jc] : 0.;
  meta[5 * mstride] = jc < m - 1 ? v[stride * jc] : 0.;
  meta[6 * mstride] = rlast;
}

/* implicit-diffusion factor tables, one job per (field, substep).
 *   x: one system of size n over i = i0..i0+n-1 with per-row stencils lap{l,c,u}[i]
 *   y: one periodic system per i (stencil lap{l,c,u}[i]) of size ny, table layout [j][tp]
 *   z: one periodic system of size nz (stencil lapz)
 * prefactor = fac * (0.5 * gamma * dt * diffusivity) (ux.c:496-506 and its siblings; fac is
 * 2 for the diagonal stress direction of each velocity and for T in y, t.c:346-349).
 * ymeta/zmeta hold the TC_PMETA meta rows of the periodic systems ([7][tp] and [7]).  All jobs of a
 * step run in one launch: each system is a short sequential recurrence, so the launch costs
 * one recurrence latency however many systems it holds. */
#define FACTOR_MAXJOB 12

struct FactorJob {
  int rkstep;
  double diff;
  double fac[3];
  int i0, n;
  const double * lxl; const double * lxc; const double * lxu;
  const double * lyl; const double * lyc; const double * lyu;
  double * xr; double * xv; double * xl;
  double * yr; double * yv; double * yq1; double * ymeta;
  double * zr; double * zv; double * zq1; double * zmeta;
};

struct FactorBatch {
  const double * dt;
  int dirs[3];
  int nx, ny, nz, tp;
  double lzl, lzc, lzu;
  int njobs;
  FactorJob jobs[FACTOR_MAXJOB];
};

__constant__ double c_ls_rka[3] = {32. / 60., 25. / 60., 45. / 60.};
__constant__ double c_ls_rkb[3] = { 0. / 60., -17. / 60., -25. / 60.};

__global__ void tck_factor(const FactorBatch B){
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(blockIdx.y >= (unsigned)B.njobs) return;
  const FactorJob & A = B.jobs[blockIdx.y];
  const double dt = *B.dt;
  const double gamma = c_ls_rka[A.rkstep] + c_ls_rkb[A.rkstep];
  const double prefactor = 0.5 * gamma * dt * A.diff;
  /* thread 0: x system; threads 1..nx: y systems for i = t; thread nx+1: z system */
  if(t == 0 && B.dirs[0]){
    const double pf = A.fac[0] * prefactor;
    factor_sys(A.n, A.lxl + A.i0, A.lxc + A.i0, A.lxu + A.i0, pf, 1, A.xr, A.xv, A.xl);
  }
  if(t >= 1 && t <= B.nx && B.dirs[1]){
    const int i = t;
    const double pf = A.fac[1] * prefactor;
    const double tl = - pf * A.lyl[i];
    const double tc = 1. - pf * A.lyc[i];
    const double tu = - pf * A.lyu[i];
    factor_periodic(B.ny, tl, tc, tu, B.tp, A.yr + i, A.yv + i, A.yq1 + i, A.ymeta + i, B.tp);
  }
  if(t == B.nx + 1 && B.dirs[2]){
    const double pf = A.fac[2] * prefactor;
    const double tl = - pf * B.lzl;
    const double tc = 1. - pf * B.lzc;
    const double tu = - pf * B.lzu;
    factor_periodic(B.nz, tl, tc, tu, 1, A.zr, A.zv, A.zq1, A.zmeta, 1);
  }
}

/* ---------------------------------------------------------------------------------------
 * wave-level scans of affine maps x -> a x + b
 * ------------------------------------------------------------------------------------- */
/* value entering this lane's segment when maps compose in increasing lane order, x0 = 0 */
template <int NC>
__device__ __forceinline__ void scan_fwd(double a, double * b, double * in){
  const int lane = threadIdx.x % TCK_WAVE;
  #pragma unroll
  for(int off = 1; off < TCK_WAVE; off <<= 1){
    const double ap = __shfl_up(a, off, TCK_WAVE);
    double bp[NC];
    #pragma unroll
    for(int q = 0; q < NC; q++) bp[q] = __shfl_up(b[q], off, TCK_WAVE);
    if(lane >= off){
      #pragma unroll
      for(int q = 0; q < NC; q++) b[q] = a * bp[q] + b[q];
      a = a * ap;
    }
  }
  #pragma unroll
  for(int q = 0; q < NC; q++){
    const double prev = __shfl_up(b[q], 1, TCK_WAVE);
    in[q] = lane == 0 ? 0. : prev;
  }
}

/* same in decreasing lane order */
template <int NC>
__device__ __forceinline__ void scan_bwd(double a, double * b, double * in){
  const int lane = threadIdx.x % TCK_WAVE;
  #pragma unroll
  for(int off = 1; off < TCK_WAVE; off <<= 1){
    const double ap = __shfl_down(a, off, TCK_WAVE);
    double bp[NC];
    #pragma unroll
    for(int q = 0; q < NC; q++) bp[q] = __shfl_down(b[q], off, TCK_WAVE);
    if(lane + off < TCK_WAVE){
      #pragma unroll
      for(int q = 0; q < NC; q++) b[q] = a * bp[q] + b[q];
      a = a * ap;
    }
  }
  #pragma unroll
  for(int q = 0; q < NC; q++){
    const double next = __shfl_down(b[q], 1, TCK_WAVE);
    in[q] = lane == TCK_WAVE - 1 ? 0. : next;
  }
}

/* Solve one tridiagonal line held by a wavefront: lane L owns rows L*SEG .. L*SEG+SEG-1
 * (rows >= n are padding and pass values through).  q[e][c]: the lane's right-hand sides
 * (NC = 1 real, 2 complex).  rr, ll, vv: the factor rows of the lane's rows (factor_sys
 * layout: rr of row 0 holds c0 itself), already in registers, so the recurrences below wait
 * on no memory. */
template <int SEG, int NC>
__device__ __forceinline__ void wave_tdm(double (*q)[NC], const double * rr, const double * ll,
    const double * vv, int n){
  const int row0 = (threadIdx.x % TCK_WAVE) * SEG;
  /* forward: q_0 = q_0 / c_0, q_i = r_i (q_i - l_i q_{i-1}) */
  double a = 1., b[NC];
  #pragma unroll
  for(int c = 0; c < NC; c++) b[c] = 0.;
  #pragma unroll
  for(int e = 0; e < SEG; e++){
    const int row = row0 + e;
    if(row < n){
      if(row == 0){
        a = 0.;
        #pragma unroll
        for(int c = 0; c < NC; c++) b[c] = q[e][c] / rr[e];
      }else{
        const double ai = - rr[e] * ll[e];
        #pragma unroll
        for(int c = 0; c < NC; c++) b[c] = ai * b[c] + rr[e] * q[e][c];
        a = ai * a;
      }
    }
  }
  double in[NC];
  scan_fwd<NC>(a, b, in);
  #pragma unroll
  for(int e = 0; e < SEG; e++){
    const int row = row0 + e;
    if(row < n){
      if(row == 0){
        #pragma unroll
        for(int c = 0; c < NC; c++) q[e][c] = q[e][c] / rr[e];
      }else{
        #pragma unroll
        for(int c = 0; c < NC; c++) q[e][c] = rr[e] * (q[e][c] - ll[e] * in[c]);
      }
      #pragma unroll
      for(int c = 0; c < NC; c++) in[c] = q[e][c];
    }
  }
  /* backward: q_i -= v_i q_{i+1}, q_{n-1} kept */
  a = 1.;
  #pragma unroll
  for(int c = 0; c < NC; c++) b[c] = 0.;
  #pragma unroll
  for(int e = SEG - 1; e >= 0; e--){
    const int row = row0 + e;
    if(row < n){
      if(row == n - 1){
        a = 0.;
        #pragma unroll
        for(int c = 0; c < NC; c++) b[c] = q[e][c];
      }else{
        const double ai = - vv[e];
        #pragma unroll
        for(int c = 0; c < NC; c++) b[c] = ai * b[c] + q[e][c];
        a = ai * a;
      }
    }
  }
  scan_bwd<NC>(a, b, in);
  #pragma unroll
  for(int e = SEG - 1; e >= 0; e--){
    const int row = row0 + e;
    if(row < n){
      if(row != n - 1){
        #pragma unroll
        for(int c = 0; c < NC; c++) q[e][c] -= vv[e] * in[c];
      }
      #pragma unroll
      for(int c = 0; c < NC; c++) in[c] = q[e][c];
    }
  }
}

/* position of line element idx in the LDS staging area of a wave: lane idx / SEG owns it,
 * rows padded to SEG + 1 so that the lanes' segment reads fall in different banks */
template <int SEG>
__device__ __forceinline__ int stage_pos(int idx){
  return (idx / SEG) * (SEG + 1) + (idx % SEG);
}

static int g_num_cus = 0;

static int num_cus(void){
  if(g_num_cus == 0){
    int dev = 0, v = 0;
    if(hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    g_num_cus = v;
  }
  return g_num_cus;
}

/* x lines of the implicit diffusion solve (solve_in_x, predict/linear_system.c:6-37).
 * d: increment in the field layout, rows i = i0 .. i0+n-1 of each (j, k) line.
 * final: store ucur + solution with halos (the field update of ux.c:544-556). */
/* the x solves of up to TCK_XS_MAXF fields in one launch, one field per blockIdx.y: the
 * fields' systems are independent, and on small grids a launch per field costs more than
 * its solve */
#define TCK_XS_MAXF 4
struct XSolveArgs {
  tc_geom_t g;
  void * d[TCK_XS_MAXF];              /* the increment: double, or float in the mixed step */
  double * uout[TCK_XS_MAXF];         /* final pass: the new field (the increment's buffer itself in FP64) */
  const double * ucur[TCK_XS_MAXF];
  const double * r[TCK_XS_MAXF]; const double * v[TCK_XS_MAXF]; const double * l[TCK_XS_MAXF];
  int i0[TCK_XS_MAXF], n[TCK_XS_MAXF];
};

/* LDS ordering between the lanes of one wavefront (the staging areas below are private to a
 * wave, so the waves of a block need not wait for each other) */
__device__ __forceinline__ void wave_sync(){
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Four waves per block, one x line per wave per iteration; blocks loop over groups of four
 * lines.  The factor rows (the same for every line) are staged once per block in LDS, in the
 * lanes' segment order, and read where the recurrences use them: held in registers for the
 * whole loop they would cost 3 SEG doubles per lane and occupancy.  Lines are staged through
 * a wave-private LDS area: global loads and stores are contiguous across the wave (lane t
 * touches elements t, t+64, ...), each lane solves its contiguous segment. */
template <int SEG, typename D, bool U>
__device__ __forceinline__ void xsolve_body(const XSolveArgs & A, D * __restrict__ dd,
    double * uout, const double * __restrict__ ucur, const double * __restrict__ rt,
    const double * __restrict__ vt, const double * __restrict__ lt, const int i0, const int n,
    double * sm, double (*tab)[TCK_WAVE * (SEG + 1)]){
  const tc_geom_t & g = A.g;
  const int w = threadIdx.x / TCK_WAVE;
  const int lane = threadIdx.x % TCK_WAVE;
  for(int idx = threadIdx.x; idx < TCK_WAVE * SEG; idx += blockDim.x){
    const int row = idx < n ? idx : n - 1;
    tab[0][stage_pos<SEG>(idx)] = rt[row];
    tab[1][stage_pos<SEG>(idx)] = lt[row];
    tab[2][stage_pos<SEG>(idx)] = vt[row];
  }
  __syncthreads();
  const long long nlines = (long long)g.ny * g.nz;
  const int wpb = TCK_BLOCK / TCK_WAVE;
  const long long stride = (long long)gridDim.x * wpb;
  /* the wave's index as a uniform value: the line, and so each buffer resource, is the wave's */
  const int wu = __builtin_amdgcn_readfirstlane(w);
  /* Buffer addressing: a line is n contiguous values from cell (i0, j, k), so a load at or past
   * n reads 0 and a store there writes nothing, and no access is conditional (a wave past the
   * last line loads and stores through a zero-size resource).  The line's increment (U: and
   * old field) are loaded into registers one line ahead, before this line's stores, so the
   * loads of the next line are in flight while this one is solved; outstanding loads and
   * stores complete in issue order as far as the wait counter is concerned, and the zero-size
   * stores before the loop give its entry the loads-then-stores sequence of an iteration's
   * end, so the put of a line waits for its loads and not for the stores before them. */
  const unsigned nbd = (unsigned)n * (unsigned)sizeof(D), nbu = (unsigned)n * 8u;
  const unsigned vod = (unsigned)lane * (unsigned)sizeof(D), vou = (unsigned)lane * 8u;
  D pd[SEG];
  double pu[SEG];
  #define XS_FETCH(ln) { \
    const long long l_ = (ln); \
    const bool in_ = l_ < nlines; \
    const long long b_ = in_ ? tck_idx(g, i0, 1 + (int)(l_ % g.ny), 1 + (int)(l_ / g.ny)) : 0; \
    const __amdgpu_buffer_rsrc_t rd_ = tck_rsrc_u(dd + b_, in_ ? nbd : 0u); \
    _Pragma("unroll") for(int e = 0; e < SEG; e++) pd[e] = tck_bldn<D>(rd_, vod, (unsigned)(e * TCK_WAVE * (int)sizeof(D))); \
    if constexpr (U){ \
      const __amdgpu_buffer_rsrc_t ru_ = tck_rsrc_u(ucur + b_, in_ ? nbu : 0u); \
      _Pragma("unroll") for(int e = 0; e < SEG; e++) pu[e] = tck_bldn<double>(ru_, vou, (unsigned)(e * TCK_WAVE * 8)); \
    } }
  XS_FETCH((long long)blockIdx.x * wpb + wu);
  {
    const __amdgpu_buffer_rsrc_t nul = tck_rsrc(dd, 0u);
    #pragma unroll
    for(int e = 0; e < SEG; e++) tck_bst0(nul, 16u * e);
  }
  for(long long g0 = (long long)blockIdx.x * wpb; g0 < nlines; g0 += stride){
    const long long line = g0 + wu;
    const bool live = line < nlines;
    const int j = live ? 1 + (int)(line % g.ny) : 1;
    const int k = live ? 1 + (int)(line / g.ny) : 1;
    double u[SEG];
    #pragma unroll
    for(int e = 0; e < SEG; e++){
      sm[stage_pos<SEG>(lane + e * TCK_WAVE)] = double(pd[e]);
      if constexpr (U) u[e] = pu[e];
    }
    XS_FETCH(line + stride);
    const long long base = tck_idx(g, i0, j, k);
    wave_sync();
    /* the table offset is opaque to the compiler, so the reads stay inside the loop */
    int toff = lane * (SEG + 1);
    asm volatile("" : "+v"(toff));
    double q[SEG][1];
    #pragma unroll
    for(int e = 0; e < SEG; e++) q[e][0] = sm[lane * (SEG + 1) + e];
    wave_tdm<SEG, 1>(q, &tab[0][toff], &tab[1][toff], &tab[2][toff], n);
    #pragma unroll
    for(int e = 0; e < SEG; e++) sm[lane * (SEG + 1) + e] = q[e][0];
    wave_sync();
    if constexpr (U){
      const __amdgpu_buffer_rsrc_t ro = tck_rsrc_u(uout + base, live ? nbu : 0u);
      #pragma unroll
      for(int e = 0; e < SEG; e++)
        tck_bstn<double>(ro, vou, (unsigned)(e * TCK_WAVE * 8), u[e] + sm[stage_pos<SEG>(lane + e * TCK_WAVE)]);
      /* the halo images of the lines next to the periodic boundaries */
      if(live && (j == 1 || j == g.ny || k == 1 || k == g.nz)){
        #pragma unroll
        for(int e = 0; e < SEG; e++){
          const int idx = lane + e * TCK_WAVE;
          if(idx < n) tck_store_halo<double, false>(uout, g, i0 + idx, j, k, u[e] + sm[stage_pos<SEG>(idx)], false);
        }
      }
    }else{
      const __amdgpu_buffer_rsrc_t ro = tck_rsrc_u(dd + base, live ? nbd : 0u);
      #pragma unroll
      for(int e = 0; e < SEG; e++)
        tck_bstn<D>(ro, vod, (unsigned)(e * TCK_WAVE * (int)sizeof(D)), D(sm[stage_pos<SEG>(lane + e * TCK_WAVE)]));
    }
    wave_sync();
  }
  #undef XS_FETCH
}

template <int SEG, typename D, bool U>
__global__ void __launch_bounds__(TCK_BLOCK) __attribute__((amdgpu_waves_per_eu(SEG <= 8 ? 4 : 1))) tck_xsolve(const XSolveArgs A){
  __shared__ double xs_sm[TCK_BLOCK / TCK_WAVE][TCK_WAVE * (SEG + 1)];
  __shared__ double xs_tab[3][TCK_WAVE * (SEG + 1)];
  const int f = blockIdx.y;
  xsolve_body<SEG, D, U>(A, (D *)A.d[f], A.uout[f], A.ucur[f], A.r[f], A.v[f], A.l[f], A.i0[f], A.n[f],
      xs_sm[threadIdx.x / TCK_WAVE], xs_tab);
}

/* periodic lines in y (solve_in_y, :39-70) or z (solve_in_z, :72-103).  A block holds TI
 * lines (one per lane column, consecutive i, so each row access is a TI * 8 B contiguous
 * chunk) split into S segments along the line; every thread keeps its segment in registers,
 * so each element is read once and written once.  The forward and backward substitutions of
 * the (n-1) leading block are affine recurrences: each segment composes its rows, a fold over
 * the earlier (later) segments in LDS gives the value entering it, and the segment is replayed
 * with the reference's update formulas (tdm.c:17-50).  The Sherman-Morrison step
 * (tdm.c:92-98) then needs q0[0], q0[n-2] and q[n-1], shared through LDS.
 * Factor rows: below the convergence row jc they come from the tables, loaded ahead of each
 * recurrence in groups of PS_PRE; from jc on they are the constants rc, vc (and rlast for the
 * last row), so most segments wait on no table load at all. */
#define PS_MAXM 32
#define PS_PRE 8

/* the solves of up to TCK_XS_MAXF fields in one direction, one field per blockIdx.z */
struct PSolveArgs {
  tc_geom_t g;
  void * d[TCK_XS_MAXF];              /* the increment: double, or float in the mixed step */
  double * uout[TCK_XS_MAXF];         /* last pass: the new field (the increment's buffer itself in FP64) */
  const double * ucur[TCK_XS_MAXF];
  const double * r[TCK_XS_MAXF]; const double * v[TCK_XS_MAXF]; const double * q1[TCK_XS_MAXF];
  const double * meta[TCK_XS_MAXF];  /* TC_PMETA rows [7][tstride] indexed by i (y) or [7] (z) */
  long long tstride;    /* table stride between rows (tp for y, 1 for z) */
  int tcol;             /* 1: tables indexed by i as well (y) */
  int dir;              /* 1: y, 2: z */
  int nti, nsg;         /* block shape (TI lines, S segments), as launched */
  int lean;             /* bit 1: waves of interior full-length segments take the lean form;
                           bit 2: the folds read LDS four segments at a time */
  int i0[TCK_XS_MAXF], isz[TCK_XS_MAXF];
};

/* factor rows row0+e0 .. row0+e0+PS_PRE-1 of table t: rows below jc from the table (loaded
 * together, ahead of use), rows from jc on the constant cval, row m-1 lval */
template <typename Q>
__device__ __forceinline__ void ps_rows(Q * out, const double * __restrict__ t, long long ts,
    int row0, int e0, int jc, int m, Q cval, Q lval){
  if(row0 + e0 < jc){
    double tmp[PS_PRE];
    #pragma unroll
    for(int e = 0; e < PS_PRE; e++){
      const int row = row0 + e0 + e;
      tmp[e] = t[ts * (row < m ? row : m - 1)];
    }
    #pragma unroll
    for(int e = 0; e < PS_PRE; e++){
      const int row = row0 + e0 + e;
      out[e] = row < jc ? Q(tmp[e]) : (row == m - 1 ? lval : cval);
    }
  }else{
    #pragma unroll
    for(int e = 0; e < PS_PRE; e++) out[e] = row0 + e0 + e == m - 1 ? lval : cval;
  }
}

/* plain store (as a halo image's) at byte offset vo + so of resource r */
__device__ __forceinline__ void ps_bsti(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, double v){
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tck_u2v, v), r, vo, so, 0);
}

/* body with restrict-qualified pointers: loads of the old field and the tables may be
 * scheduled ahead of the stores to the increment array.
 * Two forms of the same operations, chosen per wave: the general one tests every row (row 0,
 * row m-1, rows past the segment's count) and forms 64-bit addresses per row; the lean one,
 * taken when every lane of the wave holds a full segment (MAXM rows) of interior rows, has no
 * row tests and reaches the rows through buffer resources based at the wave's first lane, with
 * a 32-bit offset per lane and a uniform one per row (the launcher checks that the offsets of
 * a wave stay below 2^31).  The arithmetic of a row is the same in both. */
template <typename D, int MAXM, bool LEAN>
__device__ __forceinline__ void psolve_body(const PSolveArgs & A, D * __restrict__ d, double * uout,
    const double * __restrict__ ucur, const double * __restrict__ rt, const double * __restrict__ vt,
    const double * __restrict__ q1t, const double * __restrict__ meta, const int ib0, const int isz,
    double * ps_sm){
  const tc_geom_t & g = A.g;
  const int TI = A.nti;
  const int S = A.nsg;
  const int ti = threadIdx.x;
  const int sg = threadIdx.y;
  /* tiles start at i = 1 (a 128 B line boundary); rows below i0 are skipped (ux: i0 = 2) */
  const int i = 1 + blockIdx.x * TI + ti;
  const int other = 1 + blockIdx.y;    /* k for y lines, j for z lines */
  const bool valid = i >= ib0 && i < ib0 + isz;
  const int n = A.dir == 1 ? g.ny : g.nz;
  const int m = n - 1;
  const int seg = (m + S - 1) / S;
  const int row0 = sg * seg;
  int cnt = m - row0;
  cnt = cnt < 0 ? 0 : (cnt > seg ? seg : cnt);
  const long long step = A.dir == 1 ? (long long)g.px : g.plane;
  /* lanes outside the field's rows work on its nearest row (nothing of theirs is stored), so
   * that the addresses of a wave do not decrease with the lane */
  const int ihi = isz > 0 ? ib0 + isz - 1 : ib0;
  const int ic = i < ib0 ? ib0 : (i > ihi ? ihi : i);
  const long long base = A.dir == 1 ? tck_idx(g, ic, 1, other) : tck_idx(g, ic, other, 1);
  const int col = A.tcol ? ic : 0;
  const long long ts = A.tstride;
  const double * r = rt + col;
  const double * v = vt + col;
  const double * q1 = q1t + col;
  typedef D Q;   /* arithmetic type: the increment's own */
  const Q den = Q(meta[col]);
  const Q tl = Q(meta[ts + col]);
  const Q tu = Q(meta[2 * ts + col]);
  const int jc = (int)meta[3 * ts + col];
  const Q rc = Q(meta[4 * ts + col]);
  const Q vc = Q(meta[5 * ts + col]);
  const Q rlast = Q(meta[6 * ts + col]);
  Q * smA = (Q *)ps_sm;                  /* [S][TI] */
  Q * smB = smA + S * TI;         /* [S][TI] */
  Q * smX = smA + 2 * S * TI;     /* [3][TI]: q0[0], q0[n-2], q[n-1] */
  /* the lean form's addressing: element offsets from the wave's first lane (the smallest),
   * one row further per `step` elements; table rows at (col + ts row) */
  const long long lo = base + step * row0;
  const long long w0 = tck_uni64(lo);
  const unsigned vel = (unsigned)(lo - w0);
  const unsigned sel = (unsigned)step;
  unsigned vtab = (unsigned)((col + ts * row0) * 8);
  const unsigned stab = (unsigned)(ts * 8);
  Q q[MAXM];
  if constexpr(LEAN){
    const __amdgpu_buffer_rsrc_t rd = tck_rsrc_u(d + w0, 0xFFFFFFFFu);
    #pragma unroll
    for(int e = 0; e < MAXM; e++)
      q[e] = Q(tck_bldn<D>(rd, vel * (unsigned)sizeof(D), (unsigned)e * sel * (unsigned)sizeof(D)));
  }else{
    #pragma unroll
    for(int e = 0; e < MAXM; e++) q[e] = (valid && e < cnt) ? Q(tck_ld(&d[base + step * (row0 + e)])) : Q(0.);
  }
  if(valid && sg == S - 1) smX[2 * TI + ti] = Q(d[base + step * m]);
  Q t[PS_PRE];
  /* forward: q0 = q0 / c0, q_i = r_i (q_i - l q_{i-1}) */
  Q a = Q(1.), b = Q(0.);
  if constexpr(LEAN){
    /* rows from jc on: the factors are the constants */
    #pragma unroll
    for(int e = 0; e < MAXM; e++){
      asm volatile("" : "+v"(q[e]));   /* row by row: keeps the products from being formed all ahead */
      const Q ai = - rc * tl;
      b = ai * b + rc * q[e];
      a = ai * a;
    }
  }else{
    #pragma unroll
    for(int e0 = 0; e0 < MAXM; e0 += PS_PRE){
      if(e0 < cnt){
        ps_rows(t, r, ts, row0, e0, jc, m, rc, rlast);
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          if(e0 + e < cnt){
            const int row = row0 + e0 + e;
            if(row == 0){ a = Q(0.); b = q[e0 + e] / t[e]; }
            else {
              const Q ai = - t[e] * tl;
              b = ai * b + t[e] * q[e0 + e];
              a = ai * a;
            }
          }
        }
      }
    }
  }
  smA[sg * TI + ti] = a;
  smB[sg * TI + ti] = b;
  __syncthreads();
  /* the fold over the earlier segments, in order; LDS reads four segments at a time so that
   * one wait covers them */
  Q in = Q(0.);
  {
    int s2 = 0;
    if(A.lean & 2) for(; s2 + 4 <= sg; s2 += 4){
      Q fa[4], fb[4];
      #pragma unroll
      for(int u = 0; u < 4; u++){ fa[u] = smA[(s2 + u) * TI + ti]; fb[u] = smB[(s2 + u) * TI + ti]; }
      #pragma unroll
      for(int u = 0; u < 4; u++) in = fa[u] * in + fb[u];
    }
    for(; s2 < sg; s2++) in = smA[s2 * TI + ti] * in + smB[s2 * TI + ti];
  }
  if constexpr(LEAN){
    #pragma unroll
    for(int e = 0; e < MAXM; e++){
      asm volatile("" : "+v"(q[e]));
      q[e] = rc * (q[e] - tl * in);
      in = q[e];
    }
  }else{
    #pragma unroll
    for(int e0 = 0; e0 < MAXM; e0 += PS_PRE){
      if(e0 < cnt){
        ps_rows(t, r, ts, row0, e0, jc, m, rc, rlast);
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          if(e0 + e < cnt){
            const int row = row0 + e0 + e;
            if(row == 0) q[e0 + e] = q[e0 + e] / t[e];
            else q[e0 + e] = t[e] * (q[e0 + e] - tl * in);
            in = q[e0 + e];
            if(row == m - 1) smX[TI + ti] = q[e0 + e];
          }
        }
      }
    }
  }
  __syncthreads();
  /* backward: q_i -= v_i q_{i+1}, q0[n-2] kept */
  a = Q(1.);
  b = Q(0.);
  if constexpr(LEAN){
    #pragma unroll
    for(int e = MAXM - 1; e >= 0; e--){
      asm volatile("" : "+v"(q[e]));
      const Q ai = - vc;
      b = ai * b + q[e];
      a = ai * a;
    }
  }else{
    #pragma unroll
    for(int e0 = MAXM - PS_PRE; e0 >= 0; e0 -= PS_PRE){
      if(e0 < cnt){
        ps_rows(t, v, ts, row0, e0, jc, m, vc, Q(0.));
        #pragma unroll
        for(int e = PS_PRE - 1; e >= 0; e--){
          if(e0 + e < cnt){
            const int row = row0 + e0 + e;
            if(row == m - 1){ a = Q(0.); b = q[e0 + e]; }
            else {
              const Q ai = - t[e];
              b = ai * b + q[e0 + e];
              a = ai * a;
            }
          }
        }
      }
    }
  }
  smA[sg * TI + ti] = a;
  smB[sg * TI + ti] = b;
  __syncthreads();
  in = Q(0.);
  {
    int s2 = S - 1;
    if(A.lean & 2) for(; s2 - 4 >= sg; s2 -= 4){
      Q fa[4], fb[4];
      #pragma unroll
      for(int u = 0; u < 4; u++){ fa[u] = smA[(s2 - u) * TI + ti]; fb[u] = smB[(s2 - u) * TI + ti]; }
      #pragma unroll
      for(int u = 0; u < 4; u++) in = fa[u] * in + fb[u];
    }
    for(; s2 > sg; s2--) in = smA[s2 * TI + ti] * in + smB[s2 * TI + ti];
  }
  if constexpr(LEAN){
    #pragma unroll
    for(int e = MAXM - 1; e >= 0; e--){
      asm volatile("" : "+v"(q[e]));
      q[e] -= vc * in;
      in = q[e];
    }
  }else{
    #pragma unroll
    for(int e0 = MAXM - PS_PRE; e0 >= 0; e0 -= PS_PRE){
      if(e0 < cnt){
        ps_rows(t, v, ts, row0, e0, jc, m, vc, Q(0.));
        #pragma unroll
        for(int e = PS_PRE - 1; e >= 0; e--){
          if(e0 + e < cnt){
            const int row = row0 + e0 + e;
            if(row != m - 1) q[e0 + e] -= t[e] * in;
            in = q[e0 + e];
          }
        }
      }
    }
  }
  if(cnt > 0 && row0 == 0) smX[ti] = q[0];
  __syncthreads();
  /* x_{n-1} (tdm.c:92-94) and x_i += x_{n-1} q1_i (:96-98) */
  const Q num = smX[2 * TI + ti] - tu * smX[ti] - tl * smX[TI + ti];
  const Q xl = fabs(double(den)) < DBL_EPSILON ? Q(0.) : num / den;
  if(!valid) return;
  if constexpr(LEAN){
    /* interior rows only: no row m, and the halo images are those of the other index (a
     * uniform offset, when it is at a wall of the period) */
    const __amdgpu_buffer_rsrc_t rq = tck_rsrc_u(q1t, 0xFFFFFFFFu);
    unsigned vd = vel * (unsigned)sizeof(D);
    const unsigned sd = sel * (unsigned)sizeof(D);
    unsigned vu = vel * 8u;
    const unsigned su = sel * 8u;
    long long dimg = 0;
    if(A.dir == 1){
      if(other == 1) dimg = (long long)g.nz * g.plane;
      else if(other == g.nz) dimg = -(long long)g.nz * g.plane;
    }else{
      if(other == 1) dimg = (long long)g.ny * g.px;
      else if(other == g.ny) dimg = -(long long)g.ny * g.px;
    }
    const __amdgpu_buffer_rsrc_t rd = tck_rsrc_u(d + w0, 0xFFFFFFFFu);
    const __amdgpu_buffer_rsrc_t rc2 = tck_rsrc_u(ucur ? ucur + w0 : q1t, 0xFFFFFFFFu);
    const __amdgpu_buffer_rsrc_t ro = tck_rsrc_u(uout ? uout + w0 : q1t, 0xFFFFFFFFu);
    const __amdgpu_buffer_rsrc_t ri = tck_rsrc_u(uout ? uout + w0 + dimg : q1t, 0xFFFFFFFFu);
    if constexpr(sizeof(Q) == 8){
      /* every load before the first store, the results formed in place in q */
      #pragma unroll
      for(int e0 = 0; e0 < MAXM; e0 += PS_PRE){
        Q qq[PS_PRE];
        double uu[PS_PRE];
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          qq[e] = Q(tck_bld<double>(rq, vtab, (unsigned)(e0 + e) * stab));
          uu[e] = ucur ? tck_bldn<double>(rc2, vu, (unsigned)(e0 + e) * su) : 0.;
        }
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          const Q x = q[e0 + e] + xl * qq[e];
          q[e0 + e] = ucur ? Q(uu[e] + double(x)) : x;
        }
        /* one group's loads in registers at a time (all of them beside q would not fit): the
         * next group's offsets only after this group's results */
        static_assert(PS_PRE == 8, "row group of 8");
        asm volatile("" : "+v"(vtab), "+v"(vu), "+v"(q[e0]), "+v"(q[e0 + 1]), "+v"(q[e0 + 2]), "+v"(q[e0 + 3]),
            "+v"(q[e0 + 4]), "+v"(q[e0 + 5]), "+v"(q[e0 + 6]), "+v"(q[e0 + 7]));
      }
      if(ucur){
        #pragma unroll
        for(int e = 0; e < MAXM; e++) tck_bstn<double>(ro, vu, (unsigned)e * su, double(q[e]));
        if(dimg != 0){
          #pragma unroll
          for(int e = 0; e < MAXM; e++) ps_bsti(ri, vu, (unsigned)e * su, double(q[e]));
        }
      }else{
        #pragma unroll
        for(int e = 0; e < MAXM; e++) tck_bstn<D>(rd, vd, (unsigned)e * sd, D(q[e]));
      }
    }else{
      /* q1 and the old field, PS_PRE rows at a time */
      #pragma unroll
      for(int e0 = 0; e0 < MAXM; e0 += PS_PRE){
        Q qq[PS_PRE];
        double uu[PS_PRE];
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          qq[e] = Q(tck_bld<double>(rq, vtab, (unsigned)(e0 + e) * stab));
          uu[e] = ucur ? tck_bldn<double>(rc2, vu, (unsigned)(e0 + e) * su) : 0.;
        }
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          const Q x = q[e0 + e] + xl * qq[e];
          if(ucur){
            const double y = uu[e] + double(x);
            tck_bstn<double>(ro, vu, (unsigned)(e0 + e) * su, y);
            if(dimg != 0) ps_bsti(ri, vu, (unsigned)(e0 + e) * su, y);
          }else{
            tck_bstn<D>(rd, vd, (unsigned)(e0 + e) * sd, D(x));
          }
        }
        asm volatile("" : "+v"(vtab), "+v"(vu), "+v"(vd));   /* the next group's loads after these stores */
      }
    }
    return;
  }
  if constexpr(sizeof(Q) == 8){
    /* FP64: every load of the pass (q1, the old field) before the first store, the results
     * formed in place in q; then the stores.  A load issued after a store could be waited for
     * only together with the store. */
    double ulast = 0.;
    if(sg == S - 1 && ucur) ulast = ucur[base + step * m];
    #pragma unroll
    for(int e0 = 0; e0 < MAXM; e0 += PS_PRE){
      if(e0 < cnt){
        Q qq[PS_PRE];
        double uu[PS_PRE];
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          const int row = row0 + e0 + e < m ? row0 + e0 + e : m - 1;
          qq[e] = Q(q1[ts * row]);
          uu[e] = ucur ? ucur[base + step * row] : 0.;
        }
        #pragma unroll
        for(int e = 0; e < PS_PRE; e++){
          const Q x = q[e0 + e] + xl * qq[e];
          q[e0 + e] = ucur ? Q(uu[e] + double(x)) : x;
        }
      }
    }
    #pragma unroll
    for(int e = 0; e < MAXM; e++){
      if(e < cnt){
        const int row = row0 + e;
        if(ucur){
          const int jj = A.dir == 1 ? 1 + row : other;
          const int kk = A.dir == 1 ? other : 1 + row;
          tck_store_halo(uout, g, i, jj, kk, double(q[e]), false);
        }else{
          tck_st(&d[base + step * row], D(q[e]));
        }
      }
    }
    if(sg == S - 1){
      if(ucur){
        const int jj = A.dir == 1 ? n : other;
        const int kk = A.dir == 1 ? other : n;
        tck_store_halo(uout, g, i, jj, kk, ulast + double(xl), false);
      }else{
        tck_st(&d[base + step * m], D(xl));
      }
    }
    return;
  }
  /* q1 and the old field, PS_PRE rows at a time */
  #pragma unroll
  for(int e0 = 0; e0 < MAXM; e0 += PS_PRE){
    if(e0 < cnt){
      Q qq[PS_PRE];
      double uu[PS_PRE];
      #pragma unroll
      for(int e = 0; e < PS_PRE; e++){
        const int row = row0 + e0 + e < m ? row0 + e0 + e : m - 1;
        qq[e] = Q(q1[ts * row]);
        uu[e] = ucur ? ucur[base + step * row] : 0.;
      }
      #pragma unroll
      for(int e = 0; e < PS_PRE; e++){
        if(e0 + e < cnt){
          const int row = row0 + e0 + e;
          const Q x = q[e0 + e] + xl * qq[e];
          if(ucur){
            const int jj = A.dir == 1 ? 1 + row : other;
            const int kk = A.dir == 1 ? other : 1 + row;
            tck_store_halo(uout, g, i, jj, kk, uu[e] + double(x), false);
          }else{
            tck_st(&d[base + step * row], D(x));
          }
        }
      }
    }
  }
  if(sg == S - 1){
    if(ucur){
      const int jj = A.dir == 1 ? n : other;
      const int kk = A.dir == 1 ? other : n;
      tck_store_halo(uout, g, i, jj, kk, ucur[base + step * m] + double(xl), false);
    }else{
      tck_st(&d[base + step * m], D(xl));
    }
  }
}

template <typename D, int MAXM>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) tck_psolve(const PSolveArgs A){
  extern __shared__ double ps_sm[];
  const int f = blockIdx.z;
  /* the lean form where every lane of the wave holds a full segment of interior rows at or
   * past the convergence row jc (constant factors; a uniform branch: each form passes the
   * same three barriers) */
  const int m = (A.dir == 1 ? A.g.ny : A.g.nz) - 1;
  const int seg = (m + A.nsg - 1) / A.nsg;
  const int row0 = (int)threadIdx.y * seg;
  const int i = 1 + blockIdx.x * A.nti + threadIdx.x;
  const int ib0 = A.i0[f], ihi = A.isz[f] > 0 ? ib0 + A.isz[f] - 1 : ib0;
  const int ic = i < ib0 ? ib0 : (i > ihi ? ihi : i);
  const int jc = (int)A.meta[f][3 * A.tstride + (A.tcol ? ic : 0)];
  const bool full = (A.lean & 1) && seg == MAXM && row0 > 0 && row0 >= jc && row0 + MAXM <= m - 1;
  if(__builtin_amdgcn_ballot_w64(!full) == 0)
    psolve_body<D, MAXM, true>(A, (D *)A.d[f], A.uout[f], A.ucur[f], A.r[f], A.v[f], A.q1[f], A.meta[f],
        A.i0[f], A.isz[f], ps_sm);
  else
    psolve_body<D, MAXM, false>(A, (D *)A.d[f], A.uout[f], A.ucur[f], A.r[f], A.v[f], A.q1[f], A.meta[f],
        A.i0[f], A.isz[f], ps_sm);
}

/* ---------------------------------------------------------------------------------------
 * Poisson radial systems (compute_potential.c:445-488): per line (ky, kz)
 *   c_i = - l_i - u_i + ly(ky) / HYXC(i)^2 + lz(kz) / hz^2,  c_0 += l_0,  c_{n-1} += u_{n-1}
 * factors r precomputed once into pr (layout [line][i]); the other factor, v_i = r_i u_i
 * (v_0 = u_0 / c_0, v_{n-1} = 0), is formed again by the solve with the same operation, so
 * that only r is read; rows of the spectrum are complex.
 * ------------------------------------------------------------------------------------- */
struct PFactorArgs {
  int nx, nky, nz;
  int exact_gauge;   /* 1: the (0, 0) line takes x_{n-1} = 0 whatever its last pivot */
  const double * pl; const double * pu;     /* l_i, u_i (i = 0..nx-1) */
  const double * hyxc;                      /* HYXC(i), reference index */
  const double * evy; const double * evz;
  double hz;
  double * pr;
};

__global__ void tck_pfactor(const PFactorArgs A){
  const long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if(line >= (long long)A.nky * A.nz) return;
  const int ky = (int)(line % A.nky);
  const int kz = (int)(line / A.nky);
  const double eval_y = A.evy[ky];
  const double eval_z = A.evz[kz];
  const int n = A.nx;
  double * r = A.pr + line * n;
  const double hz = A.hz;
  /* the Neumann fix-ups are added after the sum, as c[0] += l[0] does (:481-482) */
  double c0 = - A.pl[0] - A.pu[0] + eval_y / A.hyxc[1] / A.hyxc[1] + eval_z / hz / hz;
  c0 += A.pl[0];
  double vprev = A.pu[0] / c0;
  r[0] = c0;
  for(int i = 1; i < n - 1; i++){
    const double ci = - A.pl[i] - A.pu[i] + eval_y / A.hyxc[i + 1] / A.hyxc[i + 1] + eval_z / hz / hz;
    const double val = 1. / (ci - A.pl[i] * vprev);
    vprev = val * (A.pu[i]);
    r[i] = val;
  }
  double cl = - A.pl[n - 1] - A.pu[n - 1] + eval_y / A.hyxc[n] / A.hyxc[n] + eval_z / hz / hz;
  cl += A.pu[n - 1];
  const double val = cl - A.pl[n - 1] * vprev;
  r[n - 1] = fabs(val) > DBL_EPSILON ? 1. / val : 0.;
  /* the Neumann operator of the (0, 0) mode is singular; rounding leaves its last pivot a few
   * ulps from zero, so tdm.c's rule may divide by it.  That is harmless when the right-hand
   * side sums to zero to FP64 rounding; an FP32 spectrum does not, and the division would
   * swamp psi with a huge constant.  The mixed step fixes the gauge exactly instead. */
  if(A.exact_gauge && line == 0) r[n - 1] = 0.;
}

struct PXSolveArgs {
  int nx, nlines;
  void * s;                  /* spectrum, line-major rows of nx complex (double2, or float2
                              * in the mixed step; the solve itself runs in FP64) */
  const double * pr; const double * pl; const double * pu;
};

/* one (ky, kz) line per wave per iteration, WPB waves per block looping over groups of WPB
 * lines; the spectrum row and its factor row are staged through LDS so that every global
 * access is contiguous across the wave (WPB keeps the staging area within 64 KB) */
template <int SEG, int WPB, typename C2>
__global__ void __launch_bounds__(WPB * TCK_WAVE) tck_pxsolve(const PXSolveArgs A){
  __shared__ double2 smq_all[WPB][TCK_WAVE * (SEG + 1)];
  __shared__ double smr_all[WPB][TCK_WAVE * (SEG + 1)];
  const int w = threadIdx.x / TCK_WAVE;
  const int lane = threadIdx.x % TCK_WAVE;
  double2 * smq = smq_all[w];
  double * smr = smr_all[w];
  const int n = A.nx;
  const int row0 = lane * SEG;
  double ll[SEG], uu[SEG];
  #pragma unroll
  for(int e = 0; e < SEG; e++){
    ll[e] = A.pl[row0 + e < n ? row0 + e : n - 1];
    uu[e] = A.pu[row0 + e < n ? row0 + e : n - 1];
  }
  for(long long g0 = (long long)blockIdx.x * WPB; g0 < A.nlines; g0 += (long long)gridDim.x * WPB){
    const long long line = g0 + w;
    const bool live = line < A.nlines;
    C2 * row = (C2 *)A.s + (live ? line : 0) * n;
    const double * r = A.pr + (live ? line : 0) * n;
    if(live){
      #pragma unroll
      for(int e = 0; e < SEG; e++){
        const int idx = lane + e * TCK_WAVE;
        if(idx < n){
          const int p = stage_pos<SEG>(idx);
          const C2 x = row[idx];
          smq[p] = make_double2(double(x.x), double(x.y));
          smr[p] = r[idx];
        }
      }
    }
    __syncthreads();
    double q[SEG][2], rr[SEG], vv[SEG];
    #pragma unroll
    for(int e = 0; e < SEG; e++){
      const double2 x = smq[lane * (SEG + 1) + e];
      q[e][0] = x.x;
      q[e][1] = x.y;
      rr[e] = smr[lane * (SEG + 1) + e];
      /* v as the factor kernel forms it: u_0 / c_0 (r_0 holds c_0), r_i u_i, 0 at the end */
      const int row = row0 + e;
      vv[e] = row == 0 ? uu[e] / rr[e] : (row == n - 1 ? 0. : rr[e] * uu[e]);
    }
    wave_tdm<SEG, 2>(q, rr, ll, vv, n);
    #pragma unroll
    for(int e = 0; e < SEG; e++) smq[lane * (SEG + 1) + e] = make_double2(q[e][0], q[e][1]);
    __syncthreads();
    if(live){
      #pragma unroll
      for(int e = 0; e < SEG; e++){
        const int idx = lane + e * TCK_WAVE;
        if(idx < n){
          const double2 x = smq[stage_pos<SEG>(idx)];
          row[idx].x = x.x;
          row[idx].y = x.y;
        }
      }
    }
    __syncthreads();
  }
}

/* ------------------------------------------------------------------------------------- */
static int seg_for(int n){
  int seg = 1;
  while(seg * TCK_WAVE < n) seg <<= 1;
  return seg;
}

template <int SEG>
static int launch_xsolve_t(const XSolveArgs & A, int nf, int mixed, hipStream_t stream){
  const long long nlines = (long long)A.g.ny * A.g.nz;
  const int wpb = TCK_BLOCK / TCK_WAVE;
  long long nblocks = (nlines + wpb - 1) / wpb;
  const long long cap = (long long)num_cus() * 8;
  if(nblocks > cap) nblocks = cap;
  const dim3 grd((unsigned)nblocks, (unsigned)nf);
  /* the final pass (the new fields written from the old ones) is a form of its own */
  const bool u = A.ucur[0] != NULL;
  if(mixed){
    if(u) hipLaunchKernelGGL((tck_xsolve<SEG, float, true>), grd, dim3(TCK_BLOCK), 0, stream, A);
    else hipLaunchKernelGGL((tck_xsolve<SEG, float, false>), grd, dim3(TCK_BLOCK), 0, stream, A);
  }else{
    if(u) hipLaunchKernelGGL((tck_xsolve<SEG, double, true>), grd, dim3(TCK_BLOCK), 0, stream, A);
    else hipLaunchKernelGGL((tck_xsolve<SEG, double, false>), grd, dim3(TCK_BLOCK), 0, stream, A);
  }
  TCK_CHECK_LAUNCH();
  return 0;
}

template <int SEG>
static int launch_pxsolve_t(const PXSolveArgs & A, int mixed, hipStream_t stream){
  constexpr int WPB = SEG >= 16 ? 1 : 2;
  long long nblocks = ((long long)A.nlines + WPB - 1) / WPB;
  const long long cap = (long long)num_cus() * 16 / WPB;
  if(nblocks > cap) nblocks = cap;
  if(mixed) hipLaunchKernelGGL((tck_pxsolve<SEG, WPB, float2>), dim3((unsigned)nblocks), dim3(WPB * TCK_WAVE), 0, stream, A);
  else hipLaunchKernelGGL((tck_pxsolve<SEG, WPB, double2>), dim3((unsigned)nblocks), dim3(WPB * TCK_WAVE), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

extern "C" {

/* jobs: njobs descriptors laid out as in FactorJob; fields beyond the implicit set may be null */
int tck_launch_factor(const double * dt, const int dirs[3], int nx, int ny, int nz, int tp,
    double lzl, double lzc, double lzu, int njobs, const int * rkstep, const double * diff,
    const double (*fac)[3], const int * i0, const int * n,
    const double * const (*lx)[3], const double * const (*ly)[3],
    double * const (*xt)[3], double * const (*yt)[4], double * const (*zt)[4], hipStream_t stream){
  if(njobs < 1 || njobs > FACTOR_MAXJOB) return (int)hipErrorInvalidValue;
  FactorBatch B;
  B.dt = dt;
  for(int d = 0; d < 3; d++) B.dirs[d] = dirs[d];
  B.nx = nx; B.ny = ny; B.nz = nz; B.tp = tp;
  B.lzl = lzl; B.lzc = lzc; B.lzu = lzu;
  B.njobs = njobs;
  for(int q = 0; q < njobs; q++){
    FactorJob & J = B.jobs[q];
    J.rkstep = rkstep[q]; J.diff = diff[q];
    for(int d = 0; d < 3; d++) J.fac[d] = fac[q][d];
    J.i0 = i0[q]; J.n = n[q];
    J.lxl = lx[q][0]; J.lxc = lx[q][1]; J.lxu = lx[q][2];
    J.lyl = ly[q][0]; J.lyc = ly[q][1]; J.lyu = ly[q][2];
    J.xr = xt[q][0]; J.xv = xt[q][1]; J.xl = xt[q][2];
    J.yr = yt[q][0]; J.yv = yt[q][1]; J.yq1 = yt[q][2]; J.ymeta = yt[q][3];
    J.zr = zt[q][0]; J.zv = zt[q][1]; J.zq1 = zt[q][2]; J.zmeta = zt[q][3];
  }
  const int nthreads = nx + 2;
  hipLaunchKernelGGL(tck_factor, dim3((nthreads + 63) / 64, njobs), dim3(64), 0, stream, B);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_xsolve(const tc_geom_t * g, int nf, void * const * d, double * const * uout,
    const double * const * ucur, const double * const * r, const double * const * v,
    const double * const * l, const int * i0, const int * n, int mixed, hipStream_t stream){
  if(nf < 1 || nf > TCK_XS_MAXF) return (int)hipErrorInvalidValue;
  XSolveArgs A;
  A.g = *g;
  int nmax = 0;
  for(int f = 0; f < nf; f++){
    if(n[f] > TCK_WAVE * XS_MAXSEG || n[f] < 2) return (int)hipErrorInvalidValue;
    if((ucur[f] == NULL) != (ucur[0] == NULL)) return (int)hipErrorInvalidValue;   /* all final or none */
    A.d[f] = d[f]; A.uout[f] = uout[f]; A.ucur[f] = ucur[f];
    A.r[f] = r[f]; A.v[f] = v[f]; A.l[f] = l[f]; A.i0[f] = i0[f]; A.n[f] = n[f];
    if(n[f] > nmax) nmax = n[f];
  }
  /* the segment count of the longest line: shorter lines leave their last lanes idle */
  switch(seg_for(nmax)){
    case 1: return launch_xsolve_t<1>(A, nf, mixed, stream);
    case 2: return launch_xsolve_t<2>(A, nf, mixed, stream);
    case 4: return launch_xsolve_t<4>(A, nf, mixed, stream);
    case 8: return launch_xsolve_t<8>(A, nf, mixed, stream);
    default: return launch_xsolve_t<16>(A, nf, mixed, stream);
  }
}

int tck_launch_psolve(const tc_geom_t * g, int dir, int nf, void * const * d, double * const * uout,
    const double * const * ucur, const double * const * r, const double * const * v,
    const double * const * q1, const double * const * meta, long long tstride, int tcol,
    const int * i0, const int * isz, int mixed, int lean, hipStream_t stream){
  if(nf < 1 || nf > TCK_XS_MAXF) return (int)hipErrorInvalidValue;
  PSolveArgs A;
  A.g = *g;
  A.tstride = tstride; A.tcol = tcol; A.dir = dir;
  int span = 0;   /* tiles cover i = 1 .. max(i0+isz-1) */
  for(int f = 0; f < nf; f++){
    A.d[f] = d[f]; A.uout[f] = uout[f]; A.ucur[f] = ucur[f];
    A.r[f] = r[f]; A.v[f] = v[f]; A.q1[f] = q1[f]; A.meta[f] = meta[f];
    A.i0[f] = i0[f]; A.isz[f] = isz[f];
    if(i0[f] + isz[f] - 1 > span) span = i0[f] + isz[f] - 1;
  }
  const int n = dir == 1 ? g->ny : g->nz;
  const int m = n - 1;
  /* the mixed step keeps twice the rows per thread (FP32 registers) so that a row access of
   * TI lanes is a whole 128 B line */
  const int maxm = mixed ? 2 * PS_MAXM : PS_MAXM;
  const int S = (m + maxm - 1) / maxm;
  if(S > 32) return (int)hipErrorInvalidValue;     /* n <= 1025: blocks of at most 512 threads */
  const int timin = mixed ? 32 : 16;
  int TI = 512 / S;
  TI = TI > 64 ? 64 : (TI < timin ? timin : TI);
  if(TI * S > 512) return (int)hipErrorInvalidValue;
  const int nother = dir == 1 ? g->nz : g->ny;
  const size_t lds = sizeof(double) * (size_t)(2 * S * TI + 3 * TI);
  const dim3 grd((span + TI - 1) / TI, nother, nf);
  A.nti = TI;
  A.nsg = S;
  /* the lean form's byte offsets: within a wave (its lanes span at most 64 / TI + 2 segments
   * and TI columns) plus a segment's rows, and the table rows, all below 2^31 */
  const long long step = dir == 1 ? (long long)g->px : g->plane;
  const long long seg = (m + S - 1) / S;
  const long long wspan = ((long long)(64 / TI + 2) * seg + maxm) * step + 64;
  const long long tspan = ((long long)(tcol ? tstride : 1) * (m + 1) + g->px) * 8;
  A.lean = (lean & 2) | ((lean & 1) && wspan * 8 < (1LL << 31) && tspan < (1LL << 31) ? 1 : 0);
  if(mixed) hipLaunchKernelGGL((tck_psolve<float, 2 * PS_MAXM>), grd, dim3(TI, S), lds, stream, A);
  else hipLaunchKernelGGL((tck_psolve<double, PS_MAXM>), grd, dim3(TI, S), lds, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_pfactor(int nx, int nky, int nz, const double * pl, const double * pu,
    const double * hyxc, const double * evy, const double * evz, double hz, double * pr,
    int exact_gauge, hipStream_t stream){
  PFactorArgs A;
  A.nx = nx; A.nky = nky; A.nz = nz; A.exact_gauge = exact_gauge; A.pl = pl; A.pu = pu; A.hyxc = hyxc; A.evy = evy;
  A.evz = evz; A.hz = hz; A.pr = pr;
  const long long nlines = (long long)nky * nz;
  hipLaunchKernelGGL(tck_pfactor, dim3((unsigned)((nlines + 63) / 64)), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_pxsolve(int nx, int nlines, void * s, const double * pr, const double * pl,
    const double * pu, int mixed, hipStream_t stream){
  if(nx > TCK_WAVE * XS_MAXSEG || nx < 2) return (int)hipErrorInvalidValue;
  PXSolveArgs A;
  A.nx = nx; A.nlines = nlines; A.s = s; A.pr = pr; A.pl = pl; A.pu = pu;
  switch(seg_for(nx)){
    case 1: return launch_pxsolve_t<1>(A, mixed, stream);
    case 2: return launch_pxsolve_t<2>(A, mixed, stream);
    case 4: return launch_pxsolve_t<4>(A, mixed, stream);
    case 8: return launch_pxsolve_t<8>(A, mixed, stream);
    default: return launch_pxsolve_t<16>(A, mixed, stream);
  }
}

}
