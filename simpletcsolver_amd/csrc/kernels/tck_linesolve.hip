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
__device__ void factor_sys(int n, const double * lc, const double * cc, const double * uc,
    double pf, long long stride, double * r, double * v, double * l){
  /* tdm coefficients as solve_in_x builds them (predict/linear_system.c:28-32) */
  #define TL(i) (- pf * lc[i])
  #define TC(i) (1. - pf * cc[i])
  #define TU(i) (- pf * uc[i])
  double vprev = TU(0) / TC(0);
  r[0] = TC(0);
  v[0] = vprev;
  l[0] = TL(0);
  for(int i = 1; i < n - 1; i++){
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
 * (n-1) block, the correction vector q1 and the Sherman-Morrison denominator */
__device__ void factor_periodic(int n, double tl, double tc, double tu, long long stride,
    double * r, double * v, double * q1, double * den){
  const int m = n - 1;
  /* factors of the (n-1) leading block */
  double vprev = tu / tc;
  r[0] = tc;
  v[0] = vprev;
  for(int i = 1; i < m - 1; i++){
    const double val = 1. / (tc - tl * vprev);
    vprev = val * (tu);
    r[stride * i] = val;
    v[stride * i] = vprev;
  }
  {
    const double val = tc - tl * vprev;
    r[stride * (m - 1)] = fabs(val) > DBL_EPSILON ? 1. / val : 0.;
    v[stride * (m - 1)] = 0.;
  }
  /* q1 = A_{n-1}^{-1} (-l0, 0, ..., 0, -u_{n-2}) by the same gtsv arithmetic */
  double qprev = 0.;
  for(int i = 0; i < m; i++){
    const double rhs = i == 0 ? - tl : i == m - 1 ? - tu : 0.;
    double q;
    if(i == 0) q = rhs / tc;
    else q = r[stride * i] * (rhs - tl * qprev);
    q1[stride * i] = q;
    qprev = q;
  }
  for(int i = m - 2; i >= 0; i--){
    q1[stride * i] -= v[stride * i] * q1[stride * (i + 1)];
  }
  *den = tc + tu * q1[0] + tl * q1[stride * (m - 1)];
}

/* implicit-diffusion factor tables, one job per (field, substep).
 *   x: one system of size n over i = i0..i0+n-1 with per-row stencils lap{l,c,u}[i]
 *   y: one periodic system per i (stencil lap{l,c,u}[i]) of size ny, table layout [j][tp]
 *   z: one periodic system of size nz (stencil lapz)
 * prefactor = fac * (0.5 * gamma * dt * diffusivity) (ux.c:496-506 and its siblings; fac is
 * 2 for the diagonal stress direction of each velocity and for T in y, t.c:346-349).
 * ymeta/zmeta hold {den, l, u} of the periodic systems ([3][tp] and [3]).  All jobs of a
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
    double den;
    factor_periodic(B.ny, tl, tc, tu, B.tp, A.yr + i, A.yv + i, A.yq1 + i, &den);
    A.ymeta[i] = den;
    A.ymeta[B.tp + i] = tl;
    A.ymeta[2 * B.tp + i] = tu;
  }
  if(t == B.nx + 1 && B.dirs[2]){
    const double pf = A.fac[2] * prefactor;
    const double tl = - pf * B.lzl;
    const double tc = 1. - pf * B.lzc;
    const double tu = - pf * B.lzu;
    double den;
    factor_periodic(B.nz, tl, tc, tu, 1, A.zr, A.zv, A.zq1, &den);
    A.zmeta[0] = den;
    A.zmeta[1] = tl;
    A.zmeta[2] = tu;
  }
}

/* ---------------------------------------------------------------------------------------
 * wave-level scans of affine maps x -> a x + b
 * ------------------------------------------------------------------------------------- */
/* value entering this lane's segment when maps compose in increasing lane order, x0 = 0 */
__device__ __forceinline__ void scan_fwd(double a, double * b, int nb, double * in){
  const int lane = threadIdx.x % TCK_WAVE;
  for(int off = 1; off < TCK_WAVE; off <<= 1){
    const double ap = __shfl_up(a, off, TCK_WAVE);
    double bp[2];
    for(int q = 0; q < nb; q++) bp[q] = __shfl_up(b[q], off, TCK_WAVE);
    if(lane >= off){
      for(int q = 0; q < nb; q++) b[q] = a * bp[q] + b[q];
      a = a * ap;
    }
  }
  for(int q = 0; q < nb; q++){
    const double prev = __shfl_up(b[q], 1, TCK_WAVE);
    in[q] = lane == 0 ? 0. : prev;
  }
}

/* same in decreasing lane order */
__device__ __forceinline__ void scan_bwd(double a, double * b, int nb, double * in){
  const int lane = threadIdx.x % TCK_WAVE;
  for(int off = 1; off < TCK_WAVE; off <<= 1){
    const double ap = __shfl_down(a, off, TCK_WAVE);
    double bp[2];
    for(int q = 0; q < nb; q++) bp[q] = __shfl_down(b[q], off, TCK_WAVE);
    if(lane + off < TCK_WAVE){
      for(int q = 0; q < nb; q++) b[q] = a * bp[q] + b[q];
      a = a * ap;
    }
  }
  for(int q = 0; q < nb; q++){
    const double next = __shfl_down(b[q], 1, TCK_WAVE);
    in[q] = lane == TCK_WAVE - 1 ? 0. : next;
  }
}

/* Solve one tridiagonal line held by a wavefront.  q[e][0..nc-1]: this lane's rows
 * row0 .. row0+cnt-1 (cnt may be 0).  r, v, l: factor rows as produced by factor_sys,
 * addressed by a per-row pointer offset. */
template <int NC>
__device__ __forceinline__ void wave_tdm(double (*q)[NC], const double * rr, const double * vv,
    const double * ll, int row0, int cnt, int n){
  /* forward: q_0 = q_0 / c_0, q_i = r_i (q_i - l_i q_{i-1}) */
  double a = 1., b[2] = {0., 0.};
  #pragma unroll
  for(int e = 0; e < XS_MAXSEG; e++){
    if(e < cnt){
      const int row = row0 + e;
      if(row == 0){
        a = 0.;
        for(int c = 0; c < NC; c++) b[c] = q[e][c] / rr[0];
      }else{
        const double ri = rr[row];
        const double ai = - ri * ll[row];
        for(int c = 0; c < NC; c++) b[c] = ai * b[c] + ri * q[e][c];
        a = ai * a;
      }
    }
  }
  double in[2];
  scan_fwd(a, b, NC, in);
  #pragma unroll
  for(int e = 0; e < XS_MAXSEG; e++){
    if(e < cnt){
      const int row = row0 + e;
      if(row == 0){
        for(int c = 0; c < NC; c++) q[e][c] = q[e][c] / rr[0];
      }else{
        const double ri = rr[row];
        const double li = ll[row];
        for(int c = 0; c < NC; c++) q[e][c] = ri * (q[e][c] - li * in[c]);
      }
      for(int c = 0; c < NC; c++) in[c] = q[e][c];
    }
  }
  /* backward: q_i -= v_i q_{i+1}, q_{n-1} kept */
  a = 1.;
  b[0] = b[1] = 0.;
  #pragma unroll
  for(int e = XS_MAXSEG - 1; e >= 0; e--){
    if(e < cnt){
      const int row = row0 + e;
      if(row == n - 1){
        a = 0.;
        for(int c = 0; c < NC; c++) b[c] = q[e][c];
      }else{
        const double ai = - vv[row];
        for(int c = 0; c < NC; c++) b[c] = ai * b[c] + q[e][c];
        a = ai * a;
      }
    }
  }
  scan_bwd(a, b, NC, in);
  #pragma unroll
  for(int e = XS_MAXSEG - 1; e >= 0; e--){
    if(e < cnt){
      const int row = row0 + e;
      if(row != n - 1){
        for(int c = 0; c < NC; c++) q[e][c] -= vv[row] * in[c];
      }
      for(int c = 0; c < NC; c++) in[c] = q[e][c];
    }
  }
}

/* x lines of the implicit diffusion solve (solve_in_x, predict/linear_system.c:6-37).
 * d: increment in the field layout, rows i = i0 .. i0+n-1 of each (j, k) line.
 * final: store ucur + solution with halos (the field update of ux.c:544-556). */
struct XSolveArgs {
  tc_geom_t g;
  double * d;
  const double * ucur;
  const double * r; const double * v; const double * l;
  int i0, n;
};

/* one x line per wavefront; the line is staged through LDS so that global loads and stores
 * are contiguous across the wave (lane t touches elements t, t+64, ...), while each lane
 * solves its contiguous segment from LDS.  Segment stride m+1 avoids bank aliasing. */
__device__ __forceinline__ void xsolve_body(const XSolveArgs & A, double * __restrict__ dd,
    const double * __restrict__ ucur, const double * __restrict__ rt, const double * __restrict__ vt,
    const double * __restrict__ lt, double * sm){
  const tc_geom_t & g = A.g;
  const int w = threadIdx.x / TCK_WAVE;
  const int wave = blockIdx.x * (TCK_BLOCK / TCK_WAVE) + w;
  const int lane = threadIdx.x % TCK_WAVE;
  const int nlines = g.ny * g.nz;
  const bool live = wave < nlines;
  const int j = live ? 1 + wave % g.ny : 1;
  const int k = live ? 1 + wave / g.ny : 1;
  const int n = A.n;
  const int seg = (n + TCK_WAVE - 1) / TCK_WAVE;
  const int ld = seg + 1;
  const int row0 = lane * seg;
  int cnt = n - row0;
  cnt = cnt < 0 ? 0 : (cnt > seg ? seg : cnt);
  const long long base = tck_idx(g, A.i0, j, k);
  if(live){
    for(int idx = lane; idx < n; idx += TCK_WAVE){
      const int ln = idx / seg;
      sm[ln * ld + (idx - ln * seg)] = dd[base + idx];
    }
  }
  __syncthreads();
  double q[XS_MAXSEG][1];
  #pragma unroll
  for(int e = 0; e < XS_MAXSEG; e++) q[e][0] = e < cnt ? sm[lane * ld + e] : 0.;
  wave_tdm<1>(q, rt, vt, lt, row0, cnt, n);
  #pragma unroll
  for(int e = 0; e < XS_MAXSEG; e++){
    if(e < cnt) sm[lane * ld + e] = q[e][0];
  }
  __syncthreads();
  if(!live) return;
  if(ucur){
    double u[XS_MAXSEG];
    #pragma unroll
    for(int e = 0; e < XS_MAXSEG; e++){
      const int idx = lane + e * TCK_WAVE;
      u[e] = idx < n ? ucur[base + idx] : 0.;
    }
    #pragma unroll
    for(int e = 0; e < XS_MAXSEG; e++){
      const int idx = lane + e * TCK_WAVE;
      if(idx < n){
        const int ln = idx / seg;
        tck_store_halo(dd, g, A.i0 + idx, j, k, u[e] + sm[ln * ld + (idx - ln * seg)], false);
      }
    }
  }else{
    for(int idx = lane; idx < n; idx += TCK_WAVE){
      const int ln = idx / seg;
      dd[base + idx] = sm[ln * ld + (idx - ln * seg)];
    }
  }
}

/* one x line per wavefront; the line is staged through LDS so that global loads and stores
 * are contiguous across the wave (lane t touches elements t, t+64, ...), while each lane
 * solves its contiguous segment from LDS.  Segment stride m+1 avoids bank aliasing. */
__global__ void __launch_bounds__(TCK_BLOCK) tck_xsolve(const XSolveArgs A){
  __shared__ double xs_sm[TCK_BLOCK / TCK_WAVE][TCK_WAVE * (XS_MAXSEG + 1)];
  xsolve_body(A, A.d, A.ucur, A.r, A.v, A.l, xs_sm[threadIdx.x / TCK_WAVE]);
}

/* periodic lines in y (solve_in_y, :39-70) or z (solve_in_z, :72-103).  A block holds TI
 * lines (one per lane column, consecutive i, so each row access is a TI * 8 B contiguous
 * chunk) split into S segments along the line; every thread keeps its segment in registers,
 * so each element is read once and written once.  The forward and backward substitutions of
 * the (n-1) leading block are affine recurrences: each segment composes its rows, a fold over
 * the earlier (later) segments in LDS gives the value entering it, and the segment is replayed
 * with the reference's update formulas (tdm.c:17-50).  The Sherman-Morrison step
 * (tdm.c:92-98) then needs q0[0], q0[n-2] and q[n-1], shared through LDS. */
#define PS_MAXM 32

struct PSolveArgs {
  tc_geom_t g;
  double * d;
  const double * ucur;
  const double * r; const double * v; const double * q1;
  const double * meta;  /* {den, l, u}: [3][tstride] indexed by i (y) or [3] (z) */
  long long tstride;    /* table stride between rows (tp for y, 1 for z) */
  int tcol;             /* 1: tables indexed by i as well (y) */
  int dir;              /* 1: y, 2: z */
  int i0, isz;
};

/* body with restrict-qualified pointers: loads of the old field and the tables may be
 * scheduled ahead of the stores to the increment array */
__device__ __forceinline__ void psolve_body(const PSolveArgs & A, double * __restrict__ d,
    const double * __restrict__ ucur, const double * __restrict__ rt, const double * __restrict__ vt,
    const double * __restrict__ q1t, const double * __restrict__ meta, double * ps_sm){
  const tc_geom_t & g = A.g;
  const int TI = blockDim.x;
  const int S = blockDim.y;
  const int ti = threadIdx.x;
  const int sg = threadIdx.y;
  /* tiles start at i = 1 (a 128 B line boundary); rows below i0 are skipped (ux: i0 = 2) */
  const int i = 1 + blockIdx.x * TI + ti;
  const int other = 1 + blockIdx.y;    /* k for y lines, j for z lines */
  const bool valid = i >= A.i0 && i < A.i0 + A.isz;
  const int n = A.dir == 1 ? g.ny : g.nz;
  const int m = n - 1;
  const int seg = (m + S - 1) / S;
  const int row0 = sg * seg;
  int cnt = m - row0;
  cnt = cnt < 0 ? 0 : (cnt > seg ? seg : cnt);
  const long long step = A.dir == 1 ? (long long)g.px : g.plane;
  const int ic = valid ? i : A.i0;
  const long long base = A.dir == 1 ? tck_idx(g, ic, 1, other) : tck_idx(g, ic, other, 1);
  const int col = A.tcol ? ic : 0;
  const long long ts = A.tstride;
  const double * r = rt + col;
  const double * v = vt + col;
  const double * q1 = q1t + col;
  const double den = meta[col];
  const double tl = meta[ts + col];
  const double tu = meta[2 * ts + col];
  double * smA = ps_sm;                  /* [S][TI] */
  double * smB = ps_sm + S * TI;         /* [S][TI] */
  double * smX = ps_sm + 2 * S * TI;     /* [3][TI]: q0[0], q0[n-2], q[n-1] */
  double q[PS_MAXM];
  #pragma unroll
  for(int e = 0; e < PS_MAXM; e++) q[e] = (valid && e < cnt) ? d[base + step * (row0 + e)] : 0.;
  if(valid && sg == S - 1) smX[2 * TI + ti] = d[base + step * m];
  /* forward: q0 = q0 / c0, q_i = r_i (q_i - l q_{i-1}) */
  double a = 1., b = 0.;
  #pragma unroll
  for(int e = 0; e < PS_MAXM; e++){
    if(e < cnt){
      const int row = row0 + e;
      if(row == 0){ a = 0.; b = q[e] / r[0]; }
      else {
        const double ri = r[ts * row];
        const double ai = - ri * tl;
        b = ai * b + ri * q[e];
        a = ai * a;
      }
    }
  }
  smA[sg * TI + ti] = a;
  smB[sg * TI + ti] = b;
  __syncthreads();
  double in = 0.;
  for(int s2 = 0; s2 < sg; s2++) in = smA[s2 * TI + ti] * in + smB[s2 * TI + ti];
  #pragma unroll
  for(int e = 0; e < PS_MAXM; e++){
    if(e < cnt){
      const int row = row0 + e;
      if(row == 0) q[e] = q[e] / r[0];
      else q[e] = r[ts * row] * (q[e] - tl * in);
      in = q[e];
      if(row == m - 1) smX[TI + ti] = q[e];
    }
  }
  __syncthreads();
  /* backward: q_i -= v_i q_{i+1}, q0[n-2] kept */
  a = 1.;
  b = 0.;
  #pragma unroll
  for(int e = PS_MAXM - 1; e >= 0; e--){
    if(e < cnt){
      const int row = row0 + e;
      if(row == m - 1){ a = 0.; b = q[e]; }
      else {
        const double ai = - v[ts * row];
        b = ai * b + q[e];
        a = ai * a;
      }
    }
  }
  smA[sg * TI + ti] = a;
  smB[sg * TI + ti] = b;
  __syncthreads();
  in = 0.;
  for(int s2 = S - 1; s2 > sg; s2--) in = smA[s2 * TI + ti] * in + smB[s2 * TI + ti];
  #pragma unroll
  for(int e = PS_MAXM - 1; e >= 0; e--){
    if(e < cnt){
      const int row = row0 + e;
      if(row != m - 1) q[e] -= v[ts * row] * in;
      in = q[e];
    }
  }
  if(cnt > 0 && row0 == 0) smX[ti] = q[0];
  __syncthreads();
  /* x_{n-1} (tdm.c:92-94) and x_i += x_{n-1} q1_i (:96-98) */
  const double num = smX[2 * TI + ti] - tu * smX[ti] - tl * smX[TI + ti];
  const double xl = fabs(den) < DBL_EPSILON ? 0. : num / den;
  if(!valid) return;
  if(ucur){
    double u[PS_MAXM];
    #pragma unroll
    for(int e = 0; e < PS_MAXM; e++) u[e] = e < cnt ? ucur[base + step * (row0 + e)] : 0.;
    const double ul = sg == S - 1 ? ucur[base + step * m] : 0.;
    #pragma unroll
    for(int e = 0; e < PS_MAXM; e++){
      if(e < cnt){
        const int row = row0 + e;
        const double x = q[e] + xl * q1[ts * row];
        const int jj = A.dir == 1 ? 1 + row : other;
        const int kk = A.dir == 1 ? other : 1 + row;
        tck_store_halo(d, g, i, jj, kk, u[e] + x, false);
      }
    }
    if(sg == S - 1){
      const int jj = A.dir == 1 ? n : other;
      const int kk = A.dir == 1 ? other : n;
      tck_store_halo(d, g, i, jj, kk, ul + xl, false);
    }
  }else{
    #pragma unroll
    for(int e = 0; e < PS_MAXM; e++){
      if(e < cnt){
        const int row = row0 + e;
        d[base + step * row] = q[e] + xl * q1[ts * row];
      }
    }
    if(sg == S - 1) d[base + step * m] = xl;
  }
}

__global__ void __launch_bounds__(1024) tck_psolve(const PSolveArgs A){
  extern __shared__ double ps_sm[];
  psolve_body(A, A.d, A.ucur, A.r, A.v, A.q1, A.meta, ps_sm);
}

/* ---------------------------------------------------------------------------------------
 * Poisson radial systems (compute_potential.c:445-488): per line (ky, kz)
 *   c_i = - l_i - u_i + ly(ky) / HYXC(i)^2 + lz(kz) / hz^2,  c_0 += l_0,  c_{n-1} += u_{n-1}
 * factors precomputed once into pr/pv (layout [line][i]); rows of the spectrum are complex.
 * ------------------------------------------------------------------------------------- */
struct PFactorArgs {
  int nx, nky, nz;
  const double * pl; const double * pu;     /* l_i, u_i (i = 0..nx-1) */
  const double * hyxc;                      /* HYXC(i), reference index */
  const double * evy; const double * evz;
  double hz;
  double * pr; double * pv;
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
  double * v = A.pv + line * n;
  const double hz = A.hz;
  /* the Neumann fix-ups are added after the sum, as c[0] += l[0] does (:481-482) */
  double c0 = - A.pl[0] - A.pu[0] + eval_y / A.hyxc[1] / A.hyxc[1] + eval_z / hz / hz;
  c0 += A.pl[0];
  double vprev = A.pu[0] / c0;
  r[0] = c0;
  v[0] = vprev;
  for(int i = 1; i < n - 1; i++){
    const double ci = - A.pl[i] - A.pu[i] + eval_y / A.hyxc[i + 1] / A.hyxc[i + 1] + eval_z / hz / hz;
    const double val = 1. / (ci - A.pl[i] * vprev);
    vprev = val * (A.pu[i]);
    r[i] = val;
    v[i] = vprev;
  }
  double cl = - A.pl[n - 1] - A.pu[n - 1] + eval_y / A.hyxc[n] / A.hyxc[n] + eval_z / hz / hz;
  cl += A.pu[n - 1];
  const double val = cl - A.pl[n - 1] * vprev;
  r[n - 1] = fabs(val) > DBL_EPSILON ? 1. / val : 0.;
  v[n - 1] = 0.;
}

struct PXSolveArgs {
  int nx, nlines;
  double2 * s;               /* spectrum, line-major rows of nx complex */
  const double * pr; const double * pv; const double * pl;
};

/* one (ky, kz) line per wavefront, staged through LDS as in tck_xsolve so that the spectrum
 * row (nx complex values, contiguous) is read and written with contiguous wave accesses */
__global__ void __launch_bounds__(TCK_BLOCK) tck_pxsolve(const PXSolveArgs A){
  extern __shared__ double2 px_sm[];
  const int w = threadIdx.x / TCK_WAVE;
  const long long wave = (long long)blockIdx.x * (blockDim.x / TCK_WAVE) + w;
  const int lane = threadIdx.x % TCK_WAVE;
  const bool live = wave < A.nlines;
  const int n = A.nx;
  const int seg = (n + TCK_WAVE - 1) / TCK_WAVE;
  const int ld = seg + 1;
  const int row0 = lane * seg;
  int cnt = n - row0;
  cnt = cnt < 0 ? 0 : (cnt > seg ? seg : cnt);
  const long long lw = live ? wave : 0;
  double2 * row = A.s + lw * n;
  const double * r = A.pr + lw * n;
  const double * v = A.pv + lw * n;
  double2 * sm = px_sm + (size_t)w * TCK_WAVE * ld;
  if(live){
    for(int idx = lane; idx < n; idx += TCK_WAVE){
      const int ln = idx / seg;
      sm[ln * ld + (idx - ln * seg)] = row[idx];
    }
  }
  __syncthreads();
  double q[XS_MAXSEG][2];
  #pragma unroll
  for(int e = 0; e < XS_MAXSEG; e++){
    if(e < cnt){
      const double2 x = sm[lane * ld + e];
      q[e][0] = x.x;
      q[e][1] = x.y;
    }else{
      q[e][0] = q[e][1] = 0.;
    }
  }
  wave_tdm<2>(q, r, v, A.pl, row0, cnt, n);
  #pragma unroll
  for(int e = 0; e < XS_MAXSEG; e++){
    if(e < cnt) sm[lane * ld + e] = make_double2(q[e][0], q[e][1]);
  }
  __syncthreads();
  if(!live) return;
  for(int idx = lane; idx < n; idx += TCK_WAVE){
    const int ln = idx / seg;
    row[idx] = sm[ln * ld + (idx - ln * seg)];
  }
}

/* ------------------------------------------------------------------------------------- */
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

int tck_launch_xsolve(const tc_geom_t * g, double * d, const double * ucur, const double * r,
    const double * v, const double * l, int i0, int n, hipStream_t stream){
  if(n > TCK_WAVE * XS_MAXSEG || n < 2) return (int)hipErrorInvalidValue;
  XSolveArgs A;
  A.g = *g; A.d = d; A.ucur = ucur; A.r = r; A.v = v; A.l = l; A.i0 = i0; A.n = n;
  const long long nlines = (long long)g->ny * g->nz;
  const int waves_per_block = TCK_BLOCK / TCK_WAVE;
  const long long nblocks = (nlines + waves_per_block - 1) / waves_per_block;
  hipLaunchKernelGGL(tck_xsolve, dim3((unsigned)nblocks), dim3(TCK_BLOCK), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_psolve(const tc_geom_t * g, int dir, double * d, const double * ucur,
    const double * r, const double * v, const double * q1, const double * meta,
    long long tstride, int tcol, int i0, int isz, hipStream_t stream){
  PSolveArgs A;
  A.g = *g; A.d = d; A.ucur = ucur; A.r = r; A.v = v; A.q1 = q1; A.meta = meta;
  A.tstride = tstride; A.tcol = tcol; A.dir = dir; A.i0 = i0; A.isz = isz;
  const int n = dir == 1 ? g->ny : g->nz;
  const int m = n - 1;
  const int S = (m + PS_MAXM - 1) / PS_MAXM;
  if(S > 64) return (int)hipErrorInvalidValue;
  int TI = 512 / S;
  TI = TI > 64 ? 64 : (TI < 16 ? 16 : TI);
  const int nother = dir == 1 ? g->nz : g->ny;
  const size_t lds = sizeof(double) * (size_t)(2 * S * TI + 3 * TI);
  const int span = i0 + isz - 1;   /* tiles cover i = 1 .. i0+isz-1 */
  hipLaunchKernelGGL(tck_psolve, dim3((span + TI - 1) / TI, nother), dim3(TI, S), lds, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_pfactor(int nx, int nky, int nz, const double * pl, const double * pu,
    const double * hyxc, const double * evy, const double * evz, double hz, double * pr,
    double * pv, hipStream_t stream){
  PFactorArgs A;
  A.nx = nx; A.nky = nky; A.nz = nz; A.pl = pl; A.pu = pu; A.hyxc = hyxc; A.evy = evy;
  A.evz = evz; A.hz = hz; A.pr = pr; A.pv = pv;
  const long long nlines = (long long)nky * nz;
  hipLaunchKernelGGL(tck_pfactor, dim3((unsigned)((nlines + 63) / 64)), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_pxsolve(int nx, int nlines, double * s, const double * pr, const double * pv,
    const double * pl, hipStream_t stream){
  if(nx > TCK_WAVE * XS_MAXSEG || nx < 2) return (int)hipErrorInvalidValue;
  PXSolveArgs A;
  A.nx = nx; A.nlines = nlines; A.s = (double2 *)s; A.pr = pr; A.pv = pv; A.pl = pl;
  const int seg = (nx + TCK_WAVE - 1) / TCK_WAVE;
  const int waves_per_block = TCK_BLOCK / TCK_WAVE;
  const size_t lds = sizeof(double2) * (size_t)waves_per_block * TCK_WAVE * (seg + 1);
  const long long nblocks = ((long long)nlines + waves_per_block - 1) / waves_per_block;
  if(lds > 65536 && hipFuncSetAttribute((const void *)tck_pxsolve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(tck_pxsolve, dim3((unsigned)nblocks), dim3(TCK_BLOCK), lds, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

}
