/*
 * tck_diag.hip -- boundary fill and FP64 diagnostics (gfx950).
 *
 *  tck_bc_fill     walls and periodic halos of one field (src/fluid/boundary/*.c with
 *                  src/halo.c), used after fields are replaced from the host.
 *  tck_diag_cells  per-block partial sums of the volume diagnostics:
 *                   divergence maximum   logging/divergence.c:32-91
 *                   dissipation          logging/dissipation.c:27-371 (9 tensor products
 *                                        and 3 scalar-gradient terms, tensors formed on the
 *                                        fly from the velocity as predict/l*.c do)
 *                   total energy         logging/total_energy.c:22-93
 *  tck_diag_walls  injection through the walls, logging/injection.c:31-103
 *  tck_diag_finish sums the partials in a fixed order (run-to-run reproducible).
 */
#include "tck_common.h"

/* kind: 0 ux, 1 uy, 2 uz, 3 t, 4 Neumann (p, psi) */
struct BcArgs {
  tc_geom_t g;
  double * a;
  int kind;
  double wm, wp;   /* Dirichlet wall values (uy, t) */
};

__global__ void tck_bc_fill(const BcArgs A){
  const tc_geom_t & g = A.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if(i > g.nx + 1) return;
  const int js = j == 0 ? g.ny : (j == g.ny + 1 ? 1 : j);
  const int ks = k == 0 ? g.nz : (k == g.nz + 1 ? 1 : k);
  double v;
  if(A.kind == 0){
    if(i == 0) return;
    if(i == 1 || i == g.nx + 1) v = 0.;
    else v = A.a[tck_idx(g, i, js, ks)];
  }else if(A.kind == 4){
    const int is = i == 0 ? 1 : (i == g.nx + 1 ? g.nx : i);
    v = A.a[tck_idx(g, is, js, ks)];
  }else{
    if(i == 0) v = A.wm;
    else if(i == g.nx + 1) v = A.wp;
    else v = A.a[tck_idx(g, i, js, ks)];
  }
  const bool halo = (j != js) || (k != ks);
  const bool wall = (i == 0 || i == g.nx + 1 || (A.kind == 0 && i == 1));
  if(halo || wall) A.a[tck_idx(g, i, j, k)] = v;
}

/* interior pass before the halo pass: walls on interior planes first so that the halo
 * images of the wall slots copy the final wall values (bc then halo, as the reference) */
__global__ void tck_bc_walls(const BcArgs A){
  const tc_geom_t & g = A.g;
  const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.y;
  if(j > g.ny) return;
  if(A.kind == 0){
    A.a[tck_idx(g, 1, j, k)] = 0.;
    A.a[tck_idx(g, g.nx + 1, j, k)] = 0.;
  }else if(A.kind == 4){
    A.a[tck_idx(g, 0, j, k)] = A.a[tck_idx(g, 1, j, k)];
    A.a[tck_idx(g, g.nx + 1, j, k)] = A.a[tck_idx(g, g.nx, j, k)];
  }else{
    A.a[tck_idx(g, 0, j, k)] = A.wm;
    A.a[tck_idx(g, g.nx + 1, j, k)] = A.wp;
  }
}

/* ------------------------------------------------------------------------------------- */
#define NQ_CELL 8   /* divmax, dissipation u, dissipation T, E_ux, E_uy, E_uz, E_t, spare */

struct DiagArgs {
  tc_geom_t g;
  tc_coef_t c;
  const double * ux; const double * uy; const double * uz; const double * t;
  double * part;        /* [nblocks][NQ_CELL] */
  int scalar;
};

__global__ void __launch_bounds__(TCK_BLOCK) tck_diag_cells(const DiagArgs A){
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;   /* faces 1..nx+1, centres 1..nx */
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double q[NQ_CELL] = {0., 0., 0., 0., 0., 0., 0., 0.};
  if(i <= g.nx + 1 && j <= g.ny){
    const long long o = tck_idx(g, i, j, k);
    const long long sy = g.px;
    const long long sz = g.plane;
    const double * __restrict__ ux = A.ux;
    const double * __restrict__ uy = A.uy;
    const double * __restrict__ uz = A.uz;
    const double hz = c.hz;
    /* face i quantities */
    {
      const double lyx0 = 1. / c.hyxf[i] * (- ux[o - sy] + ux[o]);
      const double lyx1 = c.l1f[i] * (+ (1 == i ? 0. : 0.5) * uy[o - 1] + (g.nx + 1 == i ? 0. : 0.5) * uy[o]);
      const double lxy = 1. / c.hxxf[i] * (- uy[o - 1] + uy[o]);
      const double lzx = 1. / hz * (- ux[o - sz] + ux[o]);
      const double lxz = 1. / c.hxxf[i] * (- uz[o - 1] + uz[o]);
      const double jd = c.jdxf[i];
      double v = 0.;
      {
        const double lij = lyx0 + lyx1;
        v += jd * lij * (lij + lxy);
      }
      v += jd * lzx * (lzx + lxz);
      {
        const double lji = lyx0 + lyx1;
        v += jd * lxy * (lxy + lji);
      }
      v += jd * lxz * (lxz + lzx);
      q[1] += v;
      if(A.scalar){
        const double d = - A.t[o - 1] + A.t[o];
        const double r = d / c.hxxf[i];
        q[2] += jd * (r * r);
      }
    }
    if(i <= g.nx){
      /* centre i quantities */
      const double jd = c.jdxc[i];
      const double lxx = 1. / c.hxxc[i] * (- ux[o] + ux[o + 1]);
      const double lyy0 = 1. / c.hyxc[i] * (- uy[o] + uy[o + sy]);
      const double lyy1 = c.l1c[i] * (+ 0.5 * ux[o] + 0.5 * ux[o + 1]);
      const double lzy = 1. / hz * (- uy[o - sz] + uy[o]);
      const double lyz = 1. / c.hyxc[i] * (- uz[o - sy] + uz[o]);
      const double lzz = 1. / hz * (- uz[o] + uz[o + sz]);
      double v = 0.;
      v += jd * lxx * (lxx + lxx);
      {
        const double lij = lyy0 + lyy1;
        v += jd * lij * (lij + lij);
      }
      v += jd * lzy * (lzy + lyz);
      v += jd * lyz * (lyz + lzy);
      v += jd * lzz * (lzz + lzz);
      q[1] += v;
      if(A.scalar){
        const double * __restrict__ t = A.t;
        const double dy = - t[o - sy] + t[o];
        const double ry = dy / c.hyxc[i];
        const double dz = - t[o - sz] + t[o];
        const double rz = dz / hz;
        q[2] += jd * (ry * ry);
        q[2] += jd * (rz * rz);
        q[6] += jd * 0.5 * (t[o] * t[o]);
      }
      /* divergence, logging/divergence.c:56-76 */
      const double div = c.ijdxc[i] * (
          - c.jhxf[i] * ux[o] + c.jhxf[i + 1] * ux[o + 1]
          - c.jhyc[i] * uy[o] + c.jhyc[i] * uy[o + sy]
          - c.jhzc[i] * uz[o] + c.jhzc[i] * uz[o + sz]);
      q[0] = fabs(div);
      /* energies, total_energy.c:45-75 */
      if(i >= 2) q[3] += c.jdxf[i] * 0.5 * (ux[o] * ux[o]);
      q[4] += jd * 0.5 * (uy[o] * uy[o]);
      q[5] += jd * 0.5 * (uz[o] * uz[o]);
    }
  }
  /* block reduction: max for q[0], sums for the rest */
  __shared__ double red[TCK_BLOCK / TCK_WAVE][NQ_CELL];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int w = tid / TCK_WAVE;
  const int nw = (blockDim.x * blockDim.y + TCK_WAVE - 1) / TCK_WAVE;
  double r0 = tck_wave_max(q[0]);
  if((tid % TCK_WAVE) == 0) red[w][0] = r0;
  #pragma unroll
  for(int n = 1; n < NQ_CELL; n++){
    const double s = tck_wave_sum(q[n]);
    if((tid % TCK_WAVE) == 0) red[w][n] = s;
  }
  __syncthreads();
  if(tid == 0){
    const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    double * out = A.part + b * NQ_CELL;
    double m = red[0][0];
    for(int ww = 1; ww < nw; ww++) m = fmax(m, red[ww][0]);
    out[0] = m;
    for(int n = 1; n < NQ_CELL; n++){
      double s = 0.;
      for(int ww = 0; ww < nw; ww++) s += red[ww][n];
      out[n] = s;
    }
  }
}

/* injection: per (j, k) wall terms at i = 1 and nx+1 */
struct WallArgs {
  tc_geom_t g;
  tc_coef_t c;
  const double * ux; const double * uy; const double * t;
  double * part;   /* [nblocks][2] */
  int scalar;
};

__global__ void __launch_bounds__(TCK_BLOCK) tck_diag_walls(const WallArgs A){
  const tc_geom_t & g = A.g;
  const tc_coef_t & c = A.c;
  const int j = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 1 + blockIdx.y;
  double q0 = 0., q1 = 0.;
  if(j <= g.ny){
    const int nx = g.nx;
    const long long sy = g.px;
    const long long om = tck_idx(g, 1, j, k);
    const long long op = tck_idx(g, nx + 1, j, k);
    const double * __restrict__ ux = A.ux;
    const double * __restrict__ uy = A.uy;
    const double hx_xm = c.hxxf[1], hx_xp = c.hxxf[nx + 1];
    const double jd_xm = c.jdxf[1], jd_xp = c.jdxf[nx + 1];
    const double lyx0_xm = 1. / c.hyxf[1] * (- ux[om - sy] + ux[om]);
    const double lyx0_xp = 1. / c.hyxf[nx + 1] * (- ux[op - sy] + ux[op]);
    const double lyx1_xm = c.l1f[1] * (+ 0. * uy[om - 1] + 0.5 * uy[om]);
    const double lyx1_xp = c.l1f[nx + 1] * (+ 0.5 * uy[op - 1] + 0. * uy[op]);
    const double lxy_xm = 1. / c.hxxf[1] * (- uy[om - 1] + uy[om]);
    const double lxy_xp = 1. / c.hxxf[nx + 1] * (- uy[op - 1] + uy[op]);
    const double tyx_xm = lyx0_xm + lyx1_xm + lxy_xm;
    const double tyx_xp = lyx0_xp + lyx1_xp + lxy_xp;
    q0 = - jd_xm / hx_xm * uy[om - 1] * tyx_xm + jd_xp / hx_xp * uy[op] * tyx_xp;
    if(A.scalar){
      const double * __restrict__ t = A.t;
      const double t_xm = t[om - 1], t_xp = t[op];
      const double dt_xm = - t[om - 1] + t[om];
      const double dt_xp = - t[op - 1] + t[op];
      q1 = - jd_xm / hx_xm * t_xm / hx_xm * dt_xm + jd_xp / hx_xp * t_xp / hx_xp * dt_xp;
    }
  }
  __shared__ double red[TCK_BLOCK / TCK_WAVE][2];
  const int tid = threadIdx.x;
  const int w = tid / TCK_WAVE;
  const int nw = (blockDim.x + TCK_WAVE - 1) / TCK_WAVE;
  const double s0 = tck_wave_sum(q0), s1 = tck_wave_sum(q1);
  if((tid % TCK_WAVE) == 0){ red[w][0] = s0; red[w][1] = s1; }
  __syncthreads();
  if(tid == 0){
    double a = 0., b = 0.;
    for(int ww = 0; ww < nw; ww++){ a += red[ww][0]; b += red[ww][1]; }
    const long long bid = blockIdx.x + (long long)gridDim.x * blockIdx.y;
    A.part[2 * bid] = a;
    A.part[2 * bid + 1] = b;
  }
}

/* fixed-order reduction of nblocks partial rows of width nq (column 0 a max when
 * first_is_max): one block of 256 threads, each thread a strided share, then a tree in LDS;
 * the order depends only on nblocks, so repeated calls give identical results */
__global__ void __launch_bounds__(256) tck_diag_finish(const double * part, long long nblocks, int nq, int first_is_max, double * out){
  __shared__ double red[256];
  for(int n = 0; n < nq; n++){
    const bool mx = first_is_max && n == 0;
    double s = 0.;
    for(long long b = threadIdx.x; b < nblocks; b += blockDim.x){
      const double v = part[b * nq + n];
      s = mx ? fmax(s, v) : s + v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for(int h = 128; h > 0; h >>= 1){
      if((int)threadIdx.x < h) red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + h]) : red[threadIdx.x] + red[threadIdx.x + h];
      __syncthreads();
    }
    if(threadIdx.x == 0) out[n] = red[0];
    __syncthreads();
  }
}

/* ------------------------------------------------------------------------------------- */
/* running statistics, src/statistics.c:111-234.  The accumulators use the reference's
 * compact layouts: ux1, ux2, adv, dif over i = 1..nx+1 and uy1, uy2, uz1, uz2 over
 * i = 0..nx+1, then j = 1..ny, k = 1..nz, i fastest. */
struct StatArgs {
  tc_geom_t g;
  tc_coef_t c;
  const double * ux; const double * uy; const double * uz;
  double * ux1; double * ux2; double * uy1; double * uy2; double * uz1; double * uz2;
  double * adv; double * dif;
  const double * difs;   /* LYX0 + LYX1 + LXY of the tensors the last substep formed */
};

/* difs(i, j, k) = LYX0 + LYX1 + LXY (predict/lyx.c:24-95, predict/lxy.c:19-49) from the
 * velocity of the start of the last substep; i = 1..nx+1 */
__global__ void __launch_bounds__(TCK_BLOCK) tck_stat_tensors(const StatArgs A){
  const tc_geom_t & g = A.g;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if(i > g.nx + 1 || j > g.ny) return;
  const long long o = tck_idx(g, i, j, k);
  const double * __restrict__ ux = A.ux;
  const double * __restrict__ uy = A.uy;
  const double lyx0 = 1. / A.c.hyxf[i] * (- ux[o - g.px] + ux[o]);
  const double lyx1 = A.c.l1f[i] * (
      + (1 == i ? 0. : 0.5) * uy[o - 1]
      + (g.nx + 1 == i ? 0. : 0.5) * uy[o]);
  const double lxy = 1. / A.c.hxxf[i] * (- uy[o - 1] + uy[o]);
  const long long q = (i - 1) + (long long)(g.nx + 1) * ((j - 1) + (long long)g.ny * (k - 1));
  ((double *)A.difs)[q] = lyx0 + lyx1 + lxy;
}

__global__ void __launch_bounds__(TCK_BLOCK) tck_stat_collect(const StatArgs A){
  const tc_geom_t & g = A.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;    /* 0 .. nx+1 */
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if(i > g.nx + 1 || j > g.ny) return;
  const long long o = tck_idx(g, i, j, k);
  const long long jk = (j - 1) + (long long)g.ny * (k - 1);
  const long long qc = i + (long long)(g.nx + 2) * jk;
  const double vy = A.uy[o], vz = A.uz[o];
  A.uy1[qc] += vy;
  A.uy2[qc] += vy * vy;
  A.uz1[qc] += vz;
  A.uz2[qc] += vz * vz;
  if(i >= 1){
    const long long qf = (i - 1) + (long long)(g.nx + 1) * jk;
    const double vx = A.ux[o];
    A.ux1[qf] += vx;
    A.ux2[qf] += vx * vx;
    const double x = 0.5 * A.ux[o - g.px] + 0.5 * vx;
    const double y = 0.5 * A.uy[o - 1] + 0.5 * vy;
    A.adv[qf] += x * y;
    A.dif[qf] += A.difs[qf];
  }
}

/* ------------------------------------------------------------------------------------- */
extern "C" {

int tck_launch_bc(const tc_geom_t * g, double * a, int kind, double wm, double wp, hipStream_t stream){
  BcArgs A;
  A.g = *g; A.a = a; A.kind = kind; A.wm = wm; A.wp = wp;
  hipLaunchKernelGGL(tck_bc_walls, dim3((g->ny + 63) / 64, g->nz), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  hipLaunchKernelGGL(tck_bc_fill, dim3((g->nx + 2 + 63) / 64, g->ny + 2, g->nz + 2), dim3(64), 0, stream, A);
  TCK_CHECK_LAUNCH();
  return 0;
}

long long tck_diag_cells_blocks(const tc_geom_t * g){
  return (long long)((g->nx + 1 + 63) / 64) * ((g->ny + 3) / 4) * g->nz;
}

long long tck_diag_walls_blocks(const tc_geom_t * g){
  return (long long)((g->ny + TCK_BLOCK - 1) / TCK_BLOCK) * g->nz;
}

int tck_launch_diag_cells(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * uz, const double * t, int scalar, double * part,
    double * out, hipStream_t stream){
  DiagArgs A;
  A.g = *g; A.c = *c; A.ux = ux; A.uy = uy; A.uz = uz; A.t = t; A.part = part; A.scalar = scalar;
  hipLaunchKernelGGL(tck_diag_cells, dim3((g->nx + 1 + 63) / 64, (g->ny + 3) / 4, g->nz), dim3(64, 4), 0, stream, A);
  TCK_CHECK_LAUNCH();
  hipLaunchKernelGGL(tck_diag_finish, dim3(1), dim3(256), 0, stream, part, tck_diag_cells_blocks(g), NQ_CELL, 1, out);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_diag_walls(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * t, int scalar, double * part, double * out, hipStream_t stream){
  WallArgs A;
  A.g = *g; A.c = *c; A.ux = ux; A.uy = uy; A.t = t; A.part = part; A.scalar = scalar;
  hipLaunchKernelGGL(tck_diag_walls, dim3((g->ny + TCK_BLOCK - 1) / TCK_BLOCK, g->nz), dim3(TCK_BLOCK), 0, stream, A);
  TCK_CHECK_LAUNCH();
  hipLaunchKernelGGL(tck_diag_finish, dim3(1), dim3(256), 0, stream, part, tck_diag_walls_blocks(g), 2, 0, out);
  TCK_CHECK_LAUNCH();
  return 0;
}

int tck_launch_stat(const tc_geom_t * g, const tc_coef_t * c, int what, const double * ux,
    const double * uy, const double * uz, double * const acc[9], hipStream_t stream){
  StatArgs A;
  A.g = *g; A.c = *c; A.ux = ux; A.uy = uy; A.uz = uz;
  A.ux1 = acc[0]; A.ux2 = acc[1]; A.uy1 = acc[2]; A.uy2 = acc[3];
  A.uz1 = acc[4]; A.uz2 = acc[5]; A.adv = acc[6]; A.dif = acc[7]; A.difs = acc[8];
  if(what == 0){
    hipLaunchKernelGGL(tck_stat_tensors, dim3((g->nx + 1 + 63) / 64, (g->ny + 3) / 4, g->nz), dim3(64, 4), 0, stream, A);
  }else{
    hipLaunchKernelGGL(tck_stat_collect, dim3((g->nx + 2 + 63) / 64, (g->ny + 3) / 4, g->nz), dim3(64, 4), 0, stream, A);
  }
  TCK_CHECK_LAUNCH();
  return 0;
}

}
