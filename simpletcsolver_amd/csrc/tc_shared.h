/*
 * tc_shared.h -- plain-C structs shared by the C host code and the HIP kernels.
 *
 * Device layout of every 3-D array (velocities, scalar, pressure, potential, RK buffers):
 *   idx(i, j, k) = (i + xoff) + px * (j + py * k),  i = 0..nx+1, j = 0..ny+1, k = 0..nz+1
 * with the reference's own indices (src/fluid/*, include/array_macros/fluid/*.h): x walls at
 * i = 0 and nx+1 (ux: faces i = 1..nx+1, slot 0 unused), one periodic halo plane on each
 * side in y and z.  xoff = 15 and px a multiple of 16 doubles put i = 1 on a 128 B cache-line
 * boundary in every row, so tiles of 16, 32 or 64 interior cells read and write whole lines.
 * All arrays share one index map so a kernel addresses ux, uy, p and the RK buffers alike.
 */
#ifndef TC_SHARED_H
#define TC_SHARED_H

#ifdef __cplusplus
extern "C" {
#endif

#define TC_XOFF 15

/* largest table length of the allocation-free initial-condition fill: nx + 2, ny and nz
 * may not exceed it (tc_ic_fields_noalloc) */
#define TC_IC_NOALLOC_N 4096

typedef struct {
  int nx, ny, nz;
  int px, py, pz;
  long long plane;      /* px * py */
  long long n3;         /* plane * pz */
} tc_geom_t;

/* per-i coefficient tables, packed in one array of T_NTAB rows of tp = nx + 2 entries
 * (row t at tab + t * tp); the rows used by the predictor come first (T_NPRED of them) so a
 * kernel can stage them in LDS as one block */
enum {
  T_HXXF, T_HXXC, T_HYXF, T_HYXC, T_JDXF, T_JDXC,
  T_AXF, T_IJDXF, T_IHXF, T_IHYF, T_JHXF, T_JHYF, T_JHZF, T_L1F, T_NUJF, T_N2JF, T_D1M, T_D1P,
  T_AYC, T_AZC, T_IJDXC, T_IHYC, T_JHYC, T_JHZC, T_L1C, T_HDC, T_NUJC, T_N2JC, T_DAM, T_DAP,
  T_LXU_L, T_LXU_C, T_LXU_U, T_LYU_L, T_LYU_C, T_LYU_U,
  T_LXC_L, T_LXC_C, T_LXC_U, T_LYC_L, T_LYC_C, T_LYC_U,
  T_PLX_L, T_PLX_U, T_PLY, T_PLZ,
  T_NTAB
};
#define T_NPRED T_PLX_L

/* per-i coefficient tables (device pointers).  Every table has nx+2 entries and is indexed
 * by the reference's i: face quantities at i = 1..nx+1, centre quantities at i = 0..nx+1.
 * Each entry is evaluated on the host with the operand order the reference uses inline, so
 * a kernel that multiplies by it rounds exactly as the reference's loop body does. */
typedef struct {
  const double * hxxf; const double * hxxc; const double * hyxf; const double * hyxc;
  const double * jdxf; const double * jdxc;
  /* face based */
  const double * axf;    /* 0.5 * JDXF / HXXF          advection of ux, uy, uz in x   */
  const double * ijdxf;  /* 1. / JDXF                                                 */
  const double * ihxf;   /* 1. / HXXF                  lxy, lxz, pressure of ux        */
  const double * ihyf;   /* 1. / HYXF                  lyx0                            */
  const double * jhxf;   /* JDXF / HXXF                                                */
  const double * jhyf;   /* JDXF / HYXF                                                */
  const double * jhzf;   /* JDXF / hz                                                  */
  const double * l1f;    /* -1. / JDXF * d(J/hx)_c     lyx1                            */
  const double * nujf;   /* nu / JDXF                                                  */
  const double * n2jf;   /* 2 nu / JDXF                                                */
  const double * d1m;    /* 0.5 * d(J/hx)_f at i-1/2   ux advection_1, diffusion_a     */
  const double * d1p;    /* 0.5 * d(J/hx)_f at i+1/2                                   */
  /* centre based */
  const double * ayc;    /* 0.5 * JDXC / HYXC                                          */
  const double * azc;    /* 0.5 * JDXC / hz                                            */
  const double * ijdxc;  /* 1. / JDXC                                                  */
  const double * ihyc;   /* 1. / HYXC                                                  */
  const double * jhyc;   /* JDXC / HYXC                                                */
  const double * jhzc;   /* JDXC / hz                                                  */
  const double * l1c;    /* 1. / JDXC * d(J/hx)_f      lyy1                            */
  const double * hdc;    /* 0.5 * d(J/hx)_f            uy advection_1                  */
  const double * nujc;   /* nu / JDXC                                                  */
  const double * n2jc;   /* 2 nu / JDXC                                                */
  const double * dam;    /* 0.5 * d(J/hx)_c at i-1/2   uy diffusion_a                  */
  const double * dap;    /* 0.5 * d(J/hx)_c at i+1/2                                   */
  /* second-derivative stencils {l, c, u} */
  const double * lxu_l; const double * lxu_c; const double * lxu_u;   /* ux, x (i = 2..nx) */
  const double * lyu_l; const double * lyu_c; const double * lyu_u;   /* ux, y             */
  const double * lxc_l; const double * lxc_c; const double * lxc_u;   /* uy uz t, x        */
  const double * lyc_l; const double * lyc_c; const double * lyc_u;   /* uy uz t, y        */
  /* implicit-diffusion correction of the pressure (update_pressure.c:39-136) */
  const double * plx_l; const double * plx_u; const double * ply; const double * plz;
  double lapz_l, lapz_c, lapz_u;
  double hz, ihz;
  double nu, kappa;
  const double * tab;    /* all rows, T_NTAB x tp */
  int tp;
} tc_coef_t;

/* boundary values and switches (src/param/*.c, overridable at run time) */
typedef struct {
  int implicit[3];
  int scalar;
  double uy_xm, uy_xp, t_xm, t_xp;
} tc_flags_t;

#define TC_NFIELD 4   /* ux uy uz t */

#ifdef __cplusplus
}
#endif
#endif
