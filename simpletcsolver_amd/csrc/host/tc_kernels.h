/*
 * tc_kernels.h -- C prototypes of the HIP kernel launchers (defined in ../kernels/*.hip).
 * Every launcher returns 0 or a hipError_t code.
 */
#ifndef TC_KERNELS_H
#define TC_KERNELS_H

#include <hip/hip_runtime_api.h>
#include "../tc_shared.h"

/* tck_step.hip.  kb, ty: launch shape of the cell kernels (planes marched per block, j rows
 * per block of 64 x ty threads; ty in {1, 2, 4}).  dtmin: TC_DTSLOTS accumulators, one
 * every TC_DTSTRIDE entries */
#define TC_DTSLOTS 64
#define TC_DTSTRIDE 16
int tck_launch_predict(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    const double * const cur[4], const double * p, double * const nxt[4],
    const double * const srcb[4], double * const srca[4], const double * dt, int rkstep,
    int final_here, int kb, int ty, hipStream_t stream);
int tck_launch_correct(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    double * ux, double * uy, double * uz, double * p, const double * psi, const double * dt,
    int rkstep, int mode, unsigned long long * dtmin, int kb, int ty, hipStream_t stream);
int tck_launch_dt(const tc_geom_t * g, const tc_coef_t * c, const double * ux, const double * uy,
    const double * uz, unsigned long long * dtmin, hipStream_t stream);
int tck_launch_dt_finalise(unsigned long long * dtmin, double * dt, double coef_adv,
    double dif_limit, double dt_fixed, hipStream_t stream);

/* tck_linesolve.hip */
int tck_launch_factor(const double * dt, const int dirs[3], int nx, int ny, int nz, int tp,
    double lzl, double lzc, double lzu, int njobs, const int * rkstep, const double * diff,
    const double (*fac)[3], const int * i0, const int * n,
    const double * const (*lx)[3], const double * const (*ly)[3],
    double * const (*xt)[3], double * const (*yt)[4], double * const (*zt)[4], hipStream_t stream);
int tck_launch_xsolve(const tc_geom_t * g, double * d, const double * ucur, const double * r,
    const double * v, const double * l, int i0, int n, hipStream_t stream);
int tck_launch_psolve(const tc_geom_t * g, int dir, double * d, const double * ucur,
    const double * r, const double * v, const double * q1, const double * meta,
    long long tstride, int tcol, int i0, int isz, hipStream_t stream);
int tck_launch_pfactor(int nx, int nky, int nz, const double * pl, const double * pu,
    const double * hyxc, const double * evy, const double * evz, double hz, double * pr,
    double * pv, hipStream_t stream);
int tck_launch_pxsolve(int nx, int nlines, double * s, const double * pr, const double * pv,
    const double * pl, hipStream_t stream);

/* tck_fft.hip */
int tck_fft_supported(int n);
int tck_launch_div_yfft(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * uz, const double * dt, int rkstep, double * S,
    const double * twM, const double * twN, hipStream_t stream);
int tck_launch_zfft(int nx, int nky, int nz, double * S, const double * tw, int sign, hipStream_t stream);
int tck_launch_yifft_psi(const tc_geom_t * g, const double * S, double * psi, const double * twM,
    const double * twN, hipStream_t stream);

/* tck_diag.hip */
int tck_launch_copy(const double * src, double * dst, long long n, hipStream_t stream);
int tck_launch_bc(const tc_geom_t * g, double * a, int kind, double wm, double wp, hipStream_t stream);
long long tck_diag_cells_blocks(const tc_geom_t * g);
long long tck_diag_walls_blocks(const tc_geom_t * g);
int tck_launch_diag_cells(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * uz, const double * t, int scalar, double * part,
    double * out, hipStream_t stream);
int tck_launch_diag_walls(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * t, int scalar, double * part, double * out, hipStream_t stream);
/* what 0: acc[8] = LYX0 + LYX1 + LXY of (ux, uy); what 1: add the samples to acc[0..7] */
int tck_launch_stat(const tc_geom_t * g, const tc_coef_t * c, int what, const double * ux,
    const double * uy, const double * uz, double * const acc[9], hipStream_t stream);

#endif
