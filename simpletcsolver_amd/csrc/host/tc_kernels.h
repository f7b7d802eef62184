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
/* mixed: the temporaries (increment dlt, srcb, srca, psi) are float and the right-hand sides
 * are evaluated in FP32; otherwise they are double and dlt may be nxt itself.  window: the
 * variant that stages the fields' planes in LDS (tck_predict_w) */
int tck_launch_predict(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    const double * const cur[4], const double * p, double * const nxt[4], void * const dlt[4],
    const void * const srcb[4], void * const srca[4], const float * const corr[4],
    const double * dt, int rkstep, int final_here, int mixed, int kb, int ty, int window, hipStream_t stream);
int tck_launch_correct(const tc_geom_t * g, const tc_coef_t * c, const tc_flags_t * f,
    double * ux, double * uy, double * uz, double * p, const void * psi, const double * dt,
    int rkstep, int mode, unsigned long long * dtmin, int mixed, int kb, int ty, hipStream_t stream);
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
/* the x solves of nf (1..4) fields in one launch; per field: d the increment (float when
 * mixed), uout where the final pass (ucur non-null, for every field or for none) writes
 * ucur + solution (FP64; may be d itself when not mixed), the factor rows r, v, l and the
 * line i0 .. i0+n-1 */
int tck_launch_xsolve(const tc_geom_t * g, int nf, void * const * d, double * const * uout,
    const double * const * ucur, const double * const * r, const double * const * v,
    const double * const * l, const int * i0, const int * n, int mixed, hipStream_t stream);
/* the periodic solves (dir 1: y, 2: z) of nf (1..4) fields in one launch, with per-field
 * arrays as tck_launch_xsolve's plus the correction vector q1 and the meta rows; lean bit 1
 * lets waves of interior full-length segments skip the row tests, bit 2 batches the folds' LDS
 * reads (same results) */
int tck_launch_psolve(const tc_geom_t * g, int dir, int nf, void * const * d, double * const * uout,
    const double * const * ucur, const double * const * r, const double * const * v,
    const double * const * q1, const double * const * meta, long long tstride, int tcol,
    const int * i0, const int * isz, int mixed, int lean, hipStream_t stream);
int tck_launch_pfactor(int nx, int nky, int nz, const double * pl, const double * pu,
    const double * hyxc, const double * evy, const double * evz, double hz, double * pr,
    int exact_gauge, hipStream_t stream);
int tck_launch_pxsolve(int nx, int nlines, void * s, const double * pr, const double * pl,
    const double * pu, int mixed, hipStream_t stream);

/* tck_fft.hip */
int tck_fft_supported(int n);
/* mixed: spectrum, transforms, twiddles and psi in FP32 */
int tck_launch_div_yfft(const tc_geom_t * g, const tc_coef_t * c, const double * ux,
    const double * uy, const double * uz, const double * dt, int rkstep, void * S,
    const void * twM, const void * twN, int mixed, int twl, hipStream_t stream);
int tck_launch_zfft(int nx, int nky, int nz, void * S, const void * tw, int sign, int mixed, int exact, hipStream_t stream);
int tck_launch_yifft_psi(const tc_geom_t * g, const void * S, void * psi, const void * twM,
    const void * twN, int mixed, int exact, hipStream_t stream);

/* tck_util.hip: device address of this device's staging array of the allocation-free
 * initial-condition fill (4 tables of TC_IC_NOALLOC_N doubles), looked up once per device */
int tck_ic_scratch(double ** ptr);

/* tck_diag.hip */
int tck_launch_copy(const double * src, double * dst, long long n, int chunks, hipStream_t stream);
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
