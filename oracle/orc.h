/*
 * orc.h -- CPU restatement ("oracle") of NaokiHori/SimpleTCSolver's RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or calls this
 * code.  It is used by tests/ (as the checker), by __graft_entry__.smoke() (as the checker)
 * and by bench.py's cpu_baseline leg (timed beside the GPU, never as the product).
 *
 * It follows the reference loop by loop and term by term, in the same accumulation order,
 * on the reference's own host array layouts (include/array_macros/fluid/*.h): every field
 * keeps its y/z halo planes and x wall slots exactly as the reference stores them, so the
 * arrays can be compared one to one with what the reference would hold.  Each function in
 * orc.c cites the reference file:line it restates.  One process, no MPI: the pencil
 * transposes of SimpleDecomp are pure permutations and are replaced by strided line access;
 * the periodic halo exchange (src/halo.c) becomes a local plane copy; FFTW is replaced by a
 * self-written mixed-radix FFT (orc_fft_*), checked against numpy.fft in the tests.
 *
 * Runtime options that the reference fixes at compile time (src/param/*.c) are fields of
 * orc_config_t; their reference defaults are implicit = {1,0,0}, uy walls = {1,0},
 * t walls = {+0.5,-0.5}, scalar advanced.
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;          /* radial, azimuthal, axial cell counts */
  int is_curved;           /* 1: Taylor-Couette, 0: planar channel */
  double lengths[3];       /* lx, ly (radians when curved), lz */
  double Re, Pr;
  int implicit[3];         /* param_implicit_{x,y,z} */
  int scalar;              /* 1: advance T (reference), 0: leave T as loaded */
  double uy_xm, uy_xp;     /* param_uy_x{m,p} */
  double t_xm, t_xp;       /* param_t_x{m,p} */
  double coef_dt_adv, coef_dt_dif;
  double dt_fixed;         /* > 0: decide_dt returns this value instead of the CFL value */
  int nthreads;            /* OpenMP threads (<= 0: library default) */
} orc_config_t;

typedef struct orc_t_ orc_t;

/* xf: nx+1 face positions (XF(1..nx+1)), xc: nx+2 centre positions (XC(0..nx+1)) */
orc_t * orc_create(const orc_config_t * cfg, const double * xf, const double * xc);
void orc_destroy(orc_t * o);

/* named arrays in the reference host layout; NULL for an unknown name.
 * fields: ux uy uz p psi t; tensors: lxx lyx0 lyx1 lxy lzx lxz lyy0 lyy1 lzy lyz lzz;
 * RK buffers: srcux0..2 srcuy0..2 srcuz0..2 srct0..2 (0=a, 1=b, 2=g);
 * metrics: xf xc hxxf hxxc hyxf hyxc jdxf jdxc */
double * orc_array(orc_t * o, const char * name, size_t * nitems);
double orc_hz(const orc_t * o);

int orc_update_boundaries(orc_t * o);   /* BCs + halos of ux uy uz p t (fluid/init.c:112-116) and psi */
int orc_decide_dt(orc_t * o, double * dt);
int orc_predict(orc_t * o, int rkstep, double dt);
int orc_potential(orc_t * o, int rkstep, double dt);
int orc_correct(orc_t * o, int rkstep, double dt);
int orc_pressure(orc_t * o, int rkstep, double dt);
int orc_integrate(orc_t * o, double * dt);
int orc_compute_tensors(orc_t * o);     /* the 11 velocity-gradient arrays + halos */

double orc_divergence(const orc_t * o);                 /* logging/divergence.c */
void orc_injection(const orc_t * o, double out[2]);     /* logging/injection.c, raw sums x diffusivity */
void orc_dissipation(const orc_t * o, double out[2]);   /* logging/dissipation.c */
void orc_total_energy(const orc_t * o, double out[4]);  /* logging/total_energy.c */
/* statistics.c collect: adds one sample to ux1 ux2 uy1 uy2 uz1 uz2 adv dif (orc_array names),
 * returns the sample count */
size_t orc_stat_collect(orc_t * o);

/* unit-level pieces, exposed so the tests can pin them on their own */
void orc_fft_c2c(int n, int sign, const double * in, double * out);  /* interleaved complex */
void orc_fft_r2c(int n, const double * in, double * out);            /* out: n/2+1 complex */
void orc_fft_c2r(int n, const double * in, double * out);            /* unnormalised */
int orc_gtsv(int n, const double * l, const double * c, const double * u, double * q);
int orc_gtsv_complex(int n, const double * l, const double * c, const double * u, double * q);
int orc_tdm_periodic(int n, const double * l, const double * c, const double * u, double * q);
void orc_rkcoefs(double out[9]);

/* Planted mistakes, for the tests that show each pin of the oracle fails on the slip it is
 * meant to catch.  ORC_MUT_NONE (the default) is the reference's algorithm; nothing but those
 * tests sets another value.  The setting is process-wide. */
enum {
  ORC_MUT_NONE = 0,
  ORC_MUT_UX_X_F,      /* ux implicit x solve with f instead of 2f (predict/ux.c:499-506) */
  ORC_MUT_UY_Y_F,      /* uy implicit y solve with f instead of 2f (predict/uy.c:539-544) */
  ORC_MUT_UZ_Z_F,      /* uz implicit z solve with f instead of 2f (predict/uz.c:449-454) */
  ORC_MUT_T_Y_F,       /* T implicit y solve with f: the predict/t.c:346-349 quirk undone */
  ORC_MUT_UX_Y_2F,     /* ux implicit y solve with 2f instead of f (predict/ux.c:508-524) */
  ORC_MUT_UZ_X_2F,     /* uz implicit x solve with 2f instead of f (predict/uz.c:418-420) */
  ORC_MUT_PF_ALPHA,    /* implicit prefactor from alpha instead of gamma (ux.c:496-497) */
  ORC_MUT_AB_SWAP,     /* RK combination weighs a with beta and b with alpha (ux.c:476-494) */
  ORC_MUT_NO_ROTATE,   /* reset_srcs zeroes a without the a/b pointer swap (predict/main.c:16-18) */
  ORC_MUT_P_F,         /* update_pressure with f instead of 2f (update_pressure.c:66, 98, 129) */
  ORC_MUT_P_NO_Z,      /* update_pressure leaves out the implicit z term (:165-167) */
  ORC_MUT_P_X_INDEX,   /* update_pressure x term with HXXF(i+1) in the i-1 coefficient (:54-59) */
  ORC_MUT_COUNT
};
void orc_set_mutation(int m);

#ifdef __cplusplus
}
#endif
#endif
