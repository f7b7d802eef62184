/*
 * tc_ic.c -- seeded initial conditions in the reference's NPY layout.
 *
 * The reference's own generator (initial_condition/main.py:52-113) draws unseeded uniform
 * noise on a clipped-Chebyshev grid; the workloads here need a seeded, reproducible field on
 * a tanh-clustered grid with a laminar profile, so this tool defines one:
 *   grid     xf[i] = ri + lx/2 (1 + tanh(s (2i/nx - 1)) / tanh(s)),  xc as in main.py:71-77
 *   laminar  circular Couette uy = A r + B / r between the wall values (curved) or the
 *            linear profile (planar); conduction T = log(r/ro)/log(ri/ro) profile
 *   noise    amp (2u - 1), u = (splitmix64(key) >> 11) 2^-53,
 *            key = seed * 0x9E3779B97F4A7C15 + (field << 56) + linear NPY index
 *   modes    mode_amp X(x) Y(y) Z(z), each factor the mean of nmodes seeded sine/cosine
 *            modes (vanishing at the walls in x), a smooth "turbulent-like" start
 * The per-index factors (profile, modes) are evaluated on the host with the C library's
 * transcendental functions; the device kernel adds the hash noise and the mode product in a
 * fixed operand order, so the field equals one a host loop would write, bit for bit.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <hip/hip_runtime_api.h>
#include "tcstep.h"
#include "tc_internal.h"
#include "tc_kernels.h"

static unsigned long long splitmix64(unsigned long long z){
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double uniform01(unsigned long long seed, unsigned field, unsigned long long lin){
  const unsigned long long key = seed * 0x9E3779B97F4A7C15ULL + ((unsigned long long)field << 56) + lin;
  return (double)(splitmix64(key) >> 11) * 0x1.0p-53;
}

int tc_ic_grid(int nx, double lx, int is_curved, double stretch, double * xf, double * xc){
  if(nx < 2) return tc_fail("nx too small");
  const double ri = is_curved ? 1. : 0.;
  for(int i = 0; i <= nx; i++){
    const double fi = (double)i;
    if(stretch > 0.) xf[i] = ri + lx * 0.5 * (1. + tanh(stretch * (2. * fi / nx - 1.)) / tanh(stretch));
    else xf[i] = ri + lx * fi / nx;
  }
  xf[0] = ri;
  xf[nx] = ri + lx;
  xc[0] = xf[0];
  for(int i = 1; i <= nx; i++) xc[i] = 0.5 * xf[i - 1] + 0.5 * xf[i];
  xc[nx + 1] = xf[nx];
  return 0;
}

static double laminar_uy(double r, double ri, double ro, double ui, double uo, int curved){
  if(curved){
    const double a = (uo * ro - ui * ri) / (ro * ro - ri * ri);
    const double b = ri * ro * (ui * ro - uo * ri) / (ro * ro - ri * ri);
    return a * r + b / r;
  }
  return uo + (ui - uo) * (ro - r) / (ro - ri);
}

static double conduction_t(double r, double ri, double ro, double ti, double to, int curved){
  if(curved) return to + (ti - to) * log(r / ro) / log(ri / ro);
  return to + (ti - to) * (ro - r) / (ro - ri);
}

/* mean of nm seeded modes; kind 0: sin(pi m xi) (walls), 1: cos(2 pi m t / n + phase) */
static void mode_table(unsigned long long seed, unsigned field, unsigned dir, int nm, int kind,
    int n, const double * pos, double * out){
  const double two_pi = 6.283185307179586476925286766559;
  const double pi = 3.14159265358979323846;
  for(int q = 0; q < n; q++) out[q] = 0.;
  for(int m = 1; m <= nm; m++){
    const unsigned long long tag = ((unsigned long long)dir << 40) + ((unsigned long long)m << 8);
    const double amp = 2. * uniform01(seed, 16 + field, tag) - 1.;
    const double phase = two_pi * uniform01(seed, 32 + field, tag);
    for(int q = 0; q < n; q++){
      const double v = kind == 0 ? sin(pi * m * pos[q]) : cos(two_pi * m * pos[q] + phase);
      out[q] += amp * v / nm;
    }
  }
}

int tck_launch_ic_fill(const tc_geom_t * g, double * dst, const double * prof, const double * mx,
    const double * my, const double * mz, double amp, double mode_amp, unsigned long long seed,
    unsigned field, int nxs, int i0, hipStream_t stream);

/* one host table into device staging; the host table is reused as soon as this returns */
static int put_table(tc_solver_t * s, double * d, const double * h, int n){
  if(hipMemcpyAsync(d, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
     hipStreamSynchronize(s->stream) != hipSuccess)
    return tc_fail("initial condition: table upload failed");
  return 0;
}

/* The fields of tc_ic_fields, given device staging d (four tables ld doubles apart: profile,
 * x, y and z mode factors) and host work space tab and pos (nmax doubles each).  The per-index
 * tables are evaluated here; the device kernel tck_ic_fill combines them with the splitmix64
 * noise in the same operand order as a host loop would, so the field equals the one a
 * host-side generator writes. */
static int ic_fill(tc_solver_t * s, unsigned long long seed, double amp, double t_amp, int nmodes,
    double mode_amp, double * d, size_t ld, double * tab, double * pos){
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
  const int curved = s->p.is_curved;
  const double * xf = s->h_xf;   /* XF(i) at [i] */
  const double * xc = s->h_xc;
  const double ri = xf[1], ro = xf[nx + 1];
  const double lx = ro - ri;
  static const int fmap[5] = {TC_FUX, TC_FUY, TC_FUZ, TC_FT, TC_FP};
  for(int f = 0; f < 5; f++){
    const int nxs = f == 0 ? nx + 1 : nx + 2;
    const int i0 = f == 0 ? 1 : 0;
    for(int q = 0; q < nxs; q++){
      const double r = f == 0 ? xf[q + 1] : xc[q];
      double v = 0.;
      if(f == 1) v = laminar_uy(r, ri, ro, s->p.uy_xm, s->p.uy_xp, curved);
      if(f == 3) v = conduction_t(r, ri, ro, s->p.t_xm, s->p.t_xp, curved);
      tab[q] = v;
    }
    if(put_table(s, d, tab, nxs)) return 1;
    const int use_modes = nmodes > 0 && mode_amp != 0. && f < 3;
    if(use_modes){
      for(int q = 0; q < nxs; q++) pos[q] = ((f == 0 ? xf[q + 1] : xc[q]) - ri) / lx;
      mode_table(seed, f, 0, nmodes, 0, nxs, pos, tab);
      if(put_table(s, d + ld, tab, nxs)) return 1;
      for(int q = 0; q < ny; q++) pos[q] = (double)q / ny;
      mode_table(seed, f, 1, nmodes, 1, ny, pos, tab);
      if(put_table(s, d + 2 * ld, tab, ny)) return 1;
      for(int q = 0; q < nz; q++) pos[q] = (double)q / nz;
      mode_table(seed, f, 2, nmodes, 1, nz, pos, tab);
      if(put_table(s, d + 3 * ld, tab, nz)) return 1;
    }
    const double a = f < 3 ? amp : (f == 3 ? t_amp : 0.);
    double * dst = tc_field_ptr(s, fmap[f]);
    if(tck_launch_ic_fill(&s->g, dst, d, use_modes ? d + ld : NULL, d + 2 * ld, d + 3 * ld, a,
         mode_amp, seed, (unsigned)f, nxs, i0, s->stream) != 0 ||
       hipStreamSynchronize(s->stream) != hipSuccess)
      return tc_fail("initial condition fill failed");
    if(tc_finish_field(s, fmap[f])) return 1;
  }
  return 0;
}

static int ic_nmax(const tc_solver_t * s){
  const int a = s->p.nx + 2, b = s->p.ny, c = s->p.nz;
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

int tc_ic_fields(tc_solver_t * s, unsigned long long seed, double amp, double t_amp, int nmodes, double mode_amp){
  HIPOK(hipSetDevice(s->p.device));
  const int nmax = ic_nmax(s);
  double * h = calloc((size_t)2 * nmax, sizeof(double));
  double * d = NULL;
  if(!h) return tc_fail("host allocation failed");
  if(hipMalloc((void **)&d, sizeof(double) * 4 * nmax) != hipSuccess){
    free(h);
    return tc_fail("hipMalloc failed");
  }
  const int rc = ic_fill(s, seed, amp, t_amp, nmodes, mode_amp, d, (size_t)nmax, h, h + nmax);
  hipFree(d);
  free(h);
  return rc;
}

/* the same fields with no allocation: device staging in the library's per-device array of
 * four TC_IC_NOALLOC_N-double tables (tck_ic_scratch), host work space on the stack */
int tc_ic_fields_noalloc(tc_solver_t * s, unsigned long long seed, double amp, double t_amp, int nmodes, double mode_amp){
  HIPOK(hipSetDevice(s->p.device));
  const int nmax = ic_nmax(s);
  if(nmax > TC_IC_NOALLOC_N)
    return tc_fail("tc_ic_fields_noalloc: grid %d x %d x %d needs tables of %d entries, the limit is %d (nx + 2, ny and nz)",
        s->p.nx, s->p.ny, s->p.nz, nmax, TC_IC_NOALLOC_N);
  double * d = NULL;
  const int e = tck_ic_scratch(&d);
  if(e) return tc_fail("tc_ic_fields_noalloc: device address of the staging array: %s", hipGetErrorString((hipError_t)e));
  double tab[TC_IC_NOALLOC_N], pos[TC_IC_NOALLOC_N];
  return ic_fill(s, seed, amp, t_amp, nmodes, mode_amp, d, TC_IC_NOALLOC_N, tab, pos);
}
