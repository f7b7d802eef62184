/*
 * tc_ref.c -- the reference's public API (domain, array, fluid_solver, integrate, decide_dt,
 * logging, save) on top of the device solver handle.
 *
 * The reference keeps its solver state in module statics initialised on first use; here
 * fluid_init binds a tc_solver_t to the caller's fluid_t (a small registry keyed by the
 * fluid_t address) and every entry point below runs on that handle.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <math.h>
#include <float.h>
#include <time.h>
#include "domain.h"
#include "array.h"
#include "fluid.h"
#include "fluid_solver.h"
#include "integrate.h"
#include "decide_dt.h"
#include "config.h"
#include "fileio.h"
#include "param.h"
#include "logging.h"
#include "save.h"
#include "statistics.h"
#include "tcstep.h"

/* ---------------------------------------------------------------------------------------
 * domain (src/domain.c:28-86, :259-432)
 * ------------------------------------------------------------------------------------- */
static double * dcalloc(size_t n){
  return calloc(n ? n : 1, sizeof(double));
}

int domain_init(const char dirname_ic[], domain_t * domain){
  memset(domain, 0, sizeof(*domain));
  bool is_curved = false;
  size_t glsizes[NDIMS] = {0};
  double lengths[NDIMS] = {0.};
  if(0 != fileio.r_serial(dirname_ic, "is_curved", 0, NULL, fileio.npy_boolean, sizeof(bool), &is_curved)) return 1;
  if(0 != fileio.r_serial(dirname_ic, "glsizes", 1, (size_t [1]){NDIMS}, fileio.npy_size_t, sizeof(size_t), glsizes)) return 1;
  if(0 != fileio.r_serial(dirname_ic, "lengths", 1, (size_t [1]){NDIMS}, fileio.npy_double, sizeof(double), lengths)) return 1;
  const size_t nx = glsizes[0], ny = glsizes[1], nz = glsizes[2];
  double * xf = dcalloc(nx + 1);
  double * xc = dcalloc(nx + 2);
  if(0 != fileio.r_serial(dirname_ic, "xf", 1, (size_t [1]){nx + 1}, fileio.npy_double, sizeof(double), xf)) return 1;
  if(0 != fileio.r_serial(dirname_ic, "xc", 1, (size_t [1]){nx + 2}, fileio.npy_double, sizeof(double), xc)) return 1;
  domain->info = NULL;
  domain->is_curved = is_curved;
  for(int d = 0; d < NDIMS; d++){
    domain->glsizes[d] = glsizes[d];
    domain->mysizes[d] = glsizes[d];
    domain->offsets[d] = 0;
    domain->lengths[d] = lengths[d];
  }
  domain->xf = xf;
  domain->xc = xc;
  /* metrics, domain.c:259-357 (the same operand order as the device tables) */
  double * hxxf = dcalloc(nx + 1), * hxxc = dcalloc(nx + 2);
  double * hyxf = dcalloc(nx + 1), * hyxc = dcalloc(nx + 2);
  double * jdxf = dcalloc(nx + 1), * jdxc = dcalloc(nx + 2);
  #define XF(i) (xf[(i) - 1])
  #define XC(i) (xc[(i)])
  for(size_t i = 1; i <= nx + 1; i++) hxxf[i - 1] = XC(i) - XC(i - 1);
  hxxc[0] = 0.5 * XF(2) - 0.5 * XF(1);
  for(size_t i = 1; i <= nx; i++) hxxc[i] = XF(i + 1) - XF(i);
  hxxc[nx + 1] = 0.5 * XF(nx + 1) - 0.5 * XF(nx);
  const double ly = lengths[1];
  for(size_t i = 1; i <= nx + 1; i++) hyxf[i - 1] = is_curved ? XF(i) * ly / ny : ly / ny;
  for(size_t i = 0; i <= nx + 1; i++) hyxc[i] = is_curved ? XC(i) * ly / ny : ly / ny;
  domain->hz = lengths[2] / nz;
  for(size_t i = 1; i <= nx + 1; i++) jdxf[i - 1] = hxxf[i - 1] * hyxf[i - 1] * domain->hz;
  for(size_t i = 0; i <= nx + 1; i++) jdxc[i] = hxxc[i] * hyxc[i] * domain->hz;
  #undef XF
  #undef XC
  domain->hxxf = hxxf; domain->hxxc = hxxc;
  domain->hyxf = hyxf; domain->hyxc = hyxc;
  domain->jdxf = jdxf; domain->jdxc = jdxc;
  printf("DOMAIN\n\t%s domain\n", is_curved ? "Cylindrical" : "Cartesian");
  for(int d = 0; d < NDIMS; d++) printf("\tglsizes[%d]: %zu\n", d, glsizes[d]);
  for(int d = 0; d < NDIMS; d++) printf("\tlengths[%d]: % .7e\n", d, lengths[d]);
  fflush(stdout);
  return 0;
}

int domain_save(const char dirname[], const domain_t * domain){
  const size_t nx = domain->glsizes[0];
  fileio.w_serial(dirname, "is_curved", 0, NULL, fileio.npy_boolean, sizeof(bool), &domain->is_curved);
  fileio.w_serial(dirname, "glsizes", 1, (size_t [1]){NDIMS}, fileio.npy_size_t, sizeof(size_t), domain->glsizes);
  fileio.w_serial(dirname, "lengths", 1, (size_t [1]){NDIMS}, fileio.npy_double, sizeof(double), domain->lengths);
  fileio.w_serial(dirname, "xf", 1, (size_t [1]){nx + 1}, fileio.npy_double, sizeof(double), domain->xf);
  fileio.w_serial(dirname, "xc", 1, (size_t [1]){nx + 2}, fileio.npy_double, sizeof(double), domain->xc);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * array (src/array.c:10-222), serial NPY
 * ------------------------------------------------------------------------------------- */
static int arr_prepare(const domain_t * domain, const int nadds[NDIMS][2], const size_t size, array_t * a){
  size_t nitems = 1;
  for(int d = 0; d < NDIMS; d++) nitems *= domain->mysizes[d] + nadds[d][0] + nadds[d][1];
  if(nitems > INT_MAX){
    printf("local array size (%zu) exceeds INT_MAX (%d)\n", nitems, INT_MAX);
    return 1;
  }
  a->size = size;
  a->nadds = calloc(NDIMS, 2 * sizeof(int));
  for(int d = 0; d < NDIMS; d++){ a->nadds[d][0] = nadds[d][0]; a->nadds[d][1] = nadds[d][1]; }
  a->datasize = nitems * size;
  a->data = calloc(nitems, size);
  if(!a->data){
    fprintf(stderr, "FATAL: allocation of %zu bytes failed\n", a->datasize);
    return 1;
  }
  return 0;
}

static int arr_destroy(array_t * a){
  free(a->nadds);
  free(a->data);
  a->nadds = NULL;
  a->data = NULL;
  a->datasize = 0;
  return 0;
}

/* element offset of (i, j, k) of the x-inclusive buffer in the halo'd array (array.c:57-68) */
static size_t arr_index(const domain_t * domain, const array_t * a, size_t i, size_t j, size_t k){
  const size_t isz = domain->mysizes[0] + a->nadds[0][0] + a->nadds[0][1];
  const size_t jsz = domain->mysizes[1] + a->nadds[1][0] + a->nadds[1][1];
  return i + (j + a->nadds[1][0]) * isz + (k + a->nadds[2][0]) * isz * jsz;
}

static int arr_load(const domain_t * domain, const char dirname[], const char dsetname[], const char dtype[], array_t * a){
  const size_t * n = domain->mysizes;
  const size_t imax = n[0] + a->nadds[0][0] + a->nadds[0][1];
  char * buf = calloc(imax * n[1] * n[2], a->size);
  if(!buf) return 1;
  if(0 != fileio.r_serial(dirname, dsetname, 3, (size_t [3]){n[2], n[1], imax}, dtype, a->size, buf)){
    free(buf);
    return 1;
  }
  size_t cnt = 0;
  for(size_t k = 0; k < n[2]; k++) for(size_t j = 0; j < n[1]; j++) for(size_t i = 0; i < imax; i++, cnt++){
    memcpy((char *)a->data + a->size * arr_index(domain, a, i, j, k), buf + a->size * cnt, a->size);
  }
  free(buf);
  return 0;
}

static int arr_dump(const domain_t * domain, const char dirname[], const char dsetname[], const char dtype[], const array_t * a){
  const size_t * n = domain->mysizes;
  const size_t imax = n[0] + a->nadds[0][0] + a->nadds[0][1];
  char * buf = calloc(imax * n[1] * n[2], a->size);
  if(!buf) return 1;
  size_t cnt = 0;
  for(size_t k = 0; k < n[2]; k++) for(size_t j = 0; j < n[1]; j++) for(size_t i = 0; i < imax; i++, cnt++){
    memcpy(buf + a->size * cnt, (const char *)a->data + a->size * arr_index(domain, a, i, j, k), a->size);
  }
  const int rc = fileio.w_serial(dirname, dsetname, 3, (size_t [3]){n[2], n[1], imax}, dtype, a->size, buf);
  free(buf);
  return rc;
}

const array_method_t array = {
  .prepare = arr_prepare,
  .destroy = arr_destroy,
  .load = arr_load,
  .dump = arr_dump,
};

/* ---------------------------------------------------------------------------------------
 * fluid <-> device solver registry
 * ------------------------------------------------------------------------------------- */
#define TC_MAXREG 16
static struct { const fluid_t * f; tc_solver_t * s; int lazy; int coefs; } g_reg[TC_MAXREG];

static int reg_find(const fluid_t * f){
  for(int n = 0; n < TC_MAXREG; n++) if(g_reg[n].f == f && g_reg[n].s) return n;
  fprintf(stderr, "fluid_t %p has no device state: call fluid_init first\n", (const void *)f);
  return -1;
}

static const char * g_names[6] = {"ux", "uy", "uz", "p", "psi", "t"};

static const array_t * field_array(const fluid_t * f, int n){
  const array_t * a[6] = {&f->ux, &f->uy, &f->uz, &f->p, &f->psi, &f->t};
  return a[n];
}

int tc_fluid_sync_device(const fluid_t * fluid){
  const int r = reg_find(fluid);
  if(r < 0) return 1;
  for(int n = 0; n < 6; n++){
    if(n == 4) continue;   /* psi is an output of the step */
    if(tc_set_field_ref(g_reg[r].s, g_names[n], field_array(fluid, n)->data)){
      fprintf(stderr, "%s\n", tc_last_error());
      return 1;
    }
  }
  return 0;
}

int tc_fluid_sync_host(fluid_t * fluid){
  const int r = reg_find(fluid);
  if(r < 0) return 1;
  for(int n = 0; n < 6; n++){
    if(tc_get_field_ref(g_reg[r].s, g_names[n], field_array(fluid, n)->data)){
      fprintf(stderr, "%s\n", tc_last_error());
      return 1;
    }
  }
  return 0;
}

int tc_fluid_finalise(fluid_t * fluid){
  for(int n = 0; n < TC_MAXREG; n++){
    if(g_reg[n].f == fluid && g_reg[n].s){
      tc_destroy(g_reg[n].s);
      g_reg[n].s = NULL;
      g_reg[n].f = NULL;
    }
  }
  return 0;
}

static int before(fluid_t * fluid, int * r){
  *r = reg_find(fluid);
  if(*r < 0) return 1;
  if(!g_reg[*r].lazy) return tc_fluid_sync_device(fluid);
  return 0;
}

static int after(fluid_t * fluid, int r, int rc){
  if(rc){
    fprintf(stderr, "%s\n", tc_last_error());
    return 1;
  }
  if(!g_reg[r].lazy) return tc_fluid_sync_host(fluid);
  return 0;
}

/* decide_dt reads its coefficients on first use (decide_dt.c:126-136) */
static int ensure_coefs(int r){
  if(g_reg[r].coefs) return 0;
  double adv = 0., dif = 0., fixed = 0.;
  if(0 != config.get_double("coef_dt_adv", &adv)) return 1;
  if(0 != config.get_double("coef_dt_dif", &dif)) return 1;
  const char * s = getenv("dt_fixed");
  if(s) fixed = strtod(s, NULL);
  printf("coefs: (adv) % .3e, (dif) % .3e\n", adv, dif);
  tc_set_dt_params(g_reg[r].s, adv, dif, fixed);
  g_reg[r].coefs = 1;
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * fluid API (src/fluid/*.c)
 * ------------------------------------------------------------------------------------- */
double fluid_compute_momentum_diffusivity(const fluid_t * fluid){
  return 1. / fluid->Re;
}

double fluid_compute_scalar_diffusivity(const fluid_t * fluid){
  return 1. / fluid->Re / fluid->Pr;
}

/* boundary conditions on host arrays, src/fluid/boundary/*.c with src/halo.c */
static void host_halo(const domain_t * d, array_t * a){
  const size_t isz = d->mysizes[0] + a->nadds[0][0] + a->nadds[0][1];
  const size_t jsz = d->mysizes[1] + 2;
  const size_t ksz = d->mysizes[2] + 2;
  double * x = a->data;
  for(size_t k = 0; k < ksz; k++){
    double * pl = x + isz * jsz * k;
    memcpy(pl, pl + isz * (jsz - 2), sizeof(double) * isz);
    memcpy(pl + isz * (jsz - 1), pl + isz, sizeof(double) * isz);
  }
  memcpy(x, x + isz * jsz * (ksz - 2), sizeof(double) * isz * jsz);
  memcpy(x + isz * jsz * (ksz - 1), x + isz * jsz, sizeof(double) * isz * jsz);
}

static int host_walls(const domain_t * d, array_t * a, int kind, double wm, double wp){
  const size_t nx = d->mysizes[0], ny = d->mysizes[1], nz = d->mysizes[2];
  const size_t isz = nx + a->nadds[0][0] + a->nadds[0][1];
  double * x = a->data;
  for(size_t k = 1; k <= nz; k++){
    for(size_t j = 1; j <= ny; j++){
      double * row = x + isz * (j + (ny + 2) * k);
      if(kind == 0){ row[0] = 0.; row[nx] = 0.; }            /* UX(1), UX(nx+1) */
      else if(kind == 4){ row[0] = row[1]; row[nx + 1] = row[nx]; }
      else { row[0] = wm; row[nx + 1] = wp; }
    }
  }
  host_halo(d, a);
  return 0;
}

int fluid_update_boundaries_ux(const domain_t * domain, array_t * a){ return host_walls(domain, a, 0, 0., 0.); }
int fluid_update_boundaries_uy(const domain_t * domain, array_t * a){ return host_walls(domain, a, 1, param_uy_xm, param_uy_xp); }
int fluid_update_boundaries_uz(const domain_t * domain, array_t * a){ return host_walls(domain, a, 2, 0., 0.); }
int fluid_update_boundaries_p(const domain_t * domain, array_t * a){ return host_walls(domain, a, 4, 0., 0.); }
int fluid_update_boundaries_psi(const domain_t * domain, array_t * a){ return host_walls(domain, a, 4, 0., 0.); }
int fluid_update_boundaries_t(const domain_t * domain, array_t * a){ return host_walls(domain, a, 3, param_t_xm, param_t_xp); }

int fluid_init(const char dirname_ic[], const domain_t * domain, fluid_t * fluid){
  tc_param_load();
  memset(fluid, 0, sizeof(*fluid));
  static const int nadds_x[NDIMS][2] = {{0, 1}, {1, 1}, {1, 1}};
  static const int nadds_c[NDIMS][2] = {{1, 1}, {1, 1}, {1, 1}};
  if(0 != array.prepare(domain, nadds_x, sizeof(double), &fluid->ux)) return 1;
  if(0 != array.prepare(domain, nadds_c, sizeof(double), &fluid->uy)) return 1;
  if(0 != array.prepare(domain, nadds_c, sizeof(double), &fluid->uz)) return 1;
  if(0 != array.prepare(domain, nadds_c, sizeof(double), &fluid->p)) return 1;
  if(0 != array.prepare(domain, nadds_c, sizeof(double), &fluid->psi)) return 1;
  if(0 != array.prepare(domain, nadds_c, sizeof(double), &fluid->t)) return 1;
  if(0 != array.load(domain, dirname_ic, "ux", fileio.npy_double, &fluid->ux)) return 1;
  if(0 != array.load(domain, dirname_ic, "uy", fileio.npy_double, &fluid->uy)) return 1;
  if(0 != array.load(domain, dirname_ic, "uz", fileio.npy_double, &fluid->uz)) return 1;
  if(0 != array.load(domain, dirname_ic, "p", fileio.npy_double, &fluid->p)) return 1;
  if(0 != array.load(domain, dirname_ic, "t", fileio.npy_double, &fluid->t)) return 1;
  fluid_update_boundaries_ux(domain, &fluid->ux);
  fluid_update_boundaries_uy(domain, &fluid->uy);
  fluid_update_boundaries_uz(domain, &fluid->uz);
  fluid_update_boundaries_p(domain, &fluid->p);
  fluid_update_boundaries_t(domain, &fluid->t);
  if(0 != config.get_double("Re", &fluid->Re)) return 1;
  if(0 != config.get_double("Pr", &fluid->Pr)) return 1;
  /* device state bound to this fluid_t */
  tc_params_t p;
  tc_params_default(&p);
  p.nx = (int)domain->glsizes[0];
  p.ny = (int)domain->glsizes[1];
  p.nz = (int)domain->glsizes[2];
  p.is_curved = domain->is_curved;
  for(int d = 0; d < NDIMS; d++) p.lengths[d] = domain->lengths[d];
  p.Re = fluid->Re;
  p.Pr = fluid->Pr;
  p.implicit[0] = param_implicit_x;
  p.implicit[1] = param_implicit_y;
  p.implicit[2] = param_implicit_z;
  p.scalar = param_scalar;
  p.uy_xm = param_uy_xm; p.uy_xp = param_uy_xp;
  p.t_xm = param_t_xm; p.t_xp = param_t_xp;
  const char * prec = getenv("TC_PRECISION");
  if(prec && strcmp(prec, "mixed") == 0) p.precision = TC_MIXED;
  if(prec && strcmp(prec, "mixed_refined") == 0) p.precision = TC_MIXED_REFINED;
  int slot = -1;
  for(int n = 0; n < TC_MAXREG; n++) if(!g_reg[n].s){ slot = n; break; }
  if(slot < 0){
    fprintf(stderr, "fluid_init: too many fluids\n");
    return 1;
  }
  tc_solver_t * s = NULL;
  if(tc_create(&p, domain->xf, domain->xc, &s)){
    fprintf(stderr, "fluid_init: %s\n", tc_last_error());
    return 1;
  }
  const char * coh = getenv("TC_COHERENCE");
  g_reg[slot].f = fluid;
  g_reg[slot].s = s;
  g_reg[slot].lazy = coh && strcmp(coh, "lazy") == 0;
  g_reg[slot].coefs = 0;
  tc_arm_diagnostics(s, 1);
  if(tc_fluid_sync_device(fluid)) return 1;
  printf("FLUID\n\tRe: % .7e\n\tPr: % .7e\n", fluid->Re, fluid->Pr);
  printf("\tdiffusive treatment in x: %s\n", param_implicit_x ? "implicit" : "explicit");
  printf("\tdiffusive treatment in y: %s\n", param_implicit_y ? "implicit" : "explicit");
  printf("\tdiffusive treatment in z: %s\n", param_implicit_z ? "implicit" : "explicit");
  fflush(stdout);
  return 0;
}

int fluid_save(const char dirname[], const domain_t * domain, const fluid_t * fluid){
  const int r = reg_find(fluid);
  if(r >= 0 && g_reg[r].lazy && tc_fluid_sync_host((fluid_t *)fluid)) return 1;
  fileio.w_serial(dirname, "Re", 0, NULL, fileio.npy_double, sizeof(double), &fluid->Re);
  fileio.w_serial(dirname, "Pr", 0, NULL, fileio.npy_double, sizeof(double), &fluid->Pr);
  array.dump(domain, dirname, "ux", fileio.npy_double, &fluid->ux);
  array.dump(domain, dirname, "uy", fileio.npy_double, &fluid->uy);
  array.dump(domain, dirname, "uz", fileio.npy_double, &fluid->uz);
  array.dump(domain, dirname, "p", fileio.npy_double, &fluid->p);
  array.dump(domain, dirname, "t", fileio.npy_double, &fluid->t);
  return 0;
}

int fluid_predict_field(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid){
  (void)domain;
  int r;
  if(before(fluid, &r)) return 1;
  return after(fluid, r, tc_predict(g_reg[r].s, (int)rkstep, dt));
}

int fluid_compute_potential(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid){
  (void)domain;
  int r;
  if(before(fluid, &r)) return 1;
  return after(fluid, r, tc_potential(g_reg[r].s, (int)rkstep, dt));
}

int fluid_correct_velocity(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid){
  (void)domain;
  int r;
  if(before(fluid, &r)) return 1;
  return after(fluid, r, tc_correct(g_reg[r].s, (int)rkstep, dt));
}

int fluid_update_pressure(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid){
  (void)domain;
  int r;
  if(before(fluid, &r)) return 1;
  return after(fluid, r, tc_pressure(g_reg[r].s, (int)rkstep, dt));
}

int decide_dt(const domain_t * domain, const fluid_t * fluid, double * dt){
  (void)domain;
  const int r = reg_find(fluid);
  if(r < 0 || ensure_coefs(r)) return 1;
  if(!g_reg[r].lazy && tc_fluid_sync_device(fluid)) return 1;
  if(tc_decide_dt(g_reg[r].s, dt)){
    fprintf(stderr, "%s\n", tc_last_error());
    return 1;
  }
  return 0;
}

int integrate(const domain_t * domain, fluid_t * fluid, double * dt){
  (void)domain;
  int r = reg_find(fluid);
  if(r < 0 || ensure_coefs(r)) return 1;
  if(before(fluid, &r)) return 1;
  return after(fluid, r, tc_step(g_reg[r].s, dt));
}

/* ---------------------------------------------------------------------------------------
 * logging (src/logging/*.c): device reductions, reference file formats
 * ------------------------------------------------------------------------------------- */
static double g_log_rate = 0., g_log_next = 0.;

static int log_init(const domain_t * domain, const double time){
  (void)domain;
  if(0 != config.get_double("log_rate", &g_log_rate)) return 1;
  g_log_next = g_log_rate * ceil(fmax(DBL_EPSILON, time) / g_log_rate);
  printf("LOGGING\n\tnext: % .3e\n\trate: % .3e\n", g_log_next, g_log_rate);
  fflush(stdout);
  return 0;
}

static tc_solver_t * solver_of(const fluid_t * fluid){
  const int r = reg_find(fluid);
  return r < 0 ? NULL : g_reg[r].s;
}

int logging_check_divergence(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid){
  (void)domain;
  tc_solver_t * s = solver_of(fluid);
  double divmax = 0.;
  if(!s || tc_divergence(s, &divmax)) return 1;
  FILE * fp = fileio.fopen(fname, "a");
  if(!fp) return 1;
  fprintf(fp, "%8.2f % .1e\n", time, divmax);
  fileio.fclose(fp);
  return 0;
}

int logging_check_injection(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid){
  (void)domain;
  tc_solver_t * s = solver_of(fluid);
  double v[2];
  if(!s || tc_injection(s, v)) return 1;
  FILE * fp = fileio.fopen(fname, "a");
  if(!fp) return 1;
  fprintf(fp, "%8.2f ", time);
  fprintf(fp, "% 18.15e ", v[0]);
  fprintf(fp, "% 18.15e ", v[1]);
  fprintf(fp, "\n");
  fileio.fclose(fp);
  return 0;
}

int logging_check_dissipation(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid){
  (void)domain;
  tc_solver_t * s = solver_of(fluid);
  double v[2];
  if(!s || tc_dissipation(s, v)) return 1;
  FILE * fp = fileio.fopen(fname, "a");
  if(!fp) return 1;
  fprintf(fp, "%8.2f ", time);
  fprintf(fp, "% 18.15e ", v[0]);
  fprintf(fp, "% 18.15e ", v[1]);
  fprintf(fp, "\n");
  fileio.fclose(fp);
  return 0;
}

int logging_check_total_energy(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid){
  (void)domain;
  tc_solver_t * s = solver_of(fluid);
  double q[4];
  if(!s || tc_total_energy(s, q)) return 1;
  FILE * fp = fileio.fopen(fname, "a");
  if(!fp) return 1;
  fprintf(fp, "%8.2f ", time);
  for(int n = 0; n < 4; n++) fprintf(fp, "% 18.15e%c", q[n], n == 3 ? '\n' : ' ');
  fileio.fclose(fp);
  return 0;
}

static void log_check_and_output(const domain_t * domain, const size_t step, const double time,
    const double dt, const double wtime, const fluid_t * fluid){
  FILE * fp = fileio.fopen("output/log/progress.dat", "a");
  if(fp){
    fprintf(fp, "step %zu, time %.1f, dt %.2e, elapsed %.1f [sec]\n", step, time, dt, wtime);
    fprintf(stdout, "step %zu, time %.1f, dt %.2e, elapsed %.1f [sec]\n", step, time, dt, wtime);
    fileio.fclose(fp);
  }
  logging_check_divergence("output/log/divergence.dat", domain, time, fluid);
  logging_check_injection("output/log/injection.dat", domain, time, fluid);
  logging_check_dissipation("output/log/dissipation.dat", domain, time, fluid);
  logging_check_total_energy("output/log/total_energy.dat", domain, time, fluid);
  g_log_next += g_log_rate;
}

static double log_next(void){
  return g_log_next;
}

const logging_t logging = {
  .init = log_init,
  .check_and_output = log_check_and_output,
  .get_next_time = log_next,
};

/* ---------------------------------------------------------------------------------------
 * save (src/save.c:31-116)
 * ------------------------------------------------------------------------------------- */
static double g_save_rate = 0., g_save_next = 0.;

static int save_init(const domain_t * domain, const double time){
  (void)domain;
  if(0 != config.get_double("save_rate", &g_save_rate)) return 1;
  double after_ = 0.;
  if(0 != config.get_double("save_after", &after_)) return 1;
  g_save_next = g_save_rate * ceil(fmax(DBL_EPSILON, fmax(time, after_)) / g_save_rate);
  printf("SAVE\n\tdest: output/save/step\n\tnext: % .3e\n\trate: % .3e\n", g_save_next, g_save_rate);
  fflush(stdout);
  return 0;
}

static int save_output(const domain_t * domain, const size_t step, const double time, const fluid_t * fluid){
  char dirname[64];
  snprintf(dirname, sizeof(dirname), "output/save/step%010zu", step);
  if(0 != fileio.mkdir(dirname)) return 1;
  fileio.w_serial(dirname, "step", 0, NULL, fileio.npy_size_t, sizeof(size_t), &step);
  fileio.w_serial(dirname, "time", 0, NULL, fileio.npy_double, sizeof(double), &time);
  domain_save(dirname, domain);
  fluid_save(dirname, domain, fluid);
  g_save_next += g_save_rate;
  return 0;
}

static double save_next(void){
  return g_save_next;
}

const save_t save = {
  .init = save_init,
  .output = save_output,
  .get_next_time = save_next,
};

/* ---------------------------------------------------------------------------------------
 * statistics (src/statistics.c:56-307): accumulated on the device of the fluid's solver
 * ------------------------------------------------------------------------------------- */
static double g_stat_rate = 0., g_stat_next = 0.;
static const fluid_t * g_stat_fluid = NULL;

static int stat_init(const domain_t * domain, const double time){
  (void)domain;
  if(0 != config.get_double("stat_rate", &g_stat_rate)) return 1;
  double after_ = 0.;
  if(0 != config.get_double("stat_after", &after_)) return 1;
  g_stat_next = g_stat_rate * ceil(fmax(DBL_EPSILON, fmax(time, after_)) / g_stat_rate);
  /* every bound solver keeps the tensors of its last substep from now on */
  for(int n = 0; n < TC_MAXREG; n++){
    if(g_reg[n].s && tc_stat_arm(g_reg[n].s, 1)){
      fprintf(stderr, "%s\n", tc_last_error());
      return 1;
    }
  }
  printf("STATISTICS\n\tdest: output/stat/step\n\tnext: % .3e\n\trate: % .3e\n", g_stat_next, g_stat_rate);
  fflush(stdout);
  return 0;
}

static int stat_collect(const domain_t * domain, const fluid_t * fluid){
  (void)domain;
  tc_solver_t * s = solver_of(fluid);
  if(!s) return 1;
  if(tc_stat_arm(s, 1) || tc_stat_collect(s)){
    fprintf(stderr, "%s\n", tc_last_error());
    return 1;
  }
  g_stat_fluid = fluid;
  g_stat_next += g_stat_rate;
  return 0;
}

static int stat_output(const domain_t * domain, const size_t step){
  if(!g_stat_fluid) return 0;
  tc_solver_t * s = solver_of(g_stat_fluid);
  if(!s) return 1;
  size_t num = 0;
  if(tc_stat_get(s, "ux1", NULL, &num)) return 1;
  if(0 == num) return 0;
  char dirname[64];
  snprintf(dirname, sizeof(dirname), "output/stat/step%010zu", step);
  fileio.mkdir(dirname);
  fileio.w_serial(dirname, "num", 0, NULL, fileio.npy_size_t, sizeof(size_t), &num);
  domain_save(dirname, domain);
  static const char * const names[8] = {"ux1", "ux2", "uy1", "uy2", "uz1", "uz2", "adv", "dif"};
  static const int face[8] = {1, 1, 0, 0, 0, 0, 1, 1};
  for(int n = 0; n < 8; n++){
    static const int nadds_f[NDIMS][2] = {{0, 1}, {0, 0}, {0, 0}};
    static const int nadds_c[NDIMS][2] = {{1, 1}, {0, 0}, {0, 0}};
    array_t a = {0};
    if(array.prepare(domain, face[n] ? nadds_f : nadds_c, sizeof(double), &a)) return 1;
    if(tc_stat_get(s, names[n], a.data, NULL)){
      fprintf(stderr, "%s\n", tc_last_error());
      array.destroy(&a);
      return 1;
    }
    array.dump(domain, dirname, names[n], fileio.npy_double, &a);
    array.destroy(&a);
  }
  return 0;
}

static double stat_next(void){
  return g_stat_next;
}

const statistics_t statistics = {
  .init = stat_init,
  .collect = stat_collect,
  .output = stat_output,
  .get_next_time = stat_next,
};
