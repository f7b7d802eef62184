/*
 * tc_npy.c -- NPY v1.0 serial I/O, environment config and run-time parameters.
 *
 * fileio (src/fileio.c:20-362, serial part): the reference reads and writes NPY headers with
 * its SimpleNpyIO submodule; this file writes the NPY v1.0 format directly (magic, version,
 * little-endian header length, Python-dict header padded to a multiple of 64 bytes) and
 * reads versions 1.0 to 3.0.  The dtype strings keep the reference's quoted form ("'<f8'").
 * config (src/config.c:14-42) and param (src/param/*.c) follow.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <stdint.h>
#include <sys/stat.h>
#include <sys/types.h>
#include "fileio.h"
#include "config.h"
#include "param.h"
#include "runge_kutta.h"

static const char NPY_BOOLEAN[] = "'|b1'";
static const char NPY_SIZE_T[] = "'<u8'";
static const char NPY_DOUBLE[] = "'<f8'";

static char * npy_path(const char dirname[], const char dsetname[]){
  const size_t n = strlen(dirname);
  const char * slash = (n > 0 && dirname[n - 1] == '/') ? "" : "/";
  const size_t len = n + strlen(slash) + strlen(dsetname) + 5;
  char * s = calloc(len, 1);
  if(s) snprintf(s, len, "%s%s%s.npy", dirname, slash, dsetname);
  return s;
}

static int fio_init(void){
  if(sizeof(size_t) != 8 || sizeof(double) != 8){
    fprintf(stderr, "fileio: size_t and double must be 8 bytes\n");
    return 1;
  }
  return 0;
}

static FILE * fio_fopen(const char * path, const char * mode){
  errno = 0;
  FILE * fp = fopen(path, mode);
  if(!fp) perror(path);
  return fp;
}

static int fio_fclose(FILE * stream){
  if(!stream) return 1;
  fclose(stream);
  return 0;
}

static int fio_mkdir(const char dirname[]){
  errno = 0;
  if(0 != mkdir(dirname, S_IRWXU | S_IRWXG | S_IRWXO)){
    perror(dirname);
    return 1;
  }
  return 0;
}

/* header: {'descr': <dtype>, 'fortran_order': False, 'shape': (...), } */
static int write_header(FILE * fp, size_t ndims, const size_t * shape, const char dtype[]){
  char dict[512];
  char shp[256];
  size_t o = 0;
  o += snprintf(shp + o, sizeof(shp) - o, "(");
  for(size_t d = 0; d < ndims; d++){
    o += snprintf(shp + o, sizeof(shp) - o, "%zu%s", shape[d], (ndims == 1) ? "," : (d + 1 < ndims ? ", " : ""));
  }
  snprintf(shp + o, sizeof(shp) - o, ")");
  int len = snprintf(dict, sizeof(dict), "{'descr': %s, 'fortran_order': False, 'shape': %s, }", dtype, shp);
  if(len < 0 || (size_t)len >= sizeof(dict) - 70) return 1;
  /* pad so that magic(6) + version(2) + length(2) + header is a multiple of 64, ending in \n */
  size_t total = 10 + (size_t)len + 1;
  size_t pad = (64 - total % 64) % 64;
  for(size_t q = 0; q < pad; q++) dict[len++] = ' ';
  dict[len++] = '\n';
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t hl = (uint16_t)len;
  const unsigned char hlb[2] = {(unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
  if(fwrite(magic, 1, 8, fp) != 8 || fwrite(hlb, 1, 2, fp) != 2 || fwrite(dict, 1, (size_t)len, fp) != (size_t)len) return 1;
  return 0;
}

/* parse the header; on success *offset is the byte offset of the data */
static int read_header(FILE * fp, size_t * ndims, size_t shape[8], char dtype[32], int * fortran, size_t * offset){
  unsigned char pre[8];
  if(fread(pre, 1, 8, fp) != 8) return 1;
  if(pre[0] != 0x93 || memcmp(pre + 1, "NUMPY", 5) != 0) return 1;
  size_t hl = 0;
  if(pre[6] == 1){
    unsigned char b[2];
    if(fread(b, 1, 2, fp) != 2) return 1;
    hl = (size_t)b[0] | ((size_t)b[1] << 8);
    *offset = 10 + hl;
  }else{
    unsigned char b[4];
    if(fread(b, 1, 4, fp) != 4) return 1;
    hl = (size_t)b[0] | ((size_t)b[1] << 8) | ((size_t)b[2] << 16) | ((size_t)b[3] << 24);
    *offset = 12 + hl;
  }
  char * h = calloc(hl + 1, 1);
  if(!h) return 1;
  if(fread(h, 1, hl, fp) != hl){ free(h); return 1; }
  int rc = 1;
  const char * p = strstr(h, "'descr'");
  const char * f = strstr(h, "'fortran_order'");
  const char * s = strstr(h, "'shape'");
  if(p && f && s){
    p = strchr(p + 7, ':');
    while(p && (*p == ':' || *p == ' ')) p++;
    if(p && (*p == '\'' || *p == '"')){
      const char q = *p;
      const char * e = strchr(p + 1, q);
      if(e && (size_t)(e - p + 1) < 32){
        memcpy(dtype, p, (size_t)(e - p + 1));
        dtype[e - p + 1] = '\0';
        dtype[0] = '\'';
        dtype[e - p] = '\'';
        f = strchr(f + 15, ':');
        while(f && (*f == ':' || *f == ' ')) f++;
        *fortran = (f && strncmp(f, "True", 4) == 0);
        s = strchr(s + 7, '(');
        if(s){
          s++;
          *ndims = 0;
          rc = 0;
          while(*s && *s != ')'){
            while(*s == ' ' || *s == ',') s++;
            if(*s == ')') break;
            char * end;
            const unsigned long long v = strtoull(s, &end, 10);
            if(end == s || *ndims >= 8){ rc = 1; break; }
            shape[(*ndims)++] = (size_t)v;
            s = end;
          }
        }
      }
    }
  }
  free(h);
  return rc;
}

static int fio_r_serial(const char dirname[], const char dsetname[], const size_t ndims,
    const size_t * shape, const char dtype[], const size_t size, void * data){
  char * fname = npy_path(dirname, dsetname);
  if(!fname) return 1;
  FILE * fp = fio_fopen(fname, "rb");
  if(!fp){ free(fname); return 1; }
  size_t nd = 0, shp[8], off = 0;
  char dt[32];
  int fortran = 0;
  int rc = read_header(fp, &nd, shp, dt, &fortran, &off);
  if(rc) fprintf(stderr, "%s: NPY header read failed\n", fname);
  if(!rc && nd != ndims){ fprintf(stderr, "%s: ndims %zu expected, %zu obtained\n", fname, ndims, nd); rc = 1; }
  for(size_t d = 0; !rc && d < ndims; d++){
    if(shp[d] != shape[d]){ fprintf(stderr, "%s: shape[%zu] %zu expected, %zu obtained\n", fname, d, shape[d], shp[d]); rc = 1; }
  }
  if(!rc && strcmp(dt, dtype) != 0){ fprintf(stderr, "%s: dtype %s expected, %s obtained\n", fname, dtype, dt); rc = 1; }
  if(!rc && fortran){ fprintf(stderr, "%s: fortran order is not supported\n", fname); rc = 1; }
  if(!rc){
    size_t n = 1;
    for(size_t d = 0; d < ndims; d++) n *= shape[d];
    if(fseek(fp, (long)off, SEEK_SET) != 0 || fread(data, size, n, fp) != n){
      fprintf(stderr, "%s: data read failed\n", fname);
      rc = 1;
    }
  }
  fclose(fp);
  free(fname);
  return rc;
}

static int fio_w_serial(const char dirname[], const char dsetname[], const size_t ndims,
    const size_t * shape, const char dtype[], const size_t size, const void * data){
  char * fname = npy_path(dirname, dsetname);
  if(!fname) return 1;
  FILE * fp = fio_fopen(fname, "wb");
  if(!fp){ free(fname); return 1; }
  int rc = write_header(fp, ndims, shape, dtype);
  size_t n = 1;
  for(size_t d = 0; d < ndims; d++) n *= shape[d];
  if(!rc && fwrite(data, size, n, fp) != n) rc = 1;
  if(rc) fprintf(stderr, "%s: write failed\n", fname);
  fclose(fp);
  free(fname);
  return rc;
}

const fileio_t fileio = {
  .npy_boolean = NPY_BOOLEAN,
  .npy_size_t = NPY_SIZE_T,
  .npy_double = NPY_DOUBLE,
  .init = fio_init,
  .fopen = fio_fopen,
  .fclose = fio_fclose,
  .mkdir = fio_mkdir,
  .r_serial = fio_r_serial,
  .w_serial = fio_w_serial,
};

/* ---------------------------------------------------------------------------------------
 * config (src/config.c:14-42)
 * ------------------------------------------------------------------------------------- */
static int get_double(const char dsetname[], double * value){
  const char * s = getenv(dsetname);
  if(!s){
    printf("%s not found\n", dsetname);
    return 1;
  }
  errno = 0;
  char * end = NULL;
  const double v = strtod(s, &end);
  if(errno != 0 || end == s){
    printf("%s: invalid value as double\n", dsetname);
    return 1;
  }
  *value = v;
  return 0;
}

const config_t config = {
  .get_double = get_double,
};

/* ---------------------------------------------------------------------------------------
 * param (src/param/implicit.c:4-6, src/param/boundary-condition.c:3-7)
 * ------------------------------------------------------------------------------------- */
bool param_implicit_x = true;
bool param_implicit_y = false;
bool param_implicit_z = false;
double param_uy_xm = 1.;
double param_uy_xp = 0.;
double param_t_xm = + 0.5;
double param_t_xp = - 0.5;
bool param_scalar = true;

static void env_bool(const char * name, bool * v){
  const char * s = getenv(name);
  if(!s) return;
  *v = !(strcmp(s, "0") == 0 || strcmp(s, "false") == 0 || strcmp(s, "False") == 0);
}

static void env_double(const char * name, double * v){
  const char * s = getenv(name);
  if(!s) return;
  char * end = NULL;
  const double x = strtod(s, &end);
  if(end != s) *v = x;
}

int tc_param_load(void){
  env_bool("implicit_x", &param_implicit_x);
  env_bool("implicit_y", &param_implicit_y);
  env_bool("implicit_z", &param_implicit_z);
  env_bool("scalar", &param_scalar);
  env_double("uy_xm", &param_uy_xm);
  env_double("uy_xp", &param_uy_xp);
  env_double("t_xm", &param_t_xm);
  env_double("t_xp", &param_t_xp);
  return 0;
}

/* src/runge_kutta.c:3-15 */
const uint_fast8_t rk_a = 0;
const uint_fast8_t rk_b = 1;
const uint_fast8_t rk_g = 2;
const rkcoef_t rkcoefs[RKSTEPMAX] = {
  {+ 32. / 60.,   0. / 60., + 32. / 60. +   0. / 60.},
  {+ 25. / 60., - 17. / 60., + 25. / 60. - 17. / 60.},
  {+ 45. / 60., - 25. / 60., + 45. / 60. - 25. / 60.},
};
