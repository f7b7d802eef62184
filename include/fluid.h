/*
 * fluid.h -- fluid_t, as include/fluid.h:1-38 of the reference.
 * The host arrays ux uy uz p psi t keep the reference's layouts (include/array_macros/fluid/*.h).
 * The RK source terms and the velocity-gradient arrays of the reference live on the device
 * in this build (they are intermediate values of the step); their array_t members are kept
 * for layout compatibility and hold no host data (data == NULL, datasize == 0).
 */
#if !defined(FLUID_H)
#define FLUID_H

#include "array.h"
#include "domain.h"

typedef struct {
  array_t ux;
  array_t uy;
  array_t uz;
  array_t p;
  array_t psi;
  array_t t;
  array_t srcux[3];
  array_t srcuy[3];
  array_t srcuz[3];
  array_t srct[3];
  array_t lxx;
  array_t lyx0, lyx1, lxy;
  array_t lzx, lxz;
  array_t lyy0, lyy1;
  array_t lzy, lyz;
  array_t lzz;
  double Re, Pr;
} fluid_t;

#endif
