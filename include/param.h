/*
 * param.h -- run-time parameters of the reference (include/param.h:1-27, src/param/*.c).
 * The reference fixes them at compile time; here they are variables with the reference's
 * values as defaults, overridable from the environment by tc_param_load (variables
 * implicit_x implicit_y implicit_z uy_xm uy_xp t_xm t_xp, and scalar = 0 to freeze T).
 */
#if !defined(PARAM_H)
#define PARAM_H

#include <stdbool.h>

/* replace param/implicit.c:4-6 (defaults true, false, false) */
extern bool param_implicit_x;
extern bool param_implicit_y;
extern bool param_implicit_z;
/* replace param/boundary-condition.c:3-7 (defaults 1, 0, +0.5, -0.5) */
extern double param_uy_xm;
extern double param_uy_xp;
extern double param_t_xm;
extern double param_t_xp;
/* no reference counterpart: 0 leaves T untouched by the step */
extern bool param_scalar;

/* no reference counterpart: read the overrides above from the environment */
extern int tc_param_load(void);

#endif
