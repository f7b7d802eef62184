/* decide_dt.h -- as include/decide_dt.h:1-11 of the reference */
#if !defined(DECIDE_DT_H)
#define DECIDE_DT_H

#include "domain.h"
#include "fluid.h"

/* replaces decide_dt, src/decide_dt.c:125-160 (coefficients coef_dt_adv, coef_dt_dif from
 * the environment, decide_dt.c:128-129; environment dt_fixed > 0 fixes dt, no reference
 * counterpart) */
extern int decide_dt(const domain_t * domain, const fluid_t * fluid, double * dt);

#endif
