/* integrate.h -- as include/integrate.h:1-11 of the reference */
#if !defined(INTEGRATE_H)
#define INTEGRATE_H

#include "domain.h"
#include "fluid.h"

/* replaces integrate, src/integrate.c:9-35: decide_dt and three RK substeps on the device */
extern int integrate(const domain_t * domain, fluid_t * fluid, double * dt);

#endif
