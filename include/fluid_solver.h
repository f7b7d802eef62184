/*
 * fluid_solver.h -- the fluid API of the reference (include/fluid_solver.h:1-94), served by
 * the MI355X library.  Each call runs on the device state bound to `fluid` by fluid_init.
 *
 * Host coherence: by default (environment TC_COHERENCE unset or "strict") every call below
 * uploads the host fields ux uy uz p t before it runs and downloads ux uy uz p psi t after,
 * so the host arrays behave exactly as the reference's.  With TC_COHERENCE=lazy the fields
 * stay on the device between calls; tc_fluid_sync_host / tc_fluid_sync_device move them.
 * fluid_save and the logging functions synchronise by themselves.
 */
#if !defined(FLUID_SOLVER_H)
#define FLUID_SOLVER_H

#include "array.h"
#include "domain.h"
#include "fluid.h"

/* replaces fluid_init, src/fluid/init.c:98-122 (+ device state of the step) */
extern int fluid_init(const char dirname_ic[], const domain_t * domain, fluid_t * fluid);

/* replaces fluid_save, src/fluid/save.c:8-28 */
extern int fluid_save(const char dirname[], const domain_t * domain, const fluid_t * fluid);

/* replace src/fluid/compute_diffusivity.c:3-13 */
extern double fluid_compute_momentum_diffusivity(const fluid_t * fluid);
extern double fluid_compute_scalar_diffusivity(const fluid_t * fluid);

/* replaces fluid_predict_field, src/fluid/predict/main.c:35-110 */
extern int fluid_predict_field(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid);

/* replaces fluid_compute_potential, src/fluid/compute_potential.c:499-583 */
extern int fluid_compute_potential(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid);

/* replaces fluid_correct_velocity, src/fluid/correct/main.c:15-28 */
extern int fluid_correct_velocity(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid);

/* replaces fluid_update_pressure, src/fluid/update_pressure.c:146-173 */
extern int fluid_update_pressure(const domain_t * domain, const size_t rkstep, const double dt, fluid_t * fluid);

/* replace src/fluid/boundary/{ux,uy,uz,p,psi,t}.c (host arrays) */
extern int fluid_update_boundaries_ux(const domain_t * domain, array_t * array);
extern int fluid_update_boundaries_uy(const domain_t * domain, array_t * array);
extern int fluid_update_boundaries_uz(const domain_t * domain, array_t * array);
extern int fluid_update_boundaries_p(const domain_t * domain, array_t * array);
extern int fluid_update_boundaries_psi(const domain_t * domain, array_t * array);
extern int fluid_update_boundaries_t(const domain_t * domain, array_t * array);

/* no reference counterpart (the reference's fields always live on the host): move the
 * fields between the host arrays and the device state */
extern int tc_fluid_sync_host(fluid_t * fluid);
extern int tc_fluid_sync_device(const fluid_t * fluid);
/* no reference counterpart: release the device state bound to fluid */
extern int tc_fluid_finalise(fluid_t * fluid);

#endif
