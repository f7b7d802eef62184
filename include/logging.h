/*
 * logging.h -- the reference's logging (include/logging.h:1-30, src/logging/*.c).  The sums
 * are FP64 reductions on the device; files and formats are the reference's
 * (output/log/{progress,divergence,injection,dissipation,total_energy}.dat).
 */
#if !defined(LOGGING_H)
#define LOGGING_H

#include <stddef.h>
#include "domain.h"
#include "fluid.h"

typedef struct {
  /* replaces init, src/logging/main.c:20-40 (log_rate from the environment) */
  int (* const init)(const domain_t * domain, const double time);
  /* replaces check_and_output, src/logging/main.c:87-101 */
  void (* const check_and_output)(const domain_t * domain, const size_t step, const double time,
      const double dt, const double wtime, const fluid_t * fluid);
  /* replaces get_next_time, src/logging/main.c:107-111 */
  double (* const get_next_time)(void);
} logging_t;

extern const logging_t logging;

/* replace src/logging/internal.h:4-30 (appending one line to fname) */
extern int logging_check_divergence(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid);
extern int logging_check_injection(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid);
extern int logging_check_dissipation(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid);
extern int logging_check_total_energy(const char fname[], const domain_t * domain, const double time, const fluid_t * fluid);

#endif
