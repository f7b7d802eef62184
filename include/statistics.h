/* statistics.h -- as include/statistics.h:1-17 of the reference */
#if !defined(STATISTICS_H)
#define STATISTICS_H

#include <stddef.h>
#include "domain.h"
#include "fluid.h"

typedef struct {
  /* replaces init, src/statistics.c:56-101 (stat_rate, stat_after from the environment) */
  int (* const init)(const domain_t * domain, const double time);
  /* replaces collect, src/statistics.c:236-255: one sample added on the device */
  int (* const collect)(const domain_t * domain, const fluid_t * fluid);
  /* replaces output, src/statistics.c:257-300: output/stat/stepNNNNNNNNNN/ with num, the
   * domain_save items and ux1 ux2 uy1 uy2 uz1 uz2 adv dif */
  int (* const output)(const domain_t * domain, const size_t step);
  /* replaces get_next_time, src/statistics.c:103-109 */
  double (* const get_next_time)(void);
} statistics_t;

extern const statistics_t statistics;

#endif
