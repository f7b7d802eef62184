/* save.h -- checkpoints, as include/save.h:1-28 of the reference */
#if !defined(SAVE_H)
#define SAVE_H

#include <stddef.h>
#include "domain.h"
#include "fluid.h"

typedef struct save_t_ {
  /* replaces init, src/save.c:31-65 (save_rate, save_after from the environment) */
  int (* const init)(const domain_t * domain, const double time);
  /* replaces output, src/save.c:74-110: output/save/stepNNNNNNNNNN/ with step, time, the
   * domain_save items and the fluid_save items, the format of an initial-condition directory */
  int (* const output)(const domain_t * domain, const size_t step, const double time, const fluid_t * fluid);
  /* replaces get_next_time, src/save.c:112-116 */
  double (* const get_next_time)(void);
} save_t;

extern const save_t save;

#endif
