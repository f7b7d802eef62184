/* runge_kutta.h -- as include/runge_kutta.h:1-17 of the reference */
#if !defined(RUNGE_KUTTA_H)
#define RUNGE_KUTTA_H

#include <stdint.h>

extern const uint_fast8_t rk_a;
extern const uint_fast8_t rk_b;
extern const uint_fast8_t rk_g;
typedef double rkcoef_t[3];
#define RKSTEPMAX 3
/* replaces rkcoefs, src/runge_kutta.c:8-15 (Wray's low-storage RK3) */
extern const rkcoef_t rkcoefs[RKSTEPMAX];

#endif
