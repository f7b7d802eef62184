/* config.h -- as include/config.h:1-14 of the reference */
#if !defined(CONFIG_H)
#define CONFIG_H

typedef struct {
  /* replaces get_double, src/config.c:14-42: environment variable as a double, non-zero
   * return when it is unset or does not parse */
  int (* const get_double)(const char dsetname[], double * value);
} config_t;

extern const config_t config;

#endif
