/*
 * array.h -- array_t and its methods, as include/array.h:1-57 of the reference.
 */
#if !defined(ARRAY_H)
#define ARRAY_H

#include <stddef.h>
#include "domain.h"

typedef struct {
  size_t size;          /* bytes per element */
  int (* nadds)[2];     /* extra cells per dimension: [dim][lower, upper] */
  size_t datasize;      /* bytes of the local array */
  void * data;          /* host memory */
} array_t;

typedef struct {
  /* replaces prepare, src/array.c:10-46 */
  int (* const prepare)(const domain_t * domain, const int nadds[NDIMS][2], const size_t size, array_t * array);
  /* replaces destroy, src/array.c:48-54 */
  int (* const destroy)(array_t * array);
  /* replaces load, src/array.c:82-153 (serial NPY read of the x-inclusive layout) */
  int (* const load)(const domain_t * domain, const char dirname[], const char dsetname[], const char dtype[], array_t * array);
  /* replaces dump, src/array.c:155-222 */
  int (* const dump)(const domain_t * domain, const char dirname[], const char dsetname[], const char dtype[], const array_t * array);
} array_method_t;

extern const array_method_t array;

#endif
