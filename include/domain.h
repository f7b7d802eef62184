/*
 * domain.h -- domain_t and its constructor, as include/domain.h:1-49 of the reference.
 * One GPU, no MPI: `info` is kept for layout compatibility and is always NULL; mysizes equal
 * glsizes and offsets are zero.
 */
#if !defined(DOMAIN_H)
#define DOMAIN_H

#include <stdbool.h>
#include <stddef.h>

#if !defined(NDIMS)
#define NDIMS 3
#endif

typedef struct {
  void * info;                         /* sdecomp_info_t * in the reference; unused */
  bool is_curved;
  size_t glsizes[NDIMS];
  size_t mysizes[NDIMS];
  size_t offsets[NDIMS];
  double lengths[NDIMS];
  double * restrict xf, * restrict xc;
  double * restrict jdxf, * restrict jdxc;
  double * restrict hxxf, * restrict hxxc;
  double * restrict hyxf, * restrict hyxc;
  double hz;
} domain_t;

/* replaces domain_init, src/domain.c:389-432: loads is_curved, glsizes, lengths, xf, xc
 * (NPY files in dirname_ic, domain.c:28-55) and builds the metrics (domain.c:259-357) */
extern int domain_init(const char dirname_ic[], domain_t * domain);

/* replaces domain_save, src/domain.c:63-86 */
extern int domain_save(const char dirname[], const domain_t * domain);

#endif
