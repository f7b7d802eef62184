/*
 * fileio.h -- serial members of the reference's fileio_t (include/fileio.h:7-47): NPY v1.0
 * read and write by one process.  The MPI-IO members (r_nd_parallel, w_nd_parallel) have no
 * counterpart on one GPU.
 */
#if !defined(FILEIO_H)
#define FILEIO_H

#include <stdio.h>
#include <stddef.h>

typedef struct {
  const char * npy_boolean;   /* "'|b1'"  (fileio.c:20) */
  const char * npy_size_t;    /* "'<u8'"  (fileio.c:21) */
  const char * npy_double;    /* "'<f8'"  (fileio.c:22) */
  /* replaces init, fileio.c:108-130 */
  int (* const init)(void);
  /* replace fopen_/fclose_/mkdir_, fileio.c:132-165 */
  FILE * (* const fopen)(const char * path, const char * mode);
  int (* const fclose)(FILE * stream);
  int (* const mkdir)(const char dirname[]);
  /* replaces r_serial, fileio.c:269-310: header checked against ndims, shape and dtype */
  int (* const r_serial)(const char dirname[], const char dsetname[], const size_t ndims,
      const size_t * shape, const char dtype[], const size_t size, void * data);
  /* replaces w_serial, fileio.c:322-362 */
  int (* const w_serial)(const char dirname[], const char dsetname[], const size_t ndims,
      const size_t * shape, const char dtype[], const size_t size, const void * data);
} fileio_t;

extern const fileio_t fileio;

#endif
