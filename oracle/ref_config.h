/*
 * ORACLE -- test infrastructure only.
 *
 * Build configuration for compiling the reference sources in place.  The
 * reference selects single-process mode by rewriting one line of its
 * src/gcge_config.h ("make without-mpi", reference Makefile:34-35), which a
 * read-only tree does not allow; this header is force-included (-include)
 * ahead of it and defines the same four settings with GCGE_USE_MPI = 0.  Its
 * include guard is the one gcge_config.h uses, so that header then adds
 * nothing.
 */
#ifndef _GCGE_CONFIG_H_
#define _GCGE_CONFIG_H_
#include <float.h>
#define GCGE_DOUBLE double
#define GCGE_INT int
#define GET_PART_TIME 1
#define GCGE_USE_MPI 0
#endif
