/*
 * C ABI of the hourly step on the host CPU: the same per-cell source as the HIP kernels
 * (ibis_step_cell.hpp), built with g++ and run over host threads.  It is the build's
 * C++ CPU path for the 1,024-cell benchmark case, an explicit entry point of its own: the
 * GPU entry points never fall back to it.
 *
 * Replaces timestep_body (python-port/ibis/timestepping/loops.py:234-1284) run on the CPU
 * (the reference's JAX-CPU path).
 */
#ifndef IBIS_CPU_H
#define IBIS_CPU_H

#include <stdint.h>
#include "ibis_step.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One hour for `ncell` cells on `nthreads` host threads (<= 0: the hardware concurrency).
 * Host pointers; arenas, slot types, table, flags and error codes as ibis_step (out may equal
 * in).  The phase workspace is allocated and freed inside the call. */
int ibis_step_cpu(const IbisArena* in, const IbisArena* out, const int8_t* slot_type,
                  int64_t slot_pitch, const double* type_table, const IbisParams* params,
                  double dtime, double nytimes, int64_t ncell, int flags, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* IBIS_CPU_H */
