/* kz_implicit.h — device-resident batches of independent implicit solves
 * (libkz_hip.so).
 *
 * A kz_implicit_batch holds n independent cells — primitives, geometry and
 * a step size each — on the device and runs the reference's radiation–gas
 * four-force solve, Solver.solveImplicitLab (koral/solve/implicit.zig:
 * 927-1009), on all of them: the same split, compacted kernels that serve
 * Sim.opImplicit (koral/sim.zig:1736-1776) in kz_step.h, so every cell's
 * outcome, rung, iteration count and solved state equal the per-record
 * kz_solve_implicit_lab_batch (kz_cellops.h) bit for bit.  Inputs stay on
 * the device between solves, so repeated solves time the kernels alone.
 *
 * Return codes are the KZ_OK / KZ_E_* values of kz_types.h.
 */
#ifndef KZ_IMPLICIT_H
#define KZ_IMPLICIT_H

#include <stdint.h>
#include "kz_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kz_implicit_batch kz_implicit_batch;

/* Room for n cells under the physics parameters of `phys` (gam, rad, opac,
 * implicit; the opacities must be on). */
kz_implicit_batch *kz_implicit_batch_create(int64_t n, const kz_physics *phys, int *err);
void kz_implicit_batch_destroy(kz_implicit_batch *b);

/* Load the inputs (host memory, synchronous): pp 13 doubles per cell (slot
 * order of kz_types.h), kz_geom per cell (koral/geometry.zig:29-49; g must
 * be symmetric bit for bit), dt per cell.  KZ_E_INVALID_CONFIG for an
 * asymmetric metric. */
int kz_implicit_batch_set(kz_implicit_batch *b, const double *pp, const kz_geom *g, const double *dt);

/* Solve every cell from the loaded inputs (solveImplicitLab: FORCEUEQPIN
 * p2u, bisection warm start, the 6-rung ladder under opacity damping).
 * Synchronous; *ms (nullable) gets the solve's kernel time from HIP events. */
int kz_implicit_batch_solve(kz_implicit_batch *b, double *ms);

/* Results of the last solve (nullable outputs): uu and pp 13 doubles per
 * cell (a failed cell keeps its input pp and a zero uu), the ladder's
 * outcome per cell, counts[3] = failures, Newton iterations, solves. */
int kz_implicit_batch_get(kz_implicit_batch *b, double *uu, double *pp, kz_implicit_result *r, uint64_t *counts);

/* No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  kz_implicit_batch_solve over the
 * caller's device arrays, with every working buffer (geometry planes, SoA
 * states, the Newton reference states, the compacted lists, the counters)
 * carved out of one static device array of this library, about 24 MB per
 * GPU, held until the process ends.  Calls no hipMalloc, hipFree or other
 * allocating routine; enqueues on `stream` (a hipStream_t; NULL: the null
 * stream) and returns without waiting.
 *
 * All pointers are device memory: pp_in 13 doubles per cell and g one
 * kz_geom per cell (only g's upper triangle is read: the metric must be
 * symmetric, as kz_implicit_batch_set requires), dt per cell; outputs uu and
 * pp 13 doubles per cell (nullable; a failed cell gets its input pp and a
 * zero uu), r per cell, counts[3] = failures, Newton iterations, solves
 * (nullable).  Results equal kz_implicit_batch_solve's bit for bit.
 * n above KZ_NOALLOC_MAX_CELLS is rejected with KZ_E_INVALID_CONFIG.  Calls
 * that overlap on one GPU share its scratch and race: serialise them. */
#define KZ_NOALLOC_MAX_CELLS 32768
int kz_implicit_solve_noalloc(int64_t n, const double *pp_in, const kz_geom *g, const double *dt, double *uu,
                              double *pp, kz_implicit_result *r, uint64_t *counts, const kz_physics *phys,
                              void *stream);

#ifdef __cplusplus
}
#endif

#endif /* KZ_IMPLICIT_H */
