/*
 * d4_dstep.h -- the implicit step of Simulation::run (simulation.cpp:88-276) resident on the device: for
 * R independent realizations at once, every stage of an iteration -- the Jacobian and speed passes, the
 * step matrices, their LU factorizations, the six solves, the g vectors and position updates, the error
 * estimate, accept/reject, the strain, the step-size control and the interrupt() limits -- is a kernel
 * launch, and one iteration is captured once as a HIP graph and replayed.  The host reads results back
 * only when asked (and once per up to 16 replays to see whether any realization is still running).
 *
 * Same semantics as include/d4_sim.h's loop, with these differences of storage only:
 *  - the step matrix is dense (n x n, zero outside the reference's storage radius, simulation.cpp:461-470)
 *    instead of CSC;
 *  - its LU is the GPU dense LU (threshold partial pivoting u = 0.001: blocked LU without interchanges,
 *    refactored with partial pivoting on the device where a multiplier exceeds 1000) in an order fixed for
 *    the run: with a storage radius below half the cell and no speed-threshold cutoff change, rows sorted
 *    by |y| inside the band envelope that order fixes for every possible pattern (y never changes, and a
 *    matrix entry needs a pair closer than the radius); otherwise the natural order and the whole matrix.
 *    Results agree with the CSC path's to rounding, not bit for bit;
 *  - the run's initial speed pass (simulation.cpp:46), whose speeds stage I recomputes before any use,
 *    is not made.
 * d4_sim_params.solver is ignored.  The dense matrix and its LU store hold 2 x 8 n^2 bytes per
 * realization (25 GB for 256 realizations of up to 4096); the largest order is bounded by the device's LDS
 * (n <= ~10000).
 */
#ifndef D4_DSTEP_H
#define D4_DSTEP_H

#include "d4.h"
#include "d4_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct d4_dstep d4_dstep;

/* R realizations, realization r = the dislocations off[r] .. off[r+1]-1 of x, y, b (any order; sorted
 * and numbered as simulation_data.cpp:76-154 does), with p[r] (params_per_realization) or p for all. */
int d4_dstep_create(d4_ctx* ctx, int R, const int* off, const double* x, const double* y, const double* b,
                    const d4_sim_params* p, int params_per_realization, d4_dstep** out);
void d4_dstep_destroy(d4_dstep* s);
const char* d4_dstep_last_error(const d4_dstep* s);

/* use_graph = 1 (the default): iterations are replays of the captured graph; 0: the same launches
 * enqueued one by one (the results are the same bits) */
int d4_dstep_set_graph(d4_dstep* s, int use_graph);

/* `iterations` loop iterations (stages I-IV) of every realization that interrupt() does not stop
 * (simulation.cpp:722-750: time, strain and step-count limits); *nactive (may be null) = realizations
 * still running afterwards.  Returns once the work is complete. */
int d4_dstep_run(d4_dstep* s, int iterations, int* nactive);

/* positions of all realizations, concatenated: sorted order, or each realization's input order */
int d4_dstep_positions(const d4_dstep* s, double* xy, int original_order);
/* per realization (R entries): as d4_sim_get_stats; pass_ms and solve_ms are 0 */
int d4_dstep_stats(const d4_dstep* s, d4_sim_stats* st);

#ifdef __cplusplus
}
#endif
#endif
