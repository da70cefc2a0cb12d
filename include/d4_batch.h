/*
 * d4_batch.h -- batched ragged ensembles of independent realizations.
 *
 * No counterpart in the reference, which runs one realization per process
 * (main.cpp, Simulation::run).  Realization r of R is the slice
 * [off[r], off[r+1]) of concatenated arrays; within its slice every pass has
 * exactly the per-realization meaning of include/d4.h (simulation.cpp:287-333
 * speed pass, 345-527 Jacobian pass, 546-608 rescale), and the results equal
 * those of the single-realization entry points bit for bit: one launch covers
 * the tiles of all realizations, and each tile's partial sums land in that
 * realization's own slots, reduced in the same fixed order.
 */
#ifndef D4_BATCH_H
#define D4_BATCH_H

#include <stdint.h>

#include "d4.h"
#include "d4_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Launch shape of the batched passes. */
typedef struct {
    int pass_grid;     /* 0: one workgroup (one wavefront) per 64x64 tile; > 0: that many
                          persistent workgroups striding over the tile table */
    int tile_order;    /* 0: realization-major; 1: largest realizations first;
                          2: round-robin over realizations */
    int use_graph;     /* 1: each batched pass is replayed from a captured hipGraph */
    int host_threads;  /* workers for the per-realization host work (0: all cores, at most 16) */
} d4_ens_config;

/* Batched speed pass (host arrays): xy concatenated sorted positions ({x, y}
 * per dislocation), tau[r] and tol_cap2[r] per realization (tol may be null). */
int d4_batch_speed_pass(d4_ctx* ctx, int R, const int* off, const double* xy, const double* tau, double* v,
                        double* tol, const double* tol_cap2);

/* Batched Jacobian pass (host arrays), p[r] per realization.  The step matrix
 * of realization r comes back as its own CSC block of a block-diagonal matrix
 * with global indices: columns off[r]..off[r+1]-1 of Ap (total+1 entries),
 * rows in the same range.  cap, nnz, D4_ERR_CAPACITY as d4_jacobian_pass. */
int d4_batch_jacobian_pass(d4_ctx* ctx, int R, const int* off, const double* xy, const d4_jac_params* p,
                           double* vA, double* vB, int* Ap, int* Ai, double* Ax, int cap, int* nnz, int* indexes,
                           double* dvec, double* tol);

typedef struct d4_ens d4_ens;

/* An ensemble of R host simulations (include/d4_sim.h semantics each), stepped
 * in lockstep: every pair pass of a step runs once for all realizations, the
 * per-realization host work (sparse LU, solves, step control) runs on a thread
 * pool.  x, y, b concatenated as off gives them; p holds R parameter sets when
 * params_per_realization != 0, else one for all. */
int d4_ens_create(d4_ctx* ctx, int R, const int* off, const double* x, const double* y, const double* b,
                  const d4_sim_params* p, int params_per_realization, d4_ens** out);
void d4_ens_destroy(d4_ens* e);
const char* d4_ens_last_error(const d4_ens* e);
/* nsteps lockstep iterations; a realization whose interrupt() holds sits out.
 * *nactive (optional): realizations that took part in the last iteration. */
int d4_ens_step(d4_ens* e, int nsteps, int* nactive);
/* concatenated positions: sorted, or each realization in its input order */
int d4_ens_positions(const d4_ens* e, double* xy, int original_order);
int d4_ens_stats(const d4_ens* e, d4_sim_stats* st); /* R entries */
int d4_ens_set_config(d4_ens* e, const d4_ens_config* c);
int d4_ens_get_config(const d4_ens* e, d4_ens_config* c);
/* Times the candidate launch shapes on the batched passes of the current
 * configuration (positions are not changed) and keeps the fastest. */
int d4_ens_autotune(d4_ens* e, d4_ens_config* best, double* best_ms);
/* accumulated wall time (ms) since creation: [0] batched pair passes incl.
 * transfers, [1] linear solves (factorizations and solves, batched on the GPU
 * for realizations with solver 1), [2] per-realization host work, [3] all of
 * d4_ens_step */
int d4_ens_timing(const d4_ens* e, double* ms4);

#ifdef __cplusplus
}
#endif
#endif
