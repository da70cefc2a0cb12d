/*
 * d4_sim.h -- C ABI of the host simulation around lib2d4's pair passes: the
 * reference's SimulationData + Simulation::run (simulation_data.cpp,
 * simulation.cpp:30-281) with every O(N^2) pass on the GPU through include/d4.h
 * and the linear solves on the host (a sparse LU in place of UMFPACK,
 * simulation.cpp:635-659).
 *
 * Parameters mirror the reference's command-line flags (project_parser.cpp:28-68)
 * as SimulationData stores them (simulation_data.h:126-328).
 */
#ifndef D4_SIM_H
#define D4_SIM_H

#include <stdint.h>

#include "d4.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct d4_sim d4_sim;

typedef struct {
    double init_step_size;   /* --initial-stepsize (default 1e-6) */
    double max_step_size;    /* -M; <= 0: no limit */
    double prec;             /* -P position precision (default 1e-5) */
    double dipole_prec;      /* -p dipole precision (default 0.05; 0 disables) */
    double cutoff_mult;      /* -u cutoff multiplier (default 1e20) */
    int heaviside;           /* -h */
    char wfunc;              /* -w weight function 'c' | 'p' | 'm' */
    int stress_kind;         /* 0 StressProtocol (always 0), 1 FixedRate, 2 Cyclic, 3 IncreasingCyclic */
    double stress_init, stress_rate, stress_period, stress_amp;
    double time_limit;       /* -t; <= 0: none */
    int64_t step_limit;      /* -C; < 0: none (stop once successful steps exceed it) */
    double strain_limit;     /* <= 0: none */
    double speed_threshold;  /* cutoff change threshold (simulation.cpp:215-219); <= 0: none */
    int solver;              /* linear solves (UMFPACK's place, simulation.cpp:635-659):
                                0 host sparse LU (left-looking, minimum-degree order, threshold
                                  pivoting u = 0.1 preferring the diagonal);
                                1 dense LU on the GPU: the matrix in reverse Cuthill-McKee order
                                  where that halves the work (kept while the pattern's envelope in it
                                  grows by at most 1/8), padded to a multiple of 64, factored
                                  first by a blocked LU without row interchanges inside its block
                                  envelope; kept when every multiplier is at most 1000 (threshold
                                  u = 0.001, the diagonal preferred), otherwise factored again with
                                  partial pivoting by rocSOLVER getrf; solves by this library's
                                  triangular-solve kernel */
} d4_sim_params;

typedef struct {
    double sim_time, step_size, strain, last_max_error_ratio_sq;
    int64_t successful_steps, failed_steps;
    int64_t pair_passes, factorizations, singular_solves;
    int64_t unpivoted_factorizations; /* of factorizations: those the GPU dense LU finished
                                         without row interchanges (every multiplier <= 1000) */
    double pass_ms, solve_ms;         /* wall time the caller spent in pair passes (with their
                                         transfers) and in linear solves, in ms */
} d4_sim_stats;

/* Fills p with the reference's defaults (project_parser.cpp:44-50, 64). */
void d4_sim_default_params(d4_sim_params* p);

/* SimulationData from n dislocations (x, y, b = +-1) in any order: sorted and
 * numbered as simulation_data.cpp:76-154 does (sum b must be 0). */
int d4_sim_create(d4_ctx* ctx, int n, const double* x, const double* y, const double* b, const d4_sim_params* p,
                  d4_sim** out);
/* the same from a .dconf file (x y b per line, simulation_data.cpp:76-113) */
int d4_sim_create_from_file(d4_ctx* ctx, const char* path, const d4_sim_params* p, d4_sim** out);
void d4_sim_destroy(d4_sim* s);
const char* d4_sim_last_error(const d4_sim* s);

/* Simulation::run (simulation.cpp:30-281) until interrupt() (simulation.cpp:722-750) */
int d4_sim_run(d4_sim* s);
/* one iteration of run()'s loop (stages I-IV); *accepted = 1 if the step was taken;
 * *stop = 1 when interrupt() holds before the step (nothing done then) */
int d4_sim_step(d4_sim* s, int* accepted, int* stop);

/* sorted positions (n {x, y}), or in the input order (disl_order, simulation_data.cpp:118-123) */
int d4_sim_positions(const d4_sim* s, double* xy, int original_order);
int d4_sim_get_stats(const d4_sim* s, d4_sim_stats* st);
/* writeDislocationDataToFile (simulation_data.cpp:177-191) */
int d4_sim_write_dconf(const d4_sim* s, const char* path);

/* The linear solves of solver 1 (UMFPACK's place, simulation.cpp:635-659) on
 * K independent systems at once: matrix k is n[k] x n[k] in CSC form (Ap[k]:
 * n[k] + 1 column pointers from 0, Ai[k] rows in [0, n[k]), Ax[k] values,
 * duplicates not allowed), b[k] its right-hand side, x[k] its solution (may be
 * b[k]).  Dense LU on the GPU with threshold partial pivoting (u = 0.001, the
 * diagonal preferred, as UMFPACK's symmetric strategy), in a symmetric
 * bandwidth-reducing order (reverse Cuthill-McKee, taken from the pattern
 * when it halves the work) whose zero blocks outside the envelope are
 * skipped; singular[k] = 1 when the factorization met a zero
 * pivot (x[k] then holds what the solve gives, as with UMFPACK's warning),
 * unpivoted[k] = 1 when no rows were interchanged; either may be null.  A
 * system's solution does not depend on the other systems of the call. */
int d4_lu_solve_batch(d4_ctx* ctx, int K, const int* n, const int* const* Ap, const int* const* Ai,
                      const double* const* Ax, const double* const* b, double* const* x, unsigned char* singular,
                      unsigned char* unpivoted);

/* The same solver on one dense n x n matrix already on the device: element (r, c) at
 * a[r * row_stride + c * col_stride] (a row-major array: n, 1; its transpose: 1, n).  No reference
 * counterpart: the adjoint's Newton steps use it.  The factorization replaces the context's previous one
 * (of this call or of d4_lu_solve_batch); *singular and *unpivoted as there (either may be null).  Both
 * calls synchronize the device first (the inputs may come from any stream) and return with the results
 * complete. */
int d4_lu_factor_dense_dev(d4_ctx* ctx, int n, const double* a, long long row_stride, long long col_stride,
                           int* singular, int* unpivoted);
/* x = A^-1 b for the matrix of the last d4_lu_factor_dense_dev; b and x on the device (may be the same) */
int d4_lu_solve_dense_dev(d4_ctx* ctx, int n, const double* b, double* x);

#ifdef __cplusplus
}
#endif
#endif
