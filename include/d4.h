/*
 * d4.h -- C ABI of lib2d4, the MI355X (gfx950) implementation of 2D4's
 * periodic pair-interaction path.
 *
 * Plain C types only (pointers, sizes, scalars); no HIP or torch types cross
 * this boundary.  Every call returns 0 on success or a negative D4_ERR_* code;
 * d4_last_error() gives the message.  Calls are synchronous with respect to
 * the host unless the name ends in _dev, in which case all array arguments are
 * device pointers and the work is ordered on the context's stream
 * (d4_ctx_set_stream); the call returns once the work is enqueued.
 *
 * Positions follow the reference's DislwoB layout (include/dislocation.h:23-34):
 * n records of {double x, double y}, sorted as SimulationData sorts them
 * (simulation_data.cpp:116), the Burgers vector implied by the index: +1 for
 * the first n/2, -1 for the rest (simulation_data.cpp:265-270).
 *
 * Tolerance arrays implement PrecisionHandler::updateTolerance
 * (precision_handler.cpp:34-38): for every pair, tol[i] = min(tol[i], d^2) and
 * tol[j] likewise when d^2 < tol_cap2; tol_cap2 <= 0 (or tol == NULL) skips it.
 */
#ifndef D4_H
#define D4_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D4_OK 0
#define D4_ERR_INVALID (-1)
#define D4_ERR_HIP (-2)
#define D4_ERR_CAPACITY (-3)
#define D4_ERR_NO_DEVICE (-4)

typedef struct d4_ctx d4_ctx;

/* Context: device, stream and the device buffers it owns (no reference
 * counterpart: the reference is single-threaded host code; one context per
 * host thread). */
int d4_ctx_create(int device, d4_ctx** out);
void d4_ctx_destroy(d4_ctx* ctx);
const char* d4_last_error(const d4_ctx* ctx);
/* the HIP device the context runs on (no reference counterpart) */
int d4_ctx_device(const d4_ctx* ctx);
/* hipStream_t passed as void*; NULL restores the context's own stream */
int d4_ctx_set_stream(d4_ctx* ctx, void* stream);
/* order the context's work on the device's legacy default (NULL) stream, the
 * one a caller's default-stream work (torch's default stream) runs on */
int d4_ctx_use_default_stream(d4_ctx* ctx);
int d4_ctx_synchronize(d4_ctx* ctx);
/* Kernel timing (no reference counterpart): when on, HIP events recorded on
 * the context's stream bracket every pair-kernel launch; d4_ctx_pair_kernel_time
 * waits for them, returns the summed milliseconds and launch count since the
 * last query, and clears the record. */
int d4_ctx_set_profiling(d4_ctx* ctx, int on);
int d4_ctx_pair_kernel_time(d4_ctx* ctx, double* total_ms, int* launches);
/* device memory helpers for callers without a HIP toolchain (ctypes, tests) */
int d4_dev_alloc(d4_ctx* ctx, int64_t bytes, void** out);
int d4_dev_free(d4_ctx* ctx, void* p);
int d4_memcpy_h2d(d4_ctx* ctx, void* dst, const void* src, int64_t bytes);
int d4_memcpy_d2h(d4_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* Per-pair field, replaces sdddstCore::Field::xy / xy_diff_x
 * (include/Fields/Field.h:40-49; src/Fields/Field.cpp:295-346).
 * For each k: sxy[k] = xy(dx[k], dy[k]) and dsxy[k] = xy_diff_x(dx[k], dy[k]).
 * When e2pix and c2piy are given they are used as the 4-arg overloads' inputs;
 * when NULL the device computes exp2(PI2LN2INV*dx) and cos(2*pi*dy)
 * (Field.cpp:335-336).  sxy or dsxy may be NULL.  flags bit 0: true derivative
 * instead of the reference's f_dx_h substitution (Field.cpp:269-283). */
int d4_field_eval(d4_ctx* ctx, int64_t n, const double* dx, const double* dy, const double* e2pix,
                  const double* c2piy, double* sxy, double* dsxy, int flags);
int d4_field_eval_dev(d4_ctx* ctx, int64_t n, const double* dx, const double* dy, const double* e2pix,
                      const double* c2piy, double* sxy, double* dsxy, int flags);

/* Speed pass, replaces Simulation::calculateSpeedsAtTime
 * (include/simulation.h:138; src/simulation.cpp:287-333):
 * v[i] = sum_{j != i} b_i b_j sigma(x_lo - x_hi, y_lo - y_hi) (+ for the lower
 * index, - for the higher) + tau b_i. */
int d4_speed_pass(d4_ctx* ctx, int n, const double* xy, double tau, double* v, double* tol, double tol_cap2);
int d4_speed_pass_dev(d4_ctx* ctx, int n, const double* xy, double tau, double* v, double* tol, double tol_cap2);

/* d4_speed_pass_dev that allocates nothing (no counterpart in the reference).  For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Inputs and outputs are the caller's device buffers, as for d4_speed_pass_dev, with the
 * same results bit for bit.  Its tile table and partial sums live in a scratch array the library holds
 * on every GPU it runs on (8.6 MB), so n is at most D4_NOALLOC_MAX_N (larger n: D4_ERR_INVALID).
 * Every call on one GPU uses that GPU's one scratch array: calls that overlap on a GPU (two streams,
 * two contexts or two threads) race and must be ordered by the caller. */
#define D4_NOALLOC_MAX_N 8192
int d4_speed_pass_dev_noalloc(d4_ctx* ctx, int n, const double* xy, double tau, double* v, double* tol,
                              double tol_cap2);

/* Parameters of a Jacobian pass (simulation.cpp:345-527 reads them from
 * SimulationData, simulation_data.h:181-197, and the stress protocol). */
typedef struct {
    double h;              /* stepsize */
    double tau_a, tau_b;   /* extStress(simTime_A), extStress(simTime_B) */
    double cutoff;         /* cutOff */
    double cut_sq;         /* cutOffSqr */
    double one_per_cut_sq; /* onePerCutOffSqr */
    int heaviside;         /* heavisideCutoff */
    char wfunc;            /* weightFunc: 'c', 'p' or 'm' */
    double tol_cap2;       /* PrecisionHandler minPrecisitySq if the dipole criterion is on, else 0 */
    int exact;             /* 0: the reference's d sigma (Field.cpp:314-331); 1: the true derivative
                              (no reference counterpart; for the adjoint's forward map) */
} d4_jac_params;

/* Jacobian pass, replaces Simulation::calcJacobianAndSpeedsAtTimes
 * (include/simulation.h:150; src/simulation.cpp:345-527) up to the UMFPACK
 * call: speeds vA (at tau_a) and vB (at tau_b), the CSC matrix
 * A = I + diag((1+d)/2) (V + diag(s)) with ascending rows per column and the
 * diagonal included (Ap n+1, Ai/Ax nnz), indexes[i] = position of A_ii in Ax,
 * dvec = weight(s).  Pairs are stored iff d^2 < (heaviside ? 1 : 49) cut_sq,
 * tapered by exp(-(d - rc)^2/rc^2) beyond rc when not heaviside.
 * Ai/Ax may be written up to cap entries (past nnz too).  If cap < nnz,
 * *nnz gets the required size and D4_ERR_CAPACITY is returned; the caller
 * grows Ai/Ax and calls again (the entries written are not the matrix). */
int d4_jacobian_pass(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                     int* Ap, int* Ai, double* Ax, int cap, int* nnz, int* indexes, double* dvec, double* tol);

/* Stage-II rescale, replaces Simulation::calcJacobianAndSpeedsFromPrev
 * (include/simulation.h:166; src/simulation.cpp:546-608) on a CSC matrix from
 * d4_jacobian_pass: factor = (new_t - base_t)/(old_t - base_t),
 * dtau = extStress(new_t) - extStress(old_t); Ax, dvec and speed in place. */
int d4_jacobian_rescale(d4_ctx* ctx, int n, const int* Ap, const int* Ai, double* Ax, double* dvec, char wfunc,
                        double factor, double dtau, double* speed);

/* Dense Jacobian pass (the reference's pass with an unlimited cutoff,
 * simulation.cpp:345-515 with cutOffSqr = infinity), device pointers:
 * M (column-major, leading dimension ld >= n) receives V + diag(s), i.e. the
 * reference's Ax before its (1+d)/2 row scaling and +1 (simulation.cpp:516-522);
 * M = -h dv/dx.  vA, vB, dvec, tol as d4_jacobian_pass. */
int d4_jacobian_dense_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                          double* M, int64_t ld, double* dvec, double* tol);
/* The reference's Jacobian pass with its cutoff (simulation.cpp:345-515) into dense storage: as
 * d4_jacobian_dense_dev, but V_ij is stored only for pairs inside the storage radius (d^2 < rc^2 with
 * the heaviside cutoff, d^2 < 49 rc^2 otherwise) and tapered by exp(-(r - rc)^2 / rc^2) beyond rc
 * (simulation.cpp:458-470), zero elsewhere: the reference's CSC Ax before its row scaling, densely. */
int d4_jacobian_dense_cut_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                              double* M, int64_t ld, double* dvec, double* tol);
/* simulation.cpp:516-522 on a dense M from d4_jacobian_dense_dev:
 * M[j,i] *= (1 + d_j)/2, M[i,i] += 1 -> the reference's final A. */
int d4_jacobian_dense_scale_dev(d4_ctx* ctx, int n, double* M, int64_t ld, const double* dvec);
/* host-pointer convenience form of the two calls above (scale != 0 applies the scaling) */
int d4_jacobian_dense(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                      double* M, double* dvec, double* tol, int scale);

/* Vector-Jacobian product of the pair-speed map (no reference counterpart; the
 * adjoint of calculateSpeedsAtTime, simulation.cpp:287-333), device pointers:
 * for L = sum_i w_i v_i:  gx = dL/dx, gy = dL/dy (gy may be NULL), *gtau = dL/dtau = sum b_i w_i.
 * flags bit 0: true derivative (else the reference's d sigma of Field.cpp:314-331,
 * whose transpose this then is). */
int d4_vjp_dev(d4_ctx* ctx, int n, const double* xy, const double* w, double* gx, double* gy, double* gtau,
               int flags);

/* The pair term of the vector-Jacobian product of the reference's implicit step (no reference
 * counterpart; the adjoint of how simulation.cpp:345-527's matrix entries
 * V_ij = k b_i b_j xy_diff_x(x_i - x_j, y_i - y_j) * taper depend on the positions, Field.cpp:314-331's
 * formula differentiated as the reference evaluates it).  p: h = the step's k, cutoff, heaviside.
 * With P_lj = -(a_l u_j + a_j u_l + c_l w_j + c_j w_l) - t_l - t_j over pairs inside the storage radius:
 *   gx_l = sum_{j > l} k b_l b_j dm/ddx (x_l - x_j, y_l - y_j) P_lj - sum_{j < l} k b_l b_j dm/ddx (x_j - x_l,
 *          y_j - y_l) P_lj  (each pair differentiated where the reference evaluates it, lower index first),
 *   gy_l the same with dm/ddy (gy may be NULL).  Device pointers, n entries each. */
int d4_theta_pair_vjp_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, const double* a,
                          const double* c, const double* u, const double* w, const double* t, double* gx, double* gy);

/* Tabulated field (no reference counterpart: the reference evaluates the
 * analytic Field::xy per pair).  A table is n x n doubles (n a power of two),
 * row-major in dy: tab[ky*n + kx] holds the value at (dx, dy) =
 * (-1/2 + kx/n, -1/2 + ky/n); between nodes the passes interpolate bilinearly
 * and periodically.  d4_table_build_dev fills tab_s with Field::xy and tab_d
 * (may be NULL) with Field::xy_diff_x (Field.cpp:295-346) at the nodes, 0 at
 * the origin.  Device pointers. */
int d4_table_build_dev(d4_ctx* ctx, int n, double* tab_s, double* tab_d);
/* bilinear table values for m pairs (dx, dy wrapped first); s or d may be NULL */
int d4_table_eval_dev(d4_ctx* ctx, int64_t m, const double* dx, const double* dy, const double* tab_s,
                      const double* tab_d, int n, double* s, double* d);
/* d4_speed_pass_dev with sigma from the table */
int d4_speed_pass_table_dev(d4_ctx* ctx, int n, const double* xy, const double* tab_s, int tab_n, double tau,
                            double* v);
/* d4_jacobian_dense_dev with sigma and d sigma from the tables (no tolerance tracking) */
int d4_jacobian_dense_table_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, const double* tab_s,
                                const double* tab_d, int tab_n, double* vA, double* vB, double* M, int64_t ld,
                                double* dvec);

#ifdef __cplusplus
}
#endif
#endif
