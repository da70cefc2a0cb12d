/*
 * d4_oracle.h -- CPU restatement of 2D4's periodic pair-interaction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (lib2d4) never links, loads or calls anything declared here.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Positions are the reference's sorted DislwoB layout: an array of n {x, y}
 * doubles, Burgers vector implied by the index (first n/2 are +1,
 * simulation_data.cpp:265-270).
 */
#ifndef D4_ORACLE_H
#define D4_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sim_utils.cpp:6-13 */
double orc_normalize(double v);
/* sim_utils.cpp:36-54, 57-75 */
double orc_weight(double subsum, char kind);
double orc_weight_inv(double w, char kind);

/* Field.cpp:295-312 (xy, 4-arg), 314-331 (xy_diff_x, 4-arg), 333-346 (2-arg).
 * exact != 0 selects the true derivative for the |shift| >= 3 images instead of
 * the reference's cosh-for-sinh substitution of Field.cpp:269-283 (used only by
 * the adjoint's finite-difference checks; the reference path uses exact == 0). */
double orc_xy4(double dx, double dy, double e2pix, double c2piy);
double orc_dxy4(double dx, double dy, double e2pix, double c2piy, int exact);
double orc_xy(double dx, double dy);
double orc_dxy(double dx, double dy);
/* the reference's Jacobian-pass inputs: exp2(PI2LN2INV*dx), cos(2*pi*dy) (simulation.cpp:456-457) */
void orc_field_inputs(double dx, double dy, double* e2pix, double* c2piy);
/* d sigma / d dy from the closed form (no reference counterpart; adjoint only) */
double orc_dyxy4(double dx, double dy, double e2pix, double c2piy);
/* batch helpers over arrays (for tests) */
void orc_inputs_batch(int64_t n, const double* dx, const double* dy, double* e, double* c);
void orc_field_batch(int64_t n, const double* dx, const double* dy, const double* e, const double* c,
                     double* sxy, double* dsxy, int exact);

/* simulation.cpp:287-333 + precision_handler.cpp:34-38.
 * tol_cap2 <= 0 disables the tolerance update (use_dipoleCrit == false). */
void orc_speed_pass(int n, const double* xy, double tau, double* v, double* tol, double tol_cap2);

/* simulation.cpp:345-527.  CSC output (Ap n+1, Ai/Ax cap), indexes n, dvec n.
 * cut_sq = cutOffSqr, cutoff = cutOff, one_per_cut_sq = onePerCutOffSqr
 * (simulation_data.cpp:253-262).  Returns nnz, or -(required nnz) if cap is short. */
int orc_jacobian_pass(int n, const double* xy, double h, double tauA, double tauB,
                      double cutoff, double cut_sq, double one_per_cut_sq, int heaviside, char wfunc,
                      double* vA, double* vB, int* Ap, int* Ai, double* Ax, int cap,
                      int* indexes, double* dvec, double* tol, double tol_cap2);

/* simulation.cpp:546-608: stage-II rescale of the CSC matrix and dvec in place. */
void orc_jacobian_from_prev(int n, int* Ap, int* Ai, double* Ax, double* dvec, char wfunc,
                            double base_t, double new_t, double old_t,
                            double tau_old, double tau_new, double* speed);

/* Dense pass: M = V + diag(s) column-major (the reference's Ax before its
 * (1+d)/2 row scaling), speeds at tauA/tauB, dvec = weight(s). */
void orc_jacobian_dense(int n, const double* xy, double h, double tauA, double tauB, char wfunc,
                        double* vA, double* vB, double* M, double* dvec, double* tol, double tol_cap2);

/* Tabulated field (the build's format, no reference counterpart): n x n node
 * values, tab[ky*n + kx] at (-1/2 + kx/n, -1/2 + ky/n), n a power of two,
 * bilinear and periodic in between; the build fills nodes with orc_xy/orc_dxy
 * (0 at the origin). */
double orc_table_bilinear(const double* tab, int n, double dx, double dy);
void orc_table_build(int n, double* tab_s, double* tab_d);
void orc_speed_pass_table(int n, const double* xy, const double* tab, int tn, double tau, double* v);

/* simulation_data.cpp:253-262 */
void orc_cutoff(int n, double mult, int heaviside, double* cutoff, double* cut_sq, double* one_per);

/* stress_protocol.cpp:13-73.  kind: 0 base (always 0), 1 fixed rate, 2 cyclic, 3 increasing cyclic */
double orc_ext_stress(int kind, double init, double rate, double period, double amp_rate, double t);

/* Whole run of simulation.cpp:30-281 (positions only; no log / sub-config files),
 * with a dense LU in place of UMFPACK.  xy (n) is sorted input and is overwritten
 * with the final sorted positions (normalised as the reference leaves them). */
typedef struct {
    double dt0;               /* --initial-stepsize */
    double max_dt;            /* -M; <= 0 for no limit */
    double prec;              /* -P */
    double dipole_prec;       /* -p */
    double cutoff_mult;       /* -u */
    int heaviside;            /* -h */
    char wfunc;               /* -w */
    int stress_kind;          /* see orc_ext_stress */
    double stress_init, stress_rate, stress_period, stress_amp;
    double time_limit;        /* -t; <= 0 none */
    long step_limit;          /* -C; < 0 none */
    double strain_limit;      /* <= 0 none */
    double speed_threshold;   /* cutoff change threshold; <= 0 none */
} orc_run_params;

typedef struct {
    double sim_time, step_size, strain;
    long ok_steps, failed_steps;
    double last_max_ratio;
} orc_run_stats;

int orc_run(int n, double* xy, const orc_run_params* p, orc_run_stats* st);

/* Reads a .dconf (x y b per line), sorts as simulation_data.cpp:116-123.
 * Returns n or <0 on error; fills xy (2*cap), order (cap) with the reference's disl_order. */
int orc_read_dconf(const char* path, double* xy, unsigned* order, int cap);

/* Multi-threaded pass variants for the CPU baseline (OpenMP over target rows,
 * each pair evaluated twice, once per row, so rows are independent). */
void orc_speed_pass_omp(int n, const double* xy, double tau, double* v);
void orc_jacobian_dense_omp(int n, const double* xy, double h, double tauA, double tauB, char wfunc,
                            double* vA, double* vB, double* M, double* dvec);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
