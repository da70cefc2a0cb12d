// d4_kernels.h -- launch interface of the gfx950 pair kernels (internal to
// lib2d4; the public boundary is include/d4.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace d4 {

constexpr int kTile = 64;  // one wavefront per 64 x 64 block of unordered pairs
constexpr int kXcds = 8;   // MI355X: workgroups are handed to its 8 XCDs round-robin by workgroup index

// One tile of a (possibly ragged, batched) pass: realization r, block rows
// [I0, I0+64) x block columns [J0, J0+64) of that realization, I0 <= J0.
// slot = canonical index of the tile's partial sums.
struct TileDesc {
    int r;
    int I0;
    int J0;
    int slot;
};

// Per-realization scalars of a pass (h = 0 for speed-only passes).
struct RealParams {
    double h;
    double tauA, tauB;
    double range_sq;    // Jacobian storage radius^2: rc^2 (heaviside) or 49 rc^2
    double cut_sq;      // rc^2
    double cutoff;      // rc
    double one_per;     // 1 / rc^2
    double tol_cap2;    // <= 0: no tolerance tracking
    int heaviside;
    int wfunc;          // 'c', 'p', 'm'
    int slot0;          // first partial slot of this realization
    int nb;             // blocks of this realization
};

struct CooEntry {
    int i;  // lower index (column of the primary copy)
    int j;  // higher index
    double v;
};

enum PassMode : int {
    kModeSpeed = 0,      // sigma only: speeds, tolerance
    kModeJacDense = 1,   // sigma + d sigma: speeds, V dense column-major, row sums
    kModeJacCoo = 2,     // sigma all pairs; the pairs inside the storage radius -> COO list, whose
                         // d sigma launch_coo_derivative adds
    kModeVjp = 3,        // d sigma/dx (+ d sigma/dy): w-weighted pair sums
    kModeJacDenseCut = 4,  // as kModeJacDense with the storage radius of kModeJacCoo:
                           // V = 0 outside it (the reference's cutoff, dense storage)
    kModeSpeedTab = 5,     // kModeSpeed with sigma interpolated in a table
    kModeJacDenseTab = 6,  // kModeJacDense with sigma and d sigma interpolated in tables
};

struct PassArgs {
    const double2* xy;        // concatenated sorted positions
    const int* off;           // R+1 realization offsets (null: one realization of n0)
    int n0;
    const TileDesc* tiles;
    int ntiles;
    const RealParams* rp;     // R (null: rp0)
    RealParams rp0;
    double* part;             // partial sums [slot][NQ][128]
    double* tol;              // per element (may be null)
    // dense Jacobian output (modes 1, 4): one matrix M (leading dimension ld),
    // or, with moff, matrix r at M + moff[r] with leading dimension n_r
    double* M;
    int64_t ld;
    const int64_t* moff;
    // COO output (mode 2)
    CooEntry* coo;
    int* coo_count;
    int coo_cap;
    // VJP input (mode 3)
    const double* w;
    // tabulated field (modes 5, 6): tab_n x tab_n cells over [-1/2, 1/2)^2 as cell records in tab_s
    // (launch_table_records): sigma's four corners (mode 5: 4 doubles), then d sigma's (mode 6: 8 doubles)
    const double* tab_s;
    const double* tab_d;
    int tab_n;
    int exact;                // mode 1/2/3: 1 = true derivative instead of the reference's
    int want_dy;              // mode 3: also d/dy
    int skip_empty;           // mode 4: leave M untouched for the pairs of a 16-step phase with none inside the
                              // storage radius (for callers that read no entry outside it)
};

// The pair-list pass's second half: d sigma of each listed pair into its entry.
struct CooDerivArgs {
    const double2* xy;        // concatenated sorted positions
    const int* off;           // R+1 realization offsets (null: one realization of n0)
    const int* elem;          // realization of each element (null: one realization)
    int n0;
    const RealParams* rp;     // R (null: rp0)
    RealParams rp0;
    CooEntry* coo;
    const int* count;         // pairs the pass listed (device)
    int cap;                  // list capacity (the grid covers it)
};
hipError_t launch_coo_derivative(const CooDerivArgs& a, int exact, hipStream_t s);

// number of accumulated quantities per mode
inline int pass_nq(int mode)
{
    return mode == kModeSpeed || mode == kModeJacCoo || mode == kModeSpeedTab ? 1 : 2;
}

// grid = 0: one workgroup per tile; grid > 0: `grid` persistent workgroups
// that stride over the tile table
hipError_t launch_pair_pass(int mode, const PassArgs& a, hipStream_t s, int grid = 0);

// Node values of the tabulated field: tab[ky * n + kx] = sigma (and d sigma/dx)
// at (dx, dy) = (-1/2 + kx/n, -1/2 + ky/n), from the analytic field (0 at the origin).
hipError_t launch_table_build(int n, double* tab_s, double* tab_d, hipStream_t s);
// cell records of node tables: rec[(ky n + kx) stride + q] = tab_s corners q = 0..3 ((kx, ky), (kx+1, ky),
// (kx, ky+1), (kx+1, ky+1), wrapped), then with tab_d (stride 8) tab_d's corners
hipError_t launch_table_records(int n, const double* tab_s, const double* tab_d, double* rec, hipStream_t s);
// bilinear values of the tables for m independent pairs
hipError_t launch_table_eval(int64_t m, const double* dx, const double* dy, const double* tab_s, const double* tab_d,
                             int tab_n, double* s_out, double* d_out, hipStream_t s);

// Field values for n independent pairs (E, c may be null -> computed on device).
hipError_t launch_field_eval(int64_t n, const double* dx, const double* dy, const double* E, const double* c,
                             double* sxy, double* dsxy, int exact, hipStream_t s);

// Deterministic reduction of tile partials into per-element sums.
// out0[e] = sum of quantity 0 (+ tau terms for speeds), out1 = quantity 1 (may be null).
struct ReduceArgs {
    const double* part;
    int nq;
    const int* off;          // null: one realization of `total` elements
    const RealParams* rp;    // null: rp0
    RealParams rp0;
    int R;
    int total;               // sum of n_r
    const int2* blocks;      // (realization, element block) per reduce block (null: one realization)
    int nblocks;
    double* vA;              // quantity 0 + tauA b
    double* vB;              // quantity 0 + tauB b (may be null)
    double* q1;              // quantity 1 raw (may be null)
    double* dvec;            // weight(q1) (may be null)
    double* M;               // dense diagonal <- q1 (may be null)
    int64_t ld;
    const int64_t* moff;     // per-realization matrix offsets (leading dimension n_r), or null
    int add_tau;             // 1: add tau b to vA/vB
};
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s);

// CSC assembly from a COO list over a block-diagonal batch (global indices).
// ncoo_dev (optional): device-side pair count, capped by ncoo (the grid is sized by ncoo)
hipError_t launch_csc_count(const CooEntry* coo, int ncoo, int* cnt, int total, hipStream_t s,
                            const int* ncoo_dev = nullptr);
// Ap[0..total] = exclusive scan of cnt; scratch (2 * ceil(total / 2048) + 1 ints, or null: one
// workgroup) lets a long scan spread over many workgroups
hipError_t launch_scan(const int* cnt, int* Ap, int total, hipStream_t s, int* scratch);
hipError_t launch_csc_fill(const CooEntry* coo, int ncoo, const int* Ap, int* cursor, int* Ai, double* Ax,
                           int total, hipStream_t s, const int* ncoo_dev = nullptr);
// the single-realization tile table of n (row-major upper triangle, slot = index) built on the device
hipError_t launch_tile_table(TileDesc* t, int n, hipStream_t s);

// the pair term of the implicit step's vector-Jacobian product (d4_adjoint.hip); rp: h = the step's k,
// the storage radius and taper
struct ThetaVjpArgs {
    int n;
    const double2* xy;
    RealParams rp;
    const double *a, *c, *u, *w, *t;
    double* gx;
    double* gy;  // may be null
};
hipError_t launch_theta_pair_vjp(const ThetaVjpArgs& a, hipStream_t s);
// p[0..n) = v (a kernel, so it can sit in a captured graph)
hipError_t launch_fill_int(int* p, int v, int n, hipStream_t s);
// wfunc per realization from wf_arr (null: wf0 for all)
hipError_t launch_csc_finish(const int* Ap, int* Ai, double* Ax, int* indexes, double* dvec, const int* elem_real,
                             const int* wf_arr, int wf0, int total, hipStream_t s);
hipError_t launch_csc_scale(const int* Ap, const int* Ai, double* Ax, const int* indexes, const double* dvec,
                            int total, hipStream_t s);
// stage-II rescale (simulation.cpp:546-608)
// per-realization factor / dtau / wfunc from the arrays, or the scalars when the arrays are null
struct RescaleArgs {
    const int* Ap;
    const int* Ai;
    double* Ax;
    double* dvec;
    double* nw;              // scratch, total
    double* speed;           // may be null
    const int* elem_real;    // null: one realization
    const int* off;          // null: one realization of `total`
    const double* factor_arr;
    const double* dtau_arr;
    const int* wf_arr;
    double factor0, dtau0;
    int wf0;
    int total;
};
hipError_t launch_rescale(const RescaleArgs& a, hipStream_t s);
// dense final scaling: M[j,i] *= (1+d_j)/2, diagonal += 1
hipError_t launch_dense_scale(double* M, int64_t ld, int n, const double* dvec, hipStream_t s);

// Batched dense matrices: realization r's n_r x n_r column-major matrix at
// moff[r]; elements g in [off[r], off[r+1]) are its rows and columns.
struct DenseBatch {
    const int* off;
    const int* elem_real;
    const int64_t* moff;
    int total;               // sum of n_r
    int maxn;                // max n_r
};
// F = the reference's final matrix from M = V + diag(s): F[j,i] = M[j,i] (1+d_j)/2 (+1 on the diagonal)
hipError_t launch_dense_scale_batched(const DenseBatch& b, const double* M, double* F, const double* dvec,
                                      hipStream_t s);
// stage-II matrix from the same M (simulation.cpp:546-598): with A = the stage-I
// final matrix, F[j,i] = A[j,i]/2 f_j off the diagonal, 1 + (A[j,j]-1)/2 f_j on it,
// f_j = (1 + dnew_j)/(1 + dold_j)
hipError_t launch_dense_rescale_batched(const DenseBatch& b, const double* M, double* F, const double* dold,
                                        const double* dnew, hipStream_t s);

}  // namespace d4
