// d4_solver.h -- host sparse LU for the implicit step's linear systems.
//
// Stands where the reference calls UMFPACK (simulation.cpp:635-659:
// umfpack_di_symbolic / _numeric / _solve(UMFPACK_A) / _free_*), which this
// image does not provide.  Same contract: factor a square CSC matrix (int32
// Ap/Ai, double Ax, ascending rows), then solve A x = b any number of times;
// a singular matrix is reported like UMFPACK_WARNING_singular_matrix and the
// solve still returns (the caller zeroes unusable entries, simulation.cpp:646-658).
//
// Method (own implementation of published algorithms):
//   * fill-reducing column order: minimum degree on the pattern of A + A^T
//     (elimination-graph form, bit-set rows);
//   * numeric factorization: left-looking Gilbert-Peierls LU (sparse
//     triangular solve per column over the reach found by depth-first search)
//     with threshold partial pivoting that keeps the diagonal when
//     |a_diag| >= 0.1 max |column| (the diagonal preference UMFPACK uses);
//   * dense LU with partial pivoting instead when the matrix is nearly full;
//   * the column order is kept while the pattern stays within 1 % (in
//     entries) of the pattern it was computed for: between two Jacobians of a
//     run only a few pairs cross the cutoff, and any column order is valid.
//     The ordering dominates at these sizes: 2.5 of 3.5 ms at N = 1024;
//   * with the pattern unchanged, each column's reach (found by the depth-first
//     search, half the work) is replayed from the last factorization for as long
//     as every pivot choice repeats: the same operations in the same order, so
//     the results are bitwise those of a full factorization; the first pivot
//     that differs sends it back to the full one.
#pragma once

#include <cstdint>
#include <vector>

// HIP's stream handle (hip_runtime_api.h declares the same), so that host-only builds of this header
// need no HIP headers
typedef struct ihipStream_t* hipStream_t;

namespace d4host {

// True on the ensemble's host worker threads.  Those threads only move host
// memory and run host solvers; HIP and rocBLAS are called from the thread
// that steps the ensemble, so that every stream dependency is one that thread
// orders (an event or a synchronization).  The GPU LU refuses to run on a
// worker thread.
inline thread_local bool t_host_worker = false;

// factor a square CSC matrix once, then solve A x = b any number of times
class LinearSolver {
public:
    virtual ~LinearSolver() = default;
    // 0 = ok, 1 = singular (some zero pivot), < 0 = failure
    virtual int factor(int n, const int* Ap, const int* Ai, const double* Ax) = 0;
    // x = A^{-1} b; 0 = ok, < 0 = failure
    virtual int solve(const double* b, double* x) = 0;
    virtual bool singular() const = 0;
    // true when the last factor() took no row interchanges by construction
    // (a solver that may skip pivoting on matrices where it is not needed)
    virtual bool unpivoted() const { return false; }
};

class SparseLU : public LinearSolver {
public:
    // returns 0 = ok, 1 = singular (some zero pivot), <0 = invalid input
    int factor(int n, const int* Ap, const int* Ai, const double* Ax) override;
    // x = A^{-1} b (b and x may alias)
    int solve(const double* b, double* x) override
    {
        solve_const(b, x);
        return 0;
    }
    void solve_const(const double* b, double* x) const;
    bool singular() const override { return singular_; }
    int64_t fill() const { return (int64_t)Li_.size() + (int64_t)Ui_.size(); }
    // factorizations that replayed the previous one's reach (same pattern and pivots)
    int64_t refactorizations() const { return refactorizations_; }
    int64_t orderings() const { return orderings_; }  // minimum degree orders computed

private:
    int n_ = 0;
    bool dense_ = false;
    bool singular_ = false;
    // column order q (factor column k = A column q[k]); row pivot pinv (row -> k)
    std::vector<int> q_, pinv_, prow_;
    int pat_n_ = -1;                     // the pattern of the last factorization
    std::vector<int> pat_p_, pat_i_;
    int ord_n_ = -1;                     // the pattern q_ was computed for
    std::vector<int> ord_p_, ord_i_;
    // L unit lower (stored without the unit diagonal), U upper with diagonal last in each column
    std::vector<int> Lp_, Li_, Up_, Ui_;
    std::vector<double> Lx_, Ux_;
    // dense fallback: column-major LU of the full matrix, row pivots
    std::vector<double> D_;
    std::vector<int> dpiv_;
    int factor_dense(int n, const int* Ap, const int* Ai, const double* Ax);
    // numeric refactorization over the last factorization's reach (see d4_solver.cpp)
    bool refactor(int n, const int* Ap, const int* Ai, const double* Ax);
    bool have_reach_ = false;
    std::vector<int> reach_p_, reach_i_;  // per column, its reach in topological order
    std::vector<int> Lrow_;               // L row indices in original rows
    std::vector<double> work_;
    int64_t refactorizations_ = 0, orderings_ = 0;
};

// minimum degree order of the symmetric pattern of A + A^T (exposed for tests)
std::vector<int> min_degree_order(int n, const int* Ap, const int* Ai);

// Reverse Cuthill-McKee order (ord[new] = old) over the graph whose
// neighbours of node c are the rows of column c: per connected component
// (taken by lowest index) a breadth-first search from a pseudo-peripheral
// node (George and Liu's search), neighbours in row order; the whole sequence
// reversed.  For a symmetric pattern (the step matrices') that is the graph
// of A; for any other pattern the order is still a permutation, and
// block_envelope measures the pattern of A + A^T under it.  Deterministic:
// the order depends on the pattern alone.
std::vector<int> rcm_order(int n, const int* Ap, const int* Ai);

// Block envelope of the pattern of A + A^T with rows and columns renumbered
// by inv (inv[old] = new; null: as they are), in blocks of bs, for a matrix
// padded with the identity to npad (a multiple of bs).  With f(i) the lowest
// column of row i's entries (f(i) <= i), Gaussian elimination without
// interchanges keeps L(i, j) = 0 for j < f(i) and, the pattern being
// symmetric, U(j, i) = 0 for j < f(i).  Per block b of nb = npad / bs:
//   first[b]: min over the rows i of block b of f(i), in blocks;
//   last[b]:  the highest block b' with first[b'] <= b: at factorization step
//             b the blocks of L below and of U right of the diagonal block
//             are zero beyond block last[b] (nondecreasing, last[b] >= b);
//   top[b]:   the lowest step p with last[p] >= b (top[b] <= first[b]).
// tiles = sum_p (last[p] - p)^2, the trailing-update tiles of a blocked LU
// that skips the zero blocks.
struct BlockEnvelope {
    std::vector<int> first, last, top;
    long long tiles = 0;
};
// symmetric: the caller knows the pattern is symmetric (entry (i, j) iff (j, i)), which lets one gather pass
// per column replace the scattered per-entry updates; the result is the same
BlockEnvelope block_envelope(int n, int npad, int bs, const int* Ap, const int* Ai, const int* inv,
                             bool symmetric = false);

// The block envelope (as block_envelope) of every matrix whose rows and columns, sorted by a key
// (key[0] <= key[1] <= ...), couple only where the keys differ by less than reach: a bound that holds for
// any pattern with that property, so it can be fixed before the pattern is known.  The implicit step's
// matrix couples dislocations closer than the storage radius, hence closer than it in y; with the key
// |y - y0| (the distance from a line, periodic) the pairs that may couple differ in key by less than the
// radius, for any x.
BlockEnvelope key_band_envelope(int n, int npad, int bs, const double* key, double reach);

// The symmetric order of a matrix for the GPU dense LU: reverse Cuthill-McKee
// when its envelope needs at most half the trailing-update tiles of the given
// order, else the given order.  Returns true and fills ord (ord[new] = old)
// and inv (inv[old] = new) when it reorders; env is the envelope of the order
// chosen.
bool band_order(int n, int npad, int bs, const int* Ap, const int* Ai, std::vector<int>& ord, std::vector<int>& inv,
                BlockEnvelope& env, bool symmetric = false);

// Dense LU on the GPU: the CSC matrix is scattered into a device matrix of
// order gpu_lu_padded(n) (identity beyond n) and factored there with
// threshold partial pivoting (u = 0.001, diagonal preferred, as UMFPACK's
// symmetric strategy: a blocked LU without interchanges, kept when every
// multiplier is at most 1000 in magnitude, else rocSOLVER getrf), then solved there by a triangular-solve
// kernel; one stream and rocBLAS handle per calling thread.  It is BatchDenseLU with one matrix, so a
// matrix gets the same bits from either.  It is the simulation's solver: its matrices are one simulation's
// sequence of step matrices, whose patterns are symmetric (Matrix::key and Matrix::symmetric set).  Defined
// in d4_gpulu.hip; null if the device cannot be used.
LinearSolver* new_gpu_dense_lu(int device);

// order of the padded device matrix for an n x n system: n rounded up to a
// multiple of 64, the factorization's block
int gpu_lu_padded(int n);

// The dense LU of new_gpu_dense_lu for K matrices at once: the unpivoted
// factorization covers every matrix in the same launches (three per 64
// columns), the matrices that need row interchanges go through rocSOLVER
// getrf (one call per matrix, spread over side streams), and all solves through one launch of the
// triangular-solve kernel.  A matrix's factors and
// solutions do not depend on which other matrices share the call.
class BatchDenseLU {
public:
    struct Matrix {
        int n;
        const int* Ap;  // the pattern on the host (always: it chooses the order)
        const int* Ai;
        const double* Ax;
        // optionally the same matrix already on the device, used instead of copying Ap/Ai/Ax there:
        // column c's entries at dAi/dAx[dAp[c] .. dAp[c + 1]), row r stored as r + row_base
        const int* dAp = nullptr;
        const int* dAi = nullptr;
        const double* dAx = nullptr;
        int row_base = 0;
        // >= 0: the caller's name for this matrix across calls (a realization, one simulation's solver).
        // Its order is then kept when its pattern changes while the pattern's envelope in that order stays
        // within 1/8 of the envelope the order was chosen with (a matrix's order then depends on its own
        // sequence of patterns, not on the other matrices of a call); < 0 (or a key named twice in one
        // call): an order is kept only while the pattern of the matrix slot repeats exactly
        int key = -1;
        // the pattern is symmetric (entry (i, j) iff (j, i)), as the step matrices' are: the envelope is
        // then found without scattered updates
        bool symmetric = false;
    };
    // runs fn(k) for k in [0, count), possibly on several threads
    using ParallelFor = void (*)(void* user, int count, void (*fn)(void*, int), void* arg);
    virtual ~BatchDenseLU() = default;
    virtual void set_parallel(ParallelFor pf, void* user) = 0;
    // 0 = ok, < 0 = failure; singular[k] and unpivoted[k] as LinearSolver reports them
    virtual int factor(int K, const Matrix* m, unsigned char* singular, unsigned char* unpivoted) = 0;
    // x[k] = A_k^{-1} b[k] for the K matrices of the last factor(); 0 = ok, < 0 = failure
    virtual int solve(int K, const double* const* b, double* const* x) = 0;
    // one dense n x n matrix on the device, element (r, c) at a[r * rs + c * cs]; the
    // same factorization as factor() gives that matrix in CSC form
    virtual int factor_dense_dev(int n, const double* a, long long rs, long long cs, unsigned char* singular,
                                 unsigned char* unpivoted) = 0;
    // x = A^{-1} b for the matrix of the last factor_dense_dev(); b and x on the device
    virtual int solve_dense_dev(int n, const double* b, double* x) = 0;
};
BatchDenseLU* new_batch_dense_lu(int device);

// The same dense LU for a batch of matrices whose orders are fixed once, by launches alone: no
// allocation, no host copy and no wait after creation, so that a caller can capture factor and solve in
// a HIP graph.  Matrix k is column-major of order gpu_lu_padded(n[k]) at store + a_off(k) (the caller
// assembles it there, identity beyond n[k]).  Natural order (no reordering), threshold partial pivoting
// with u = 0.001: factor_unpivoted() factors every matrix without interchanges and sets flag[k] where a
// pivot is not finite and nonzero or a multiplier exceeds 1000; the caller then assembles the flagged
// matrices again and factor_pivoted() factors them with partial pivoting (singular[k]: a zero pivot was
// met); solve() overwrites each right-hand side, packed at b_off[k], with its solution.  Null if the
// batch does not fit (an order above the device's LDS bound of the fallback) or the device fails.
class DenseLUGraph {
public:
    virtual ~DenseLUGraph() = default;
    virtual long long a_off(int k) const = 0;
    virtual long long store_size() const = 0;  // doubles
    // the device array of orders and envelopes, and matrix k's places in it (-1: natural order / whole
    // matrix), laid out as the CSC path's aux: ord (new -> old) then inv; last, top, first per block
    virtual const int* aux() const = 0;
    virtual int ord_off(int k) const = 0;
    virtual int env_off(int k) const = 0;
    virtual int factor_unpivoted(double* store, int* flag, hipStream_t st) = 0;
    virtual int factor_pivoted(double* store, const int* flag, int* singular, hipStream_t st) = 0;
    virtual int solve(const double* store, double* b, hipStream_t st) = 0;
};
// ord[k] (ord[new] = old; empty: the natural order) and env[k] (empty last: the whole matrix) are each
// matrix's symmetric order and block envelope, fixed for every factorization; either pointer may be null.
// The caller assembles each matrix in its order, at least inside its envelope (blocks top .. last of each
// column's block), and in full before factor_pivoted for the matrices it flags.
DenseLUGraph* new_dense_lu_graph(int device, int K, const int* n, const long long* b_off,
                                 const std::vector<std::vector<int>>* ord = nullptr,
                                 const std::vector<BlockEnvelope>* env = nullptr);

}  // namespace d4host
