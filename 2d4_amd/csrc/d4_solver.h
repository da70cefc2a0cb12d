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
//   * dense LU with partial pivoting instead when the matrix is nearly full.
#pragma once

#include <cstdint>
#include <vector>

namespace d4host {

// factor a square CSC matrix once, then solve A x = b any number of times
class LinearSolver {
public:
    virtual ~LinearSolver() = default;
    // 0 = ok, 1 = singular (some zero pivot), < 0 = failure
    virtual int factor(int n, const int* Ap, const int* Ai, const double* Ax) = 0;
    // x = A^{-1} b; 0 = ok, < 0 = failure
    virtual int solve(const double* b, double* x) = 0;
    virtual bool singular() const = 0;
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

private:
    int n_ = 0;
    bool dense_ = false;
    bool singular_ = false;
    // column order q (factor column k = A column q[k]); row pivot pinv (row -> k)
    std::vector<int> q_, pinv_, prow_;
    // L unit lower (stored without the unit diagonal), U upper with diagonal last in each column
    std::vector<int> Lp_, Li_, Up_, Ui_;
    std::vector<double> Lx_, Ux_;
    // dense fallback: column-major LU of the full matrix, row pivots
    std::vector<double> D_;
    std::vector<int> dpiv_;
    int factor_dense(int n, const int* Ap, const int* Ai, const double* Ax);
};

// minimum degree order of the symmetric pattern of A + A^T (exposed for tests)
std::vector<int> min_degree_order(int n, const int* Ap, const int* Ai);

// Dense LU with partial pivoting on the GPU (rocSOLVER getrf / getrs): the CSC
// matrix is scattered into an n x n device matrix, factored there and solved
// there; one stream and rocBLAS handle per calling thread.  Defined in
// d4_gpulu.hip; null if the device cannot be used.
LinearSolver* new_gpu_dense_lu(int device);

}  // namespace d4host
