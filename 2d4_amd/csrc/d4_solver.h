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

class SparseLU {
public:
    // returns 0 = ok, 1 = singular (some zero pivot), <0 = invalid input
    int factor(int n, const int* Ap, const int* Ai, const double* Ax);
    // x = A^{-1} b (b and x may alias)
    void solve(const double* b, double* x) const;
    bool singular() const { return singular_; }
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

}  // namespace d4host
