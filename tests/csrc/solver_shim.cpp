// solver_shim.cpp -- test-only C entry points over the host sparse LU
// (2d4_amd/csrc/d4_solver.cpp), so it can be checked on the CPU.
#include "../../2d4_amd/csrc/d4_solver.h"

extern "C" int ts_solve(int n, const int* Ap, const int* Ai, const double* Ax, const double* b, double* x,
                        long long* fill)
{
    d4host::SparseLU lu;
    int rc = lu.factor(n, Ap, Ai, Ax);
    if (rc < 0) return rc;
    lu.solve_const(b, x);
    if (fill) *fill = lu.fill();
    return rc;
}
