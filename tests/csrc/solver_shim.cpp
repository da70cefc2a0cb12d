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

// the GPU dense LU's symmetric order and block envelope (host code)
extern "C" int ts_band_order(int n, int npad, int bs, const int* Ap, const int* Ai, int* ord, int* first, int* last,
                             int* top, long long* tiles)
{
    std::vector<int> o, inv;
    d4host::BlockEnvelope e;
    const bool re = d4host::band_order(n, npad, bs, Ap, Ai, o, inv, e);
    for (int i = 0; i < n; ++i) ord[i] = re ? o[i] : i;
    for (int b = 0; b < npad / bs; ++b) {
        first[b] = e.first[b];
        last[b] = e.last[b];
        top[b] = e.top[b];
    }
    *tiles = e.tiles;
    return re ? 1 : 0;
}

// the block envelope of a pattern in an order (inv[old] = new, or null), by the general or the symmetric pass
extern "C" void ts_block_envelope(int n, int npad, int bs, const int* Ap, const int* Ai, const int* inv, int symmetric,
                                  int* first, int* last, int* top, long long* tiles)
{
    const d4host::BlockEnvelope e = d4host::block_envelope(n, npad, bs, Ap, Ai, inv, symmetric != 0);
    for (int b = 0; b < npad / bs; ++b) {
        first[b] = e.first[b];
        last[b] = e.last[b];
        top[b] = e.top[b];
    }
    *tiles = e.tiles;
}

// the envelope fixed by sorted keys that couple only within reach (host code)
extern "C" void ts_key_band_envelope(int n, int npad, int bs, const double* key, double reach, int* first, int* last,
                                     int* top, long long* tiles)
{
    const d4host::BlockEnvelope e = d4host::key_band_envelope(n, npad, bs, key, reach);
    for (int b = 0; b < npad / bs; ++b) {
        first[b] = e.first[b];
        last[b] = e.last[b];
        top[b] = e.top[b];
    }
    *tiles = e.tiles;
}

extern "C" void ts_rcm_order(int n, const int* Ap, const int* Ai, int* ord)
{
    const std::vector<int> o = d4host::rcm_order(n, Ap, Ai);
    for (int i = 0; i < n; ++i) ord[i] = o[i];
}

// two factorizations with one solver (the same pattern, values Ax1 then Ax2),
// then x = A2^{-1} b; *replayed: how many of them replayed the first's reach
extern "C" int ts_refactor(int n, const int* Ap, const int* Ai, const double* Ax1, const double* Ax2, const double* b,
                           double* x, long long* replayed)
{
    d4host::SparseLU lu;
    if (lu.factor(n, Ap, Ai, Ax1) < 0) return -1;
    const int rc = lu.factor(n, Ap, Ai, Ax2);
    if (rc < 0) return rc;
    lu.solve_const(b, x);
    *replayed = lu.refactorizations();
    return rc;
}

// two factorizations with one solver, of two matrices that may differ in
// pattern, then x = A2^{-1} b; the counts of orders computed and of replays
extern "C" int ts_two_patterns(int n, const int* Ap1, const int* Ai1, const double* Ax1, const int* Ap2,
                               const int* Ai2, const double* Ax2, const double* b, double* x, long long* orderings,
                               long long* replayed)
{
    d4host::SparseLU lu;
    if (lu.factor(n, Ap1, Ai1, Ax1) < 0) return -1;
    const int rc = lu.factor(n, Ap2, Ai2, Ax2);
    if (rc < 0) return rc;
    lu.solve_const(b, x);
    *orderings = lu.orderings();
    *replayed = lu.refactorizations();
    return rc;
}
