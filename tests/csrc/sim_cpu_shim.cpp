// sim_cpu_shim.cpp -- test-only: the five include/d4.h entry points that the
// host simulation (2d4_amd/csrc/d4_sim.cpp) calls, answered by the oracle on
// the CPU, so the simulation's host logic (stages I-IV, step control, CSC
// capacity growth, sparse LU) can be checked without a GPU.  Linked only into
// tests/csrc/libsim_cpu.so together with d4_sim.cpp, d4_solver.cpp and the
// oracle; the product library never contains this file.
#include <cmath>

#include "../../include/d4.h"
#include "../../oracle/d4_oracle.h"
#include "../../2d4_amd/csrc/d4_solver.h"

extern "C" {

const char* d4_last_error(const d4_ctx*) { return "oracle shim"; }
int d4_ctx_device(const d4_ctx*) { return 0; }

int d4_field_eval(d4_ctx*, int64_t n, const double* dx, const double* dy, const double* e2pix, const double* c2piy,
                  double* sxy, double* dsxy, int flags)
{
    for (int64_t k = 0; k < n; ++k) {
        double e, c;
        if (e2pix && c2piy) {
            e = e2pix[k];
            c = c2piy[k];
        } else {
            orc_field_inputs(dx[k], dy[k], &e, &c);
        }
        if (sxy) sxy[k] = orc_xy4(dx[k], dy[k], e, c);
        if (dsxy) dsxy[k] = orc_dxy4(dx[k], dy[k], e, c, flags & 1);
    }
    return D4_OK;
}

int d4_speed_pass(d4_ctx*, int n, const double* xy, double tau, double* v, double* tol, double tol_cap2)
{
    orc_speed_pass(n, xy, tau, v, tol, tol_cap2);
    return D4_OK;
}

int d4_jacobian_pass(d4_ctx*, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB, int* Ap,
                     int* Ai, double* Ax, int cap, int* nnz, int* indexes, double* dvec, double* tol)
{
    const int r = orc_jacobian_pass(n, xy, p->h, p->tau_a, p->tau_b, p->cutoff, p->cut_sq, p->one_per_cut_sq,
                                    p->heaviside, p->wfunc, vA, vB, Ap, Ai, Ax, cap, indexes, dvec, tol, p->tol_cap2);
    if (r < 0) {
        *nnz = -r;
        return D4_ERR_CAPACITY;
    }
    *nnz = r;
    return D4_OK;
}

// factor = (new - base)/(old - base) and dtau = tau_new - tau_old are passed
// through exactly as base = 0, new = factor, old = 1, tau_old = 0, tau_new = dtau
int d4_jacobian_rescale(d4_ctx*, int n, const int* Ap, const int* Ai, double* Ax, double* dvec, char wfunc,
                        double factor, double dtau, double* speed)
{
    orc_jacobian_from_prev(n, const_cast<int*>(Ap), const_cast<int*>(Ai), Ax, dvec, wfunc, 0.0, factor, 1.0, 0.0,
                           dtau, speed);
    return D4_OK;
}

}  // extern "C"

// the GPU dense LU is not part of this CPU build
d4host::LinearSolver* d4host::new_gpu_dense_lu(int) { return nullptr; }
