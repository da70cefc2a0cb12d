// d4_ctx.h -- the opaque context behind include/d4.h: device, stream and
// grow-only device buffers reused across calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/d4.h"
#include "d4_kernels.h"

namespace d4 {
// RealParams of one realization of n from a Jacobian pass's parameters, or
// (p == null) of a speed pass at stress tau
inline RealParams real_params(int n, const d4_jac_params* p, double tau)
{
    RealParams r{};
    r.nb = (n + kTile - 1) / kTile;
    r.slot0 = 0;
    if (p) {
        r.h = p->h;
        r.tauA = p->tau_a;
        r.tauB = p->tau_b;
        r.cut_sq = p->cut_sq;
        r.cutoff = p->cutoff;
        r.one_per = p->one_per_cut_sq;
        r.range_sq = (p->heaviside ? 1.0 : 49.0) * p->cut_sq;  // simulation.cpp:461-465
        r.heaviside = p->heaviside;
        r.wfunc = p->wfunc;
        r.tol_cap2 = p->tol_cap2;
    } else {
        r.tauA = r.tauB = tau;
        r.wfunc = 'c';
    }
    return r;
}
}  // namespace d4

struct BatchPasses;
void batch_passes_free(BatchPasses* b);
namespace d4host {
class BatchDenseLU;
}
void batch_lu_free(d4host::BatchDenseLU* lu);  // d4_gpulu.hip

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct d4_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    int tiles_n = -1;      // n of the cached single-realization tile table (row-major)
    int tiles_off_n = -1;  // n of the cached table in offset order (single_tiles)
    DevBuf part, tiles, tiles_off, xy, v0, v1, tol, coo, coo_count, cnt, cursor, Ap, Ai, Ax, idx, dvec, nw, aux0, aux1, aux2,
        tab_rec;
    // pair-kernel timing: event pairs recorded around launches while profiling is on
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_marks;
    // this device's copy of the allocation-free entry point's scratch array (d4_capi.hip), looked up
    // once per device when a context is created on it
    void* noalloc_scratch = nullptr;
    BatchPasses* batch = nullptr;  // tables and buffers of d4_batch_* (d4_batch.hip)
    d4host::BatchDenseLU* lu = nullptr;  // d4_lu_solve_batch's solver (d4_gpulu.hip)
    // tabulated field for the table pass modes (set around single_pass)
    const double* tab_s = nullptr;  // cell records (launch_table_records)
    const double* tab_d = nullptr;
    int tab_n = 0;
    hipEvent_t take_event();
    int mark_begin(hipEvent_t* e);
    int mark_end(hipEvent_t b);

    ~d4_ctx();
    int fail(int code, const std::string& msg);
    int hip_fail(hipError_t e, const char* where);
    int ensure(DevBuf& b, size_t bytes, const char* what);
    int single_tiles(int n, int* ntiles, bool by_offset = false);
    int single_pass(int mode, int n, const double* d_xy, const d4::RealParams& rp, double* d_tol, double* d_vA,
                    double* d_vB, double* d_q1, double* d_dvec, double* d_M, int64_t ld, const double* d_w, int exact,
                    int want_dy, int add_tau);
    int sync(const char* where);
    int h2d(void* dst, const void* src, size_t bytes, const char* where);
    int d2h(void* dst, const void* src, size_t bytes, const char* where);

    // Pinned staging for the single-simulation passes, whose host arrays are
    // ordinary (pageable) memory: an upload is copied into the arena and sent
    // asynchronously; a download lands in the arena and is copied to the
    // caller's array when sync() returns.  A pass calls stage_begin() first.
    struct Pending {
        void* dst;
        size_t off, bytes;
    };
    void* pin = nullptr;
    size_t pin_cap = 0, pin_used = 0;
    std::vector<Pending> pending;
    void stage_begin();
    int stage(size_t bytes, size_t* off, const char* where);
    int h2d_s(void* dst, const void* src, size_t bytes, const char* where);
    int d2h_s(void* dst, const void* src, size_t bytes, const char* where);
};
