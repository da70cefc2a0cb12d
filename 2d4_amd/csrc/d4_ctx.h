// d4_ctx.h -- the opaque context behind include/d4.h: device, stream and
// grow-only device buffers reused across calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "d4_kernels.h"

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
    int tiles_n = -1;  // n of the cached single-realization tile table
    DevBuf part, tiles, xy, v0, v1, tol, coo, coo_count, cnt, cursor, Ap, Ai, Ax, idx, dvec, nw, aux0, aux1, aux2;
    // pair-kernel timing: event pairs recorded around launches while profiling is on
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_marks;
    hipEvent_t take_event();
    int mark_begin(hipEvent_t* e);
    int mark_end(hipEvent_t b);

    ~d4_ctx();
    int fail(int code, const std::string& msg);
    int hip_fail(hipError_t e, const char* where);
    int ensure(DevBuf& b, size_t bytes, const char* what);
    int single_tiles(int n, int* ntiles);
    int single_pass(int mode, int n, const double* d_xy, const d4::RealParams& rp, double* d_tol, double* d_vA,
                    double* d_vB, double* d_q1, double* d_dvec, double* d_M, int64_t ld, const double* d_w, int exact,
                    int want_dy, int add_tau);
    int sync(const char* where);
    int h2d(void* dst, const void* src, size_t bytes, const char* where);
    int d2h(void* dst, const void* src, size_t bytes, const char* where);
};
