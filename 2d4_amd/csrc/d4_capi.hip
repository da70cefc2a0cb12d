// d4_capi.hip -- the C ABI of include/d4.h over the gfx950 kernels.
#include "../../include/d4.h"
#include "d4_ctx.h"
#include "d4_field.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace d4;

// Scratch of d4_speed_pass_dev_noalloc: the tile table and the tile partial sums of one speed pass
// over at most D4_NOALLOC_MAX_N dislocations, in device memory the library holds from load to exit
// (8.6 MB on each GPU it runs on).  Every call on a GPU shares that GPU's copy.
namespace {
constexpr int kNoAllocMaxTiles = ((D4_NOALLOC_MAX_N + kTile - 1) / kTile) * ((D4_NOALLOC_MAX_N + kTile - 1) / kTile + 1) / 2;
constexpr size_t kNoAllocTileDoubles = (sizeof(TileDesc) * kNoAllocMaxTiles + 255) / 256 * 256 / sizeof(double);
constexpr size_t kNoAllocDoubles = kNoAllocTileDoubles + 128 * (size_t)kNoAllocMaxTiles;
}  // namespace
static __device__ __align__(256) double scratch[kNoAllocDoubles];

// this device's address of `scratch` (host code never passes the symbol itself to a kernel): looked up
// once per device, with that device current
static int noalloc_scratch_ptr(int device, void** out)
{
    static std::mutex m;
    static std::vector<void*> per_device;
    std::lock_guard<std::mutex> l(m);
    if (device < 0) return D4_ERR_INVALID;
    if ((int)per_device.size() <= device) per_device.resize(device + 1, nullptr);
    if (!per_device[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return D4_ERR_HIP;
        per_device[device] = p;
    }
    *out = per_device[device];
    return D4_OK;
}

d4_ctx::~d4_ctx()
{
    if (batch) batch_passes_free(batch);
    if (lu) batch_lu_free(lu);
    for (auto* b : {&part, &tiles, &tiles_off, &xy, &v0, &v1, &tol, &coo, &coo_count, &cnt, &cursor, &Ap, &Ai, &Ax, &idx, &dvec,
                    &nw, &aux0, &aux1, &aux2, &tab_rec})
        b->release();
    for (auto& m : ev_marks) {
        ev_pool.push_back(m.first);
        ev_pool.push_back(m.second);
    }
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (pin) {
        (void)hipStreamSynchronize(stream);
        (void)hipHostFree(pin);
    }
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

int d4_ctx::fail(int code, const std::string& msg)
{
    err = msg;
    return code;
}

int d4_ctx::hip_fail(hipError_t e, const char* where)
{
    err = std::string(where) + ": " + hipGetErrorString(e);
    return D4_ERR_HIP;
}

int d4_ctx::ensure(DevBuf& b, size_t bytes, const char* what)
{
    if (b.bytes >= bytes) return D4_OK;
    b.release();
    size_t cap = std::max(bytes, (size_t)256);
    hipError_t e = hipMalloc(&b.p, cap);
    if (e != hipSuccess) {
        b.p = nullptr;
        return hip_fail(e, what);
    }
    b.bytes = cap;
    return D4_OK;
}

hipEvent_t d4_ctx::take_event()
{
    if (!ev_pool.empty()) {
        hipEvent_t e = ev_pool.back();
        ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int d4_ctx::mark_begin(hipEvent_t* e)
{
    *e = nullptr;
    if (!profiling) return D4_OK;
    *e = take_event();
    if (!*e) return fail(D4_ERR_HIP, "event create");
    hipError_t r = hipEventRecord(*e, stream);
    if (r != hipSuccess) return hip_fail(r, "event record");
    return D4_OK;
}

int d4_ctx::mark_end(hipEvent_t b)
{
    if (!profiling || !b) return D4_OK;
    hipEvent_t e = take_event();
    if (!e) return fail(D4_ERR_HIP, "event create");
    hipError_t r = hipEventRecord(e, stream);
    if (r != hipSuccess) return hip_fail(r, "event record");
    ev_marks.emplace_back(b, e);
    return D4_OK;
}

// Upper-triangle tile table of one realization of n (row-major, slot = index).
int d4_ctx::single_tiles(int n, int* ntiles, bool by_offset)
{
    const int nb = (n + kTile - 1) / kTile;
    const int nt = nb * (nb + 1) / 2;
    *ntiles = nt;
    DevBuf& buf = by_offset ? tiles_off : tiles;
    int& cached = by_offset ? tiles_off_n : tiles_n;
    if (cached == n) return D4_OK;
    // slot = the tile's row-major index in the upper triangle, whatever the launch order: partial sums
    // land in the same places and the reduction adds them in the same order, so both orders give the
    // same bits
    auto slot_of = [nb](int I, int J) { return I * nb - I * (I - 1) / 2 + (J - I); };
    std::vector<TileDesc> t;
    t.reserve(nt);
    if (!by_offset) {
        for (int I = 0; I < nb; ++I)
            for (int J = I; J < nb; ++J) t.push_back(TileDesc{0, I * kTile, J * kTile, slot_of(I, J)});
    } else {
        // tiles listed by their offset J - I, and the list cut into kXcds contiguous runs, run x taking
        // the launch positions p with p % kXcds == x (the order in which workgroups are handed to the
        // XCDs): each XCD walks its run in order, so the tiles it holds at once lie on one or two offsets
        std::vector<TileDesc> diag;
        diag.reserve(nt);
        for (int d = 0; d < nb; ++d)
            for (int I = 0; I + d < nb; ++I) diag.push_back(TileDesc{0, I * kTile, (I + d) * kTile, slot_of(I, I + d)});
        t.resize(nt);
        int start = 0;
        for (int x = 0; x < kXcds; ++x) {
            const int cnt = nt / kXcds + (x < nt % kXcds ? 1 : 0);
            for (int k = 0; k < cnt; ++k) t[(size_t)k * kXcds + x] = diag[(size_t)start + k];
            start += cnt;
        }
    }
    int rc = ensure(buf, sizeof(TileDesc) * (size_t)nt, "tile table");
    if (rc) return rc;
    hipError_t e = hipMemcpy(buf.p, t.data(), sizeof(TileDesc) * (size_t)nt, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "tile table upload");
    cached = n;
    return D4_OK;
}

// Runs one pair pass over a single realization whose positions are already on
// the device, plus the reduction.  Outputs are device pointers.
int d4_ctx::single_pass(int mode, int n, const double* d_xy, const RealParams& rp, double* d_tol, double* d_vA,
                        double* d_vB, double* d_q1, double* d_dvec, double* d_M, int64_t ld, const double* d_w,
                        int exact, int want_dy, int add_tau)
{
    int ntiles = 0;
    // the tabulated field's passes read cells at the pairs' separations: launched by tile offset, the
    // tiles an XCD holds at once read the same few table rows (positions sorted along y, as
    // SimulationData sorts them, give every tile of one offset about the same separation in y)
    const bool by_offset = mode == kModeSpeedTab || mode == kModeJacDenseTab;
    int rc = single_tiles(n, &ntiles, by_offset);
    if (rc) return rc;
    const int nq = pass_nq(mode);
    rc = ensure(part, sizeof(double) * 128 * (size_t)nq * (size_t)ntiles, "partials");
    if (rc) return rc;
    PassArgs a{};
    a.xy = reinterpret_cast<const double2*>(d_xy);
    a.off = nullptr;
    a.n0 = n;
    a.tiles = static_cast<const TileDesc*>(by_offset ? tiles_off.p : tiles.p);
    a.ntiles = ntiles;
    a.rp = nullptr;
    a.rp0 = rp;
    a.part = static_cast<double*>(part.p);
    a.tol = d_tol;
    a.M = d_M;
    a.ld = ld;
    a.coo = static_cast<CooEntry*>(coo.p);
    a.coo_count = static_cast<int*>(coo_count.p);
    a.coo_cap = (int)(coo.bytes / sizeof(CooEntry));
    a.w = d_w;
    a.tab_s = tab_s;
    a.tab_d = tab_d;
    a.tab_n = tab_n;
    a.exact = exact;
    a.want_dy = want_dy;
    hipEvent_t ev = nullptr;
    rc = mark_begin(&ev);
    if (rc) return rc;
    hipError_t e = launch_pair_pass(mode, a, stream);
    if (e != hipSuccess) return hip_fail(e, "pair pass launch");
    rc = mark_end(ev);
    if (rc) return rc;
    ReduceArgs r{};
    r.part = a.part;
    r.nq = nq;
    r.off = nullptr;
    r.rp = nullptr;
    r.rp0 = rp;
    r.R = 1;
    r.total = n;
    r.blocks = nullptr;
    r.nblocks = 0;
    r.vA = d_vA;
    r.vB = d_vB;
    r.q1 = d_q1;
    r.dvec = d_dvec;
    r.M = d_M;
    r.ld = ld;
    r.add_tau = add_tau;
    e = launch_reduce(r, stream);
    if (e != hipSuccess) return hip_fail(e, "reduce launch");
    return D4_OK;
}

int d4_ctx::sync(const char* where)
{
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        pending.clear();
        pin_used = 0;
        return hip_fail(e, where);
    }
    for (const Pending& p : pending) std::memcpy(p.dst, static_cast<char*>(pin) + p.off, p.bytes);
    pending.clear();
    pin_used = 0;
    return D4_OK;
}

void d4_ctx::stage_begin()
{
    if (pending.empty() && pin_used == 0) return;
    (void)hipStreamSynchronize(stream);  // a pass that failed midway: drop what it left
    pending.clear();
    pin_used = 0;
}

// A staged entry point's scope: it starts with an empty arena, and an error
// return (no successful sync() since) drops the downloads it queued, so that
// no later sync() copies them into the caller's arrays of this failed call.
struct StageScope {
    d4_ctx* c;
    explicit StageScope(d4_ctx* ctx) : c(ctx) { c->stage_begin(); }
    ~StageScope() { c->stage_begin(); }
};

int d4_ctx::stage(size_t bytes, size_t* off, const char* where)
{
    const size_t at = (pin_used + 255) & ~(size_t)255;
    if (at + bytes > pin_cap) {
        // grow once everything using the arena is done (the copies pending so far complete here)
        const int rc = sync(where);
        if (rc) return rc;
        const size_t cap = std::max<size_t>(2 * pin_cap, std::max<size_t>(bytes, 1 << 20));
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_cap = 0;
        hipError_t e = hipHostMalloc(&pin, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            pin = nullptr;
            return hip_fail(e, where);
        }
        pin_cap = cap;
        *off = 0;
        pin_used = bytes;
        return D4_OK;
    }
    *off = at;
    pin_used = at + bytes;
    return D4_OK;
}

int d4_ctx::h2d_s(void* dst, const void* src, size_t bytes, const char* where)
{
    if (!bytes) return D4_OK;
    size_t off = 0;
    const int rc = stage(bytes, &off, where);
    if (rc) return rc;
    std::memcpy(static_cast<char*>(pin) + off, src, bytes);
    hipError_t e = hipMemcpyAsync(dst, static_cast<char*>(pin) + off, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, where);
    return D4_OK;
}

int d4_ctx::d2h_s(void* dst, const void* src, size_t bytes, const char* where)
{
    if (!bytes) return D4_OK;
    size_t off = 0;
    const int rc = stage(bytes, &off, where);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(static_cast<char*>(pin) + off, src, bytes, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return hip_fail(e, where);
    pending.push_back({dst, off, bytes});
    return D4_OK;
}

int d4_ctx::h2d(void* dst, const void* src, size_t bytes, const char* where)
{
    if (!bytes) return D4_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, where);
    return D4_OK;
}

int d4_ctx::d2h(void* dst, const void* src, size_t bytes, const char* where)
{
    if (!bytes) return D4_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return hip_fail(e, where);
    return D4_OK;
}

#define D4_TRY(x)                 \
    do {                          \
        int rc_ = (x);            \
        if (rc_) return rc_;      \
    } while (0)

extern "C" {

int d4_ctx_create(int device, d4_ctx** out)
{
    if (!out) return D4_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return D4_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return D4_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return D4_ERR_HIP;
    d4_ctx* c = new d4_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return D4_ERR_HIP;
    }
    c->stream = c->own_stream;
    if (c->ensure(c->coo_count, sizeof(int) * 4, "coo counter") || noalloc_scratch_ptr(device, &c->noalloc_scratch)) {
        delete c;
        return D4_ERR_HIP;
    }
    *out = c;
    return D4_OK;
}

void d4_ctx_destroy(d4_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
}

const char* d4_last_error(const d4_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int d4_ctx_device(const d4_ctx* ctx) { return ctx ? ctx->device : -1; }

int d4_ctx_set_stream(d4_ctx* ctx, void* stream)
{
    if (!ctx) return D4_ERR_INVALID;
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return D4_OK;
}

int d4_ctx_use_default_stream(d4_ctx* ctx)
{
    if (!ctx) return D4_ERR_INVALID;
    ctx->stream = nullptr;
    return D4_OK;
}

int d4_ctx_synchronize(d4_ctx* ctx)
{
    if (!ctx) return D4_ERR_INVALID;
    return ctx->sync("synchronize");
}

int d4_ctx_set_profiling(d4_ctx* ctx, int on)
{
    if (!ctx) return D4_ERR_INVALID;
    ctx->profiling = on != 0;
    return D4_OK;
}

int d4_ctx_pair_kernel_time(d4_ctx* ctx, double* total_ms, int* launches)
{
    if (!ctx || !total_ms || !launches) return D4_ERR_INVALID;
    double tot = 0.0;
    int cnt = 0;
    for (auto& m : ctx->ev_marks) {
        hipError_t r = hipEventSynchronize(m.second);
        if (r != hipSuccess) return ctx->hip_fail(r, "event sync");
        float ms = 0.f;
        r = hipEventElapsedTime(&ms, m.first, m.second);
        if (r != hipSuccess) return ctx->hip_fail(r, "event elapsed");
        tot += ms;
        cnt++;
        ctx->ev_pool.push_back(m.first);
        ctx->ev_pool.push_back(m.second);
    }
    ctx->ev_marks.clear();
    *total_ms = tot;
    *launches = cnt;
    return D4_OK;
}

int d4_dev_alloc(d4_ctx* ctx, int64_t bytes, void** out)
{
    if (!ctx || !out || bytes < 0) return D4_ERR_INVALID;
    hipError_t e = hipMalloc(out, (size_t)std::max<int64_t>(bytes, 8));
    if (e != hipSuccess) return ctx->hip_fail(e, "d4_dev_alloc");
    return D4_OK;
}

int d4_dev_free(d4_ctx* ctx, void* p)
{
    if (!ctx) return D4_ERR_INVALID;
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return ctx->hip_fail(e, "d4_dev_free");
    return D4_OK;
}

int d4_memcpy_h2d(d4_ctx* ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || bytes < 0) return D4_ERR_INVALID;
    D4_TRY(ctx->h2d(dst, src, (size_t)bytes, "d4_memcpy_h2d"));
    return ctx->sync("d4_memcpy_h2d");
}

int d4_memcpy_d2h(d4_ctx* ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || bytes < 0) return D4_ERR_INVALID;
    D4_TRY(ctx->d2h(dst, src, (size_t)bytes, "d4_memcpy_d2h"));
    return ctx->sync("d4_memcpy_d2h");
}

int d4_field_eval_dev(d4_ctx* ctx, int64_t n, const double* dx, const double* dy, const double* e2pix,
                      const double* c2piy, double* sxy, double* dsxy, int flags)
{
    if (!ctx || n < 0 || (n > 0 && (!dx || !dy))) return D4_ERR_INVALID;
    if ((e2pix == nullptr) != (c2piy == nullptr)) return ctx->fail(D4_ERR_INVALID, "give both e2pix and c2piy or neither");
    hipError_t e = launch_field_eval(n, dx, dy, e2pix, c2piy, sxy, dsxy, flags & 1, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "field eval launch");
    return D4_OK;
}

int d4_field_eval(d4_ctx* ctx, int64_t n, const double* dx, const double* dy, const double* e2pix,
                  const double* c2piy, double* sxy, double* dsxy, int flags)
{
    if (!ctx || n < 0 || (n > 0 && (!dx || !dy))) return D4_ERR_INVALID;
    if ((e2pix == nullptr) != (c2piy == nullptr)) return ctx->fail(D4_ERR_INVALID, "give both e2pix and c2piy or neither");
    if (n == 0) return D4_OK;
    const size_t b = sizeof(double) * (size_t)n;
    D4_TRY(ctx->ensure(ctx->aux0, 6 * b, "field eval buffers"));
    double* d = static_cast<double*>(ctx->aux0.p);
    double *ddx = d, *ddy = d + n, *de = d + 2 * n, *dc = d + 3 * n, *ds = d + 4 * n, *dd = d + 5 * n;
    D4_TRY(ctx->h2d(ddx, dx, b, "field eval upload"));
    D4_TRY(ctx->h2d(ddy, dy, b, "field eval upload"));
    if (e2pix) {
        D4_TRY(ctx->h2d(de, e2pix, b, "field eval upload"));
        D4_TRY(ctx->h2d(dc, c2piy, b, "field eval upload"));
    }
    D4_TRY(d4_field_eval_dev(ctx, n, ddx, ddy, e2pix ? de : nullptr, e2pix ? dc : nullptr, sxy ? ds : nullptr,
                             dsxy ? dd : nullptr, flags));
    if (sxy) D4_TRY(ctx->d2h(sxy, ds, b, "field eval download"));
    if (dsxy) D4_TRY(ctx->d2h(dsxy, dd, b, "field eval download"));
    return ctx->sync("field eval");
}

int d4_speed_pass_dev(d4_ctx* ctx, int n, const double* xy, double tau, double* v, double* tol, double tol_cap2)
{
    if (!ctx || n < 0 || (n > 0 && (!xy || !v))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    RealParams rp = real_params(n, nullptr, tau);
    rp.tol_cap2 = tol ? tol_cap2 : 0.0;
    return ctx->single_pass(kModeSpeed, n, xy, rp, tol, v, nullptr, nullptr, nullptr, nullptr, 0, nullptr, 0, 0, 1);
}

int d4_speed_pass(d4_ctx* ctx, int n, const double* xy, double tau, double* v, double* tol, double tol_cap2)
{
    if (!ctx || n < 0 || (n > 0 && (!xy || !v))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    const size_t b = sizeof(double) * (size_t)n;
    StageScope staged(ctx);
    D4_TRY(ctx->ensure(ctx->xy, 2 * b, "positions"));
    D4_TRY(ctx->ensure(ctx->v0, b, "speeds"));
    double* dtol = nullptr;
    if (tol && tol_cap2 > 0) {
        D4_TRY(ctx->ensure(ctx->tol, b, "tolerance"));
        dtol = static_cast<double*>(ctx->tol.p);
        D4_TRY(ctx->h2d_s(dtol, tol, b, "tolerance upload"));
    }
    D4_TRY(ctx->h2d_s(ctx->xy.p, xy, 2 * b, "position upload"));
    D4_TRY(d4_speed_pass_dev(ctx, n, static_cast<double*>(ctx->xy.p), tau, static_cast<double*>(ctx->v0.p), dtol,
                             tol_cap2));
    D4_TRY(ctx->d2h_s(v, ctx->v0.p, b, "speed download"));
    if (dtol) D4_TRY(ctx->d2h_s(tol, dtol, b, "tolerance download"));
    return ctx->sync("speed pass");
}

int d4_speed_pass_dev_noalloc(d4_ctx* ctx, int n, const double* xy, double tau, double* v, double* tol,
                              double tol_cap2)
{
    if (!ctx || n < 0 || (n > 0 && (!xy || !v))) return D4_ERR_INVALID;
    if (n > D4_NOALLOC_MAX_N)
        return ctx->fail(D4_ERR_INVALID, "d4_speed_pass_dev_noalloc: n = " + std::to_string(n) + " exceeds " +
                                             std::to_string(D4_NOALLOC_MAX_N) + ", the size of its scratch");
    if (n == 0) return D4_OK;
    if (!ctx->noalloc_scratch) return ctx->fail(D4_ERR_HIP, "d4_speed_pass_dev_noalloc: no scratch on this device");
    // the same tiles, slots and kernels as d4_speed_pass_dev, so the same bits
    TileDesc* tiles = static_cast<TileDesc*>(ctx->noalloc_scratch);
    double* part = static_cast<double*>(ctx->noalloc_scratch) + kNoAllocTileDoubles;
    const int nb = (n + kTile - 1) / kTile;
    RealParams rp = real_params(n, nullptr, tau);
    rp.tol_cap2 = tol ? tol_cap2 : 0.0;
    hipError_t e = launch_tile_table(tiles, n, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "noalloc tile table");
    PassArgs a{};
    a.xy = reinterpret_cast<const double2*>(xy);
    a.n0 = n;
    a.tiles = tiles;
    a.ntiles = nb * (nb + 1) / 2;
    a.rp0 = rp;
    a.part = part;
    a.tol = tol;
    e = launch_pair_pass(kModeSpeed, a, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "noalloc speed pass");
    ReduceArgs r{};
    r.part = part;
    r.nq = 1;
    r.rp0 = rp;
    r.R = 1;
    r.total = n;
    r.vA = v;
    r.add_tau = 1;
    e = launch_reduce(r, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "noalloc reduce");
    return D4_OK;
}

int d4_jacobian_dense_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                          double* M, int64_t ld, double* dvec, double* tol)
{
    if (!ctx || !p || n < 0 || (n > 0 && (!xy || !vA || !M || ld < n))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    RealParams rp = real_params(n, p, 0.0);
    if (!tol) rp.tol_cap2 = 0.0;
    return ctx->single_pass(kModeJacDense, n, xy, rp, tol, vA, vB, nullptr, dvec, M, ld, nullptr, p->exact ? 1 : 0, 0,
                            1);
}

int d4_jacobian_dense_cut_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                              double* M, int64_t ld, double* dvec, double* tol)
{
    if (!ctx || !p || n < 0 || (n > 0 && (!xy || !vA || !M || ld < n))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    RealParams rp = real_params(n, p, 0.0);
    if (!tol) rp.tol_cap2 = 0.0;
    return ctx->single_pass(kModeJacDenseCut, n, xy, rp, tol, vA, vB, nullptr, dvec, M, ld, nullptr, p->exact ? 1 : 0,
                            0, 1);
}

int d4_theta_pair_vjp_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, const double* a,
                          const double* c, const double* u, const double* w, const double* t, double* gx, double* gy)
{
    if (!ctx || !p || n < 0 || (n > 0 && (!xy || !a || !c || !u || !w || !t || !gx))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    ThetaVjpArgs va{};
    va.n = n;
    va.xy = reinterpret_cast<const double2*>(xy);
    va.rp = real_params(n, p, 0.0);
    va.a = a;
    va.c = c;
    va.u = u;
    va.w = w;
    va.t = t;
    va.gx = gx;
    va.gy = gy;
    hipError_t e = launch_theta_pair_vjp(va, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "implicit-step pair VJP launch");
    return D4_OK;
}

int d4_jacobian_dense_scale_dev(d4_ctx* ctx, int n, double* M, int64_t ld, const double* dvec)
{
    if (!ctx || n < 0 || (n > 0 && (!M || !dvec || ld < n))) return D4_ERR_INVALID;
    hipError_t e = launch_dense_scale(M, ld, n, dvec, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "dense scale launch");
    return D4_OK;
}

int d4_jacobian_dense(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                      double* M, double* dvec, double* tol, int scale)
{
    if (!ctx || !p || n < 0 || (n > 0 && (!xy || !vA || !M))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    const size_t b = sizeof(double) * (size_t)n;
    D4_TRY(ctx->ensure(ctx->xy, 2 * b, "positions"));
    D4_TRY(ctx->ensure(ctx->v0, b, "speeds A"));
    D4_TRY(ctx->ensure(ctx->v1, b, "speeds B"));
    D4_TRY(ctx->ensure(ctx->dvec, b, "weights"));
    D4_TRY(ctx->ensure(ctx->aux1, b * (size_t)n, "dense matrix"));
    double* dtol = nullptr;
    if (tol && p->tol_cap2 > 0) {
        D4_TRY(ctx->ensure(ctx->tol, b, "tolerance"));
        dtol = static_cast<double*>(ctx->tol.p);
        D4_TRY(ctx->h2d(dtol, tol, b, "tolerance upload"));
    }
    D4_TRY(ctx->h2d(ctx->xy.p, xy, 2 * b, "position upload"));
    double* dM = static_cast<double*>(ctx->aux1.p);
    double* dd = static_cast<double*>(ctx->dvec.p);
    D4_TRY(d4_jacobian_dense_dev(ctx, n, static_cast<double*>(ctx->xy.p), p, static_cast<double*>(ctx->v0.p),
                                 static_cast<double*>(ctx->v1.p), dM, n, dd, dtol));
    if (scale) D4_TRY(d4_jacobian_dense_scale_dev(ctx, n, dM, n, dd));
    D4_TRY(ctx->d2h(vA, ctx->v0.p, b, "speed download"));
    if (vB) D4_TRY(ctx->d2h(vB, ctx->v1.p, b, "speed download"));
    D4_TRY(ctx->d2h(M, dM, b * (size_t)n, "matrix download"));
    if (dvec) D4_TRY(ctx->d2h(dvec, dd, b, "weight download"));
    if (dtol) D4_TRY(ctx->d2h(tol, dtol, b, "tolerance download"));
    return ctx->sync("dense jacobian pass");
}

int d4_jacobian_pass(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, double* vA, double* vB,
                     int* Ap, int* Ai, double* Ax, int cap, int* nnz, int* indexes, double* dvec, double* tol)
{
    if (!ctx || !p || n < 0 || !nnz) return D4_ERR_INVALID;
    if (n > 0 && (!xy || !vA || !vB || !Ap || !indexes || !dvec)) return D4_ERR_INVALID;
    if (p->wfunc != 'c' && p->wfunc != 'p' && p->wfunc != 'm') return ctx->fail(D4_ERR_INVALID, "wfunc must be c, p or m");
    *nnz = 0;
    if (n == 0) return D4_OK;
    const size_t b = sizeof(double) * (size_t)n;
    StageScope staged(ctx);
    D4_TRY(ctx->ensure(ctx->xy, 2 * b, "positions"));
    D4_TRY(ctx->ensure(ctx->v0, b, "speeds A"));
    D4_TRY(ctx->ensure(ctx->v1, b, "speeds B"));
    D4_TRY(ctx->ensure(ctx->dvec, b, "weights"));
    D4_TRY(ctx->ensure(ctx->idx, sizeof(int) * (size_t)n, "diagonal indexes"));
    D4_TRY(ctx->ensure(ctx->cnt, sizeof(int) * (size_t)n, "column counts"));
    D4_TRY(ctx->ensure(ctx->cursor, sizeof(int) * (size_t)n, "column cursors"));
    D4_TRY(ctx->ensure(ctx->Ap, sizeof(int) * ((size_t)n + 1), "Ap"));
    // Pair-list capacity: from the caller's CSC capacity, grown on overflow.  The
    // device CSC holds n + 2 * capacity entries (each listed pair gives at most
    // two, beside the n diagonals), so every write of the assembly stays inside
    // it whatever the pass finds; the one synchronization at the end tells
    // whether the list overflowed (pass run again) or the caller's CSC is too
    // small (D4_ERR_CAPACITY with *nnz set, as before).
    const size_t max_pairs = (size_t)n * (size_t)(n - 1) / 2;  // bounds for a caller capacity far above need
    size_t coo_cap = std::max<size_t>(std::min<size_t>((size_t)std::max(cap, n) / 2 + 1024, max_pairs + 1024),
                                      ctx->coo.bytes / sizeof(CooEntry));
    double* dtol = nullptr;
    if (tol && p->tol_cap2 > 0) {
        D4_TRY(ctx->ensure(ctx->tol, b, "tolerance"));
        dtol = static_cast<double*>(ctx->tol.p);
    }
    D4_TRY(ctx->h2d_s(ctx->xy.p, xy, 2 * b, "position upload"));
    RealParams rp = real_params(n, p, 0.0);
    if (!dtol) rp.tol_cap2 = 0.0;
    int* dcnt = static_cast<int*>(ctx->cnt.p);
    int* dAp = static_cast<int*>(ctx->Ap.p);
    int* didx = static_cast<int*>(ctx->idx.p);
    double* dd = static_cast<double*>(ctx->dvec.p);
    const int* dcount = static_cast<const int*>(ctx->coo_count.p);
    // CSC entries returned to the caller: its capacity, at most n^2 (and inside the device CSC, which holds
    // n + 2 * coo_cap >= min(cap, n^2) entries)
    const size_t down = (Ai && Ax && cap > 0) ? std::min<size_t>((size_t)cap, (size_t)n * (size_t)n) : 0;
    int ncoo = 0, total_nnz = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        D4_TRY(ctx->ensure(ctx->coo, sizeof(CooEntry) * coo_cap, "pair list"));
        const size_t csc_cap = (size_t)n + 2 * coo_cap;
        D4_TRY(ctx->ensure(ctx->Ai, sizeof(int) * csc_cap, "Ai"));
        D4_TRY(ctx->ensure(ctx->Ax, sizeof(double) * csc_cap, "Ax"));
        int* dAi = static_cast<int*>(ctx->Ai.p);
        double* dAx = static_cast<double*>(ctx->Ax.p);
        const CooEntry* dcoo = static_cast<const CooEntry*>(ctx->coo.p);
        // the pass updates the tolerance in place; on a second attempt the caller's array already holds
        // the first attempt's result, and the update (a minimum over the same pairs) leaves it unchanged
        if (dtol) D4_TRY(ctx->h2d_s(dtol, tol, b, "tolerance upload"));
        hipError_t e = hipMemsetAsync(ctx->coo_count.p, 0, sizeof(int), ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "counter reset");
        D4_TRY(ctx->single_pass(kModeJacCoo, n, static_cast<double*>(ctx->xy.p), rp, dtol,
                                static_cast<double*>(ctx->v0.p), static_cast<double*>(ctx->v1.p), nullptr, nullptr,
                                nullptr, 0, nullptr, p->exact ? 1 : 0, 0, 1));
        CooDerivArgs da{};
        da.xy = static_cast<const double2*>(ctx->xy.p);
        da.n0 = n;
        da.rp0 = rp;
        da.coo = static_cast<CooEntry*>(ctx->coo.p);
        da.count = dcount;
        da.cap = (int)coo_cap;
        e = launch_coo_derivative(da, p->exact ? 1 : 0, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "pair list derivative");
        // CSC assembly: counts (diagonal + both copies), scan, fill, sort/finish, scale
        e = launch_fill_int(dcnt, 1, n, ctx->stream);  // the diagonal of every column
        if (e == hipSuccess) e = launch_csc_count(dcoo, (int)coo_cap, dcnt, n, ctx->stream, dcount);
        if (e == hipSuccess) e = launch_scan(dcnt, dAp, n, ctx->stream, static_cast<int*>(ctx->cursor.p));
        if (e == hipSuccess)
            e = launch_csc_fill(dcoo, (int)coo_cap, dAp, static_cast<int*>(ctx->cursor.p), dAi, dAx, n, ctx->stream,
                                dcount);
        if (e == hipSuccess) e = launch_csc_finish(dAp, dAi, dAx, didx, dd, nullptr, nullptr, p->wfunc, n, ctx->stream);
        if (e == hipSuccess) e = launch_csc_scale(dAp, dAi, dAx, didx, dd, n, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "csc assembly");
        D4_TRY(ctx->d2h_s(&ncoo, dcount, sizeof(int), "pair count"));
        D4_TRY(ctx->d2h_s(&total_nnz, dAp + n, sizeof(int), "nnz"));
        D4_TRY(ctx->d2h_s(vA, ctx->v0.p, b, "speed download"));
        D4_TRY(ctx->d2h_s(vB, ctx->v1.p, b, "speed download"));
        D4_TRY(ctx->d2h_s(Ap, dAp, sizeof(int) * ((size_t)n + 1), "Ap download"));
        if (down) {
            D4_TRY(ctx->d2h_s(Ai, dAi, sizeof(int) * down, "Ai download"));
            D4_TRY(ctx->d2h_s(Ax, dAx, sizeof(double) * down, "Ax download"));
        }
        D4_TRY(ctx->d2h_s(indexes, didx, sizeof(int) * (size_t)n, "indexes download"));
        D4_TRY(ctx->d2h_s(dvec, dd, b, "weights download"));
        if (dtol) D4_TRY(ctx->d2h_s(tol, dtol, b, "tolerance download"));
        D4_TRY(ctx->sync("jacobian pass"));
        if ((size_t)ncoo <= coo_cap) break;
        if (attempt == 1) return ctx->fail(D4_ERR_HIP, "pair list overflow after resize");
        coo_cap = (size_t)ncoo + 1024;
    }
    *nnz = total_nnz;
    if (total_nnz > cap || !Ai || !Ax) {
        ctx->err = "CSC capacity too small";
        return D4_ERR_CAPACITY;
    }
    return D4_OK;
}

int d4_jacobian_rescale(d4_ctx* ctx, int n, const int* Ap, const int* Ai, double* Ax, double* dvec, char wfunc,
                        double factor, double dtau, double* speed)
{
    if (!ctx || n < 0 || (n > 0 && (!Ap || !Ai || !Ax || !dvec))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    const int nnz = Ap[n];
    const size_t b = sizeof(double) * (size_t)n;
    StageScope staged(ctx);
    D4_TRY(ctx->ensure(ctx->Ap, sizeof(int) * ((size_t)n + 1), "Ap"));
    D4_TRY(ctx->ensure(ctx->Ai, sizeof(int) * (size_t)std::max(nnz, 1), "Ai"));
    D4_TRY(ctx->ensure(ctx->Ax, sizeof(double) * (size_t)std::max(nnz, 1), "Ax"));
    D4_TRY(ctx->ensure(ctx->dvec, b, "weights"));
    D4_TRY(ctx->ensure(ctx->nw, b, "new weights"));
    D4_TRY(ctx->ensure(ctx->v0, b, "speeds"));
    D4_TRY(ctx->h2d_s(ctx->Ap.p, Ap, sizeof(int) * ((size_t)n + 1), "Ap upload"));
    D4_TRY(ctx->h2d_s(ctx->Ai.p, Ai, sizeof(int) * (size_t)nnz, "Ai upload"));
    D4_TRY(ctx->h2d_s(ctx->Ax.p, Ax, sizeof(double) * (size_t)nnz, "Ax upload"));
    D4_TRY(ctx->h2d_s(ctx->dvec.p, dvec, b, "weights upload"));
    if (speed) D4_TRY(ctx->h2d_s(ctx->v0.p, speed, b, "speed upload"));
    RescaleArgs a{};
    a.Ap = static_cast<const int*>(ctx->Ap.p);
    a.Ai = static_cast<const int*>(ctx->Ai.p);
    a.Ax = static_cast<double*>(ctx->Ax.p);
    a.dvec = static_cast<double*>(ctx->dvec.p);
    a.nw = static_cast<double*>(ctx->nw.p);
    a.speed = speed ? static_cast<double*>(ctx->v0.p) : nullptr;
    a.factor0 = factor;
    a.dtau0 = dtau;
    a.wf0 = wfunc;
    a.total = n;
    hipError_t e = launch_rescale(a, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "rescale launch");
    D4_TRY(ctx->d2h_s(Ax, ctx->Ax.p, sizeof(double) * (size_t)nnz, "Ax download"));
    D4_TRY(ctx->d2h_s(dvec, ctx->dvec.p, b, "weights download"));
    if (speed) D4_TRY(ctx->d2h_s(speed, ctx->v0.p, b, "speed download"));
    return ctx->sync("rescale");
}

int d4_vjp_dev(d4_ctx* ctx, int n, const double* xy, const double* w, double* gx, double* gy, double* gtau, int flags)
{
    if (!ctx || n < 0 || (n > 0 && (!xy || !w || !gx))) return D4_ERR_INVALID;
    if (n == 0) {
        if (gtau) *gtau = 0;
        return D4_OK;
    }
    RealParams rp = real_params(n, nullptr, 0.0);
    D4_TRY(ctx->single_pass(kModeVjp, n, xy, rp, nullptr, gx, nullptr, gy, nullptr, nullptr, 0, w, flags & 1,
                            gy != nullptr, 0));
    if (gtau) {  // dL/dtau = sum_i b_i w_i: host-side sum over a downloaded w
        std::vector<double> hw(n);
        D4_TRY(ctx->d2h(hw.data(), w, sizeof(double) * (size_t)n, "w download"));
        D4_TRY(ctx->sync("vjp"));
        double s = 0;
        for (int i = 0; i < n; ++i) s += (i < n / 2 ? 1.0 : -1.0) * hw[i];
        *gtau = s;
    }
    return D4_OK;
}

static bool table_size_ok(int n) { return n >= 2 && n <= (1 << 14) && (n & (n - 1)) == 0; }

int d4_table_build_dev(d4_ctx* ctx, int n, double* tab_s, double* tab_d)
{
    if (!ctx || !tab_s || !table_size_ok(n)) return D4_ERR_INVALID;
    hipError_t e = launch_table_build(n, tab_s, tab_d, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "table build");
    return D4_OK;
}

int d4_table_eval_dev(d4_ctx* ctx, int64_t m, const double* dx, const double* dy, const double* tab_s,
                      const double* tab_d, int n, double* s, double* d)
{
    if (!ctx || m < 0 || !table_size_ok(n) || (m > 0 && (!dx || !dy)) || (s && !tab_s) || (d && !tab_d))
        return D4_ERR_INVALID;
    hipError_t e = launch_table_eval(m, dx, dy, tab_s, tab_d, n, s, d, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "table eval");
    return D4_OK;
}

int d4_speed_pass_table_dev(d4_ctx* ctx, int n, const double* xy, const double* tab_s, int tab_n, double tau,
                            double* v)
{
    if (!ctx || n < 0 || !table_size_ok(tab_n) || (n > 0 && (!xy || !v || !tab_s))) return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    // the table as cell records (one 32-byte record per pair instead of four scattered nodes)
    D4_TRY(ctx->ensure(ctx->tab_rec, sizeof(double) * 4 * (size_t)tab_n * tab_n, "table cell records"));
    hipError_t e = launch_table_records(tab_n, tab_s, nullptr, static_cast<double*>(ctx->tab_rec.p), ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "table cell records");
    ctx->tab_s = static_cast<const double*>(ctx->tab_rec.p);
    ctx->tab_d = nullptr;
    ctx->tab_n = tab_n;
    RealParams rp = real_params(n, nullptr, tau);
    const int rc = ctx->single_pass(kModeSpeedTab, n, xy, rp, nullptr, v, nullptr, nullptr, nullptr, nullptr, 0,
                                    nullptr, 0, 0, 1);
    ctx->tab_s = nullptr;
    return rc;
}

int d4_jacobian_dense_table_dev(d4_ctx* ctx, int n, const double* xy, const d4_jac_params* p, const double* tab_s,
                                const double* tab_d, int tab_n, double* vA, double* vB, double* M, int64_t ld,
                                double* dvec)
{
    if (!ctx || !p || n < 0 || !table_size_ok(tab_n) || (n > 0 && (!xy || !vA || !M || ld < n || !tab_s || !tab_d)))
        return D4_ERR_INVALID;
    if (n == 0) return D4_OK;
    // sigma's and d sigma's corners of a cell in one 64-byte record
    D4_TRY(ctx->ensure(ctx->tab_rec, sizeof(double) * 8 * (size_t)tab_n * tab_n, "table cell records"));
    hipError_t e = launch_table_records(tab_n, tab_s, tab_d, static_cast<double*>(ctx->tab_rec.p), ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "table cell records");
    ctx->tab_s = static_cast<const double*>(ctx->tab_rec.p);
    ctx->tab_d = nullptr;
    ctx->tab_n = tab_n;
    RealParams rp = real_params(n, p, 0.0);
    rp.tol_cap2 = 0.0;
    const int rc = ctx->single_pass(kModeJacDenseTab, n, xy, rp, nullptr, vA, vB, nullptr, dvec, M, ld, nullptr, 0,
                                    0, 1);
    ctx->tab_s = ctx->tab_d = nullptr;
    return rc;
}

}  // extern "C"
