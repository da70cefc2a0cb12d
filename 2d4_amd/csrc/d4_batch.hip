// d4_batch.hip -- batched ragged passes (include/d4_batch.h) and the lockstep
// ensemble of host simulations that runs every pair pass of a step once for
// all of its realizations.
//
// Layout in HBM: positions of all realizations concatenated ({x, y} per
// dislocation, each realization sorted as SimulationData sorts it), per-
// realization RealParams, one tile table over all realizations (a tile never
// straddles two), partial sums in per-realization slot ranges, and the step
// matrices as one block-diagonal CSC with global indices.
#include "../../include/d4_batch.h"
#include "d4_ctx.h"
#include "d4_host.h"
#include "d4_solver.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using namespace d4;

namespace {

// pinned host memory (hipHostMalloc), grow-only
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t b)
    {
        if (bytes >= b) return true;
        release();
        b = std::max(b, (size_t)256);
        if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b;
        return true;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~HostBuf() { release(); }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// Fixed set of workers; parallel_for hands out indices through an atomic counter.
class ThreadPool {
public:
    explicit ThreadPool(int n)
    {
        for (int i = 0; i < n - 1; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~ThreadPool()
    {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    int size() const { return (int)workers_.size() + 1; }
    void parallel_for(int count, const std::function<void(int)>& fn)
    {
        if (count <= 0) return;
        if (workers_.empty() || count == 1) {
            for (int i = 0; i < count; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> l(m_);
            fn_ = &fn;
            count_ = count;
            next_.store(0);
            pending_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [this] { return pending_ == 0; });
        fn_ = nullptr;
    }

private:
    void work()
    {
        for (;;) {
            const int i = next_.fetch_add(1);
            if (i >= count_) break;
            (*fn_)(i);
        }
    }
    void loop()
    {
        d4host::t_host_worker = true;
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
            }
            work();
            std::lock_guard<std::mutex> l(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    int count_ = 0;
    std::atomic<int> next_{0};
    int pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int default_threads()
{
    const int hw = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(hw > 0 ? hw : 1, 16));
}

}  // namespace

#define D4_TRY(x)            \
    do {                     \
        int rc_ = (x);       \
        if (rc_) return rc_; \
    } while (0)

// Tables, buffers and captured graphs of the batched passes over one set of
// realizations (sizes n_r given by off).
struct BatchPasses {
    explicit BatchPasses(d4_ctx* c) : ctx(c) {}
    ~BatchPasses()
    {
        drop_graphs();
        for (auto* b : {&off_d, &elem_d, &wf_d, &tiles_d, &blocks_d, &rp_d, &xy_d, &vA_d, &vB_d, &tol_d, &part_d,
                        &coo_d, &ncoo_d, &cnt_d, &cursor_d, &Ap_d, &Ai_d, &Ax_d, &idx_d, &dvec_d, &nw_d, &fac_d,
                        &dtau_d})
            b->release();
    }

    d4_ctx* ctx;
    d4_ens_config cfg{0, 0, 0, 0};
    int R = 0, total = 0, maxn = 0, ntiles = 0, nslots = 0, nblocks = 0, tiles_order = -1;
    std::vector<int> off;
    DevBuf off_d, elem_d, wf_d, tiles_d, blocks_d, rp_d, xy_d, vA_d, vB_d, tol_d, part_d, coo_d, ncoo_d, cnt_d,
        cursor_d, Ap_d, Ai_d, Ax_d, idx_d, dvec_d, nw_d, fac_d, dtau_d;
    // staging: the passes read inputs from and write outputs to these
    HostBuf xy_h, rp_h, wf_h, tolin_h, tol_h, vA_h, vB_h, Ap_h, Ai_h, Ax_h, idx_h, dvec_h, ncoo_h, fac_h, dtau_h;
    size_t coo_cap = 0;
    hipGraphExec_t g_speed[2] = {nullptr, nullptr}, g_jac[2] = {nullptr, nullptr};  // [with tolerance]
    uint64_t csc_gen = 0;  // counts the passes that rewrote the device CSC (Ap_d, Ai_d, Ax_d)

    void drop_graphs()
    {
        for (auto* g : {&g_speed[0], &g_speed[1], &g_jac[0], &g_jac[1]}) {
            if (*g) (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
    }

    int fail(int code, const std::string& m) { return ctx->fail(code, m); }

    int set_config(const d4_ens_config& c)
    {
        if (c.pass_grid < 0 || c.tile_order < 0 || c.tile_order > 2) return fail(D4_ERR_INVALID, "bad ensemble config");
        const bool retile = c.tile_order != cfg.tile_order;
        if (c.pass_grid != cfg.pass_grid || c.use_graph != cfg.use_graph || retile) drop_graphs();
        cfg = c;
        if (retile && R > 0) return build_tiles();
        return D4_OK;
    }

    int slot0(int r) const
    {
        int s = 0;
        for (int q = 0; q < r; ++q) {
            const int nb = (off[q + 1] - off[q] + kTile - 1) / kTile;
            s += nb * (nb + 1) / 2;
        }
        return s;
    }

    int build_tiles()
    {
        std::vector<int> nbs(R), s0(R);
        int s = 0;
        for (int r = 0; r < R; ++r) {
            nbs[r] = (off[r + 1] - off[r] + kTile - 1) / kTile;
            s0[r] = s;
            s += nbs[r] * (nbs[r] + 1) / 2;
        }
        std::vector<std::vector<TileDesc>> per(R);
        for (int r = 0; r < R; ++r) {
            int slot = s0[r];
            for (int I = 0; I < nbs[r]; ++I)
                for (int J = I; J < nbs[r]; ++J) per[r].push_back(TileDesc{r, I * kTile, J * kTile, slot++});
        }
        std::vector<TileDesc> t;
        t.reserve(s);
        std::vector<int> order(R);
        std::iota(order.begin(), order.end(), 0);
        if (cfg.tile_order == 1)
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nbs[a] > nbs[b]; });
        if (cfg.tile_order == 2) {
            size_t longest = 0;
            for (auto& v : per) longest = std::max(longest, v.size());
            for (size_t k = 0; k < longest; ++k)
                for (int r : order)
                    if (k < per[r].size()) t.push_back(per[r][k]);
        } else {
            for (int r : order) t.insert(t.end(), per[r].begin(), per[r].end());
        }
        ntiles = (int)t.size();
        D4_TRY(ctx->ensure(tiles_d, sizeof(TileDesc) * std::max<size_t>(t.size(), 1), "batch tile table"));
        hipError_t e = hipMemcpy(tiles_d.p, t.data(), sizeof(TileDesc) * t.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return ctx->hip_fail(e, "batch tile upload");
        tiles_order = cfg.tile_order;
        return D4_OK;
    }

    // (re)builds the tables when the realization sizes change
    int set_layout(int R_, const int* off_)
    {
        if (R_ == R && std::equal(off.begin(), off.end(), off_, off_ + R_ + 1)) return D4_OK;
        if (R_ <= 0) return fail(D4_ERR_INVALID, "no realizations");
        for (int r = 0; r < R_; ++r)
            if (off_[r + 1] < off_[r] || off_[0] != 0) return fail(D4_ERR_INVALID, "offsets must start at 0 and not decrease");
        drop_graphs();
        R = R_;
        off.assign(off_, off_ + R + 1);
        total = off[R];
        maxn = 0;
        std::vector<int> elem(std::max(total, 1));
        std::vector<int2> blocks;
        nslots = 0;
        for (int r = 0; r < R; ++r) {
            const int n = off[r + 1] - off[r];
            maxn = std::max(maxn, n);
            for (int k = off[r]; k < off[r + 1]; ++k) elem[k] = r;
            const int nb = (n + kTile - 1) / kTile;
            for (int b = 0; b < nb; ++b) blocks.push_back(make_int2(r, b));
            nslots += nb * (nb + 1) / 2;
        }
        nblocks = (int)blocks.size();
        const size_t T = (size_t)std::max(total, 1);
        D4_TRY(ctx->ensure(off_d, sizeof(int) * (R + 1), "batch offsets"));
        D4_TRY(ctx->ensure(elem_d, sizeof(int) * T, "batch element table"));
        D4_TRY(ctx->ensure(blocks_d, sizeof(int2) * std::max<size_t>(blocks.size(), 1), "batch reduce blocks"));
        D4_TRY(ctx->ensure(rp_d, sizeof(RealParams) * R, "batch parameters"));
        D4_TRY(ctx->ensure(wf_d, sizeof(int) * R, "batch weight kinds"));
        D4_TRY(ctx->ensure(fac_d, sizeof(double) * R, "batch factors"));
        D4_TRY(ctx->ensure(dtau_d, sizeof(double) * R, "batch stress steps"));
        D4_TRY(ctx->ensure(xy_d, 2 * sizeof(double) * T, "batch positions"));
        for (DevBuf* b : {&vA_d, &vB_d, &tol_d, &dvec_d, &nw_d})
            D4_TRY(ctx->ensure(*b, sizeof(double) * T, "batch vectors"));
        for (DevBuf* b : {&cnt_d, &cursor_d, &idx_d}) D4_TRY(ctx->ensure(*b, sizeof(int) * T, "batch columns"));
        D4_TRY(ctx->ensure(Ap_d, sizeof(int) * (T + 1), "batch Ap"));
        D4_TRY(ctx->ensure(ncoo_d, sizeof(int) * 4, "batch pair count"));
        D4_TRY(ctx->ensure(part_d, sizeof(double) * 256 * (size_t)std::max(nslots, 1), "batch partials"));
        bool ok = xy_h.ensure(2 * sizeof(double) * T) && rp_h.ensure(sizeof(RealParams) * R) &&
                  wf_h.ensure(sizeof(int) * R) && tolin_h.ensure(sizeof(double) * T) && tol_h.ensure(sizeof(double) * T) &&
                  vA_h.ensure(sizeof(double) * T) && vB_h.ensure(sizeof(double) * T) &&
                  Ap_h.ensure(sizeof(int) * (T + 1)) && idx_h.ensure(sizeof(int) * T) &&
                  dvec_h.ensure(sizeof(double) * T) && ncoo_h.ensure(sizeof(int) * 4) &&
                  fac_h.ensure(sizeof(double) * R) && dtau_h.ensure(sizeof(double) * R);
        if (!ok) return fail(D4_ERR_HIP, "pinned host allocation failed");
        hipError_t e = hipMemcpy(off_d.p, off.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(elem_d.p, elem.data(), sizeof(int) * T, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(blocks_d.p, blocks.data(), sizeof(int2) * blocks.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return ctx->hip_fail(e, "batch table upload");
        return build_tiles();
    }

    // per-realization parameters into rp_h, with the layout's slot ranges
    RealParams* params() { return rp_h.as<RealParams>(); }
    void finish_params()
    {
        RealParams* rp = params();
        int s = 0;
        for (int r = 0; r < R; ++r) {
            const int nb = (off[r + 1] - off[r] + kTile - 1) / kTile;
            rp[r].nb = nb;
            rp[r].slot0 = s;
            s += nb * (nb + 1) / 2;
        }
    }

    PassArgs pass_args(bool with_tol)
    {
        PassArgs a{};
        a.xy = static_cast<const double2*>(xy_d.p);
        a.off = static_cast<const int*>(off_d.p);
        a.n0 = 0;
        a.tiles = static_cast<const TileDesc*>(tiles_d.p);
        a.ntiles = ntiles;
        a.rp = static_cast<const RealParams*>(rp_d.p);
        a.part = static_cast<double*>(part_d.p);
        a.tol = with_tol ? static_cast<double*>(tol_d.p) : nullptr;
        a.coo = static_cast<CooEntry*>(coo_d.p);
        a.coo_count = static_cast<int*>(ncoo_d.p);
        a.coo_cap = (int)coo_cap;
        return a;
    }

    ReduceArgs reduce_args(int nq)
    {
        ReduceArgs r{};
        r.part = static_cast<const double*>(part_d.p);
        r.nq = nq;
        r.off = static_cast<const int*>(off_d.p);
        r.rp = static_cast<const RealParams*>(rp_d.p);
        r.R = R;
        r.total = total;
        r.blocks = static_cast<const int2*>(blocks_d.p);
        r.nblocks = nblocks;
        r.add_tau = 1;
        return r;
    }

    int h2d(DevBuf& d, const HostBuf& h, size_t bytes, const char* w) { return ctx->h2d(d.p, h.p, bytes, w); }
    int d2h(HostBuf& h, const DevBuf& d, size_t bytes, const char* w) { return ctx->d2h(h.p, d.p, bytes, w); }

    int enqueue_speed(bool with_tol)
    {
        const size_t T = total;
        D4_TRY(h2d(xy_d, xy_h, 2 * sizeof(double) * T, "batch position upload"));
        D4_TRY(h2d(rp_d, rp_h, sizeof(RealParams) * R, "batch parameter upload"));
        if (with_tol) D4_TRY(h2d(tol_d, tolin_h, sizeof(double) * T, "batch tolerance upload"));
        hipError_t e = launch_pair_pass(kModeSpeed, pass_args(with_tol), ctx->stream, cfg.pass_grid);
        if (e != hipSuccess) return ctx->hip_fail(e, "batch speed pass");
        ReduceArgs r = reduce_args(1);
        r.vA = static_cast<double*>(vA_d.p);
        e = launch_reduce(r, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "batch reduce");
        D4_TRY(d2h(vA_h, vA_d, sizeof(double) * T, "batch speed download"));
        if (with_tol) D4_TRY(d2h(tol_h, tol_d, sizeof(double) * T, "batch tolerance download"));
        return D4_OK;
    }

    int enqueue_jacobian(bool with_tol)
    {
        const size_t T = total;
        hipStream_t s = ctx->stream;
        D4_TRY(h2d(xy_d, xy_h, 2 * sizeof(double) * T, "batch position upload"));
        D4_TRY(h2d(rp_d, rp_h, sizeof(RealParams) * R, "batch parameter upload"));
        D4_TRY(h2d(wf_d, wf_h, sizeof(int) * R, "batch weight kinds upload"));
        if (with_tol) D4_TRY(h2d(tol_d, tolin_h, sizeof(double) * T, "batch tolerance upload"));
        int* ncoo = static_cast<int*>(ncoo_d.p);
        hipError_t e = launch_fill_int(ncoo, 0, 1, s);
        if (e == hipSuccess) e = launch_pair_pass(kModeJacCoo, pass_args(with_tol), s, cfg.pass_grid);
        if (e == hipSuccess) {
            CooDerivArgs da{};
            da.xy = static_cast<const double2*>(xy_d.p);
            da.off = static_cast<const int*>(off_d.p);
            da.elem = static_cast<const int*>(elem_d.p);
            da.rp = static_cast<const RealParams*>(rp_d.p);
            da.coo = static_cast<CooEntry*>(coo_d.p);
            da.count = ncoo;
            da.cap = (int)coo_cap;
            e = launch_coo_derivative(da, 0, s);
        }
        if (e != hipSuccess) return ctx->hip_fail(e, "batch jacobian pass");
        ReduceArgs r = reduce_args(1);
        r.vA = static_cast<double*>(vA_d.p);
        r.vB = static_cast<double*>(vB_d.p);
        e = launch_reduce(r, s);
        int* cnt = static_cast<int*>(cnt_d.p);
        int* Ap = static_cast<int*>(Ap_d.p);
        int* Ai = static_cast<int*>(Ai_d.p);
        double* Ax = static_cast<double*>(Ax_d.p);
        int* idx = static_cast<int*>(idx_d.p);
        double* dv = static_cast<double*>(dvec_d.p);
        const CooEntry* coo = static_cast<const CooEntry*>(coo_d.p);
        if (e == hipSuccess) e = launch_fill_int(cnt, 1, total, s);
        if (e == hipSuccess) e = launch_csc_count(coo, (int)coo_cap, cnt, total, s, ncoo);
        if (e == hipSuccess) e = launch_scan(cnt, Ap, total, s, static_cast<int*>(cursor_d.p));  // cursors: free until the fill
        if (e == hipSuccess)
            e = launch_csc_fill(coo, (int)coo_cap, Ap, static_cast<int*>(cursor_d.p), Ai, Ax, total, s, ncoo);
        if (e == hipSuccess)
            e = launch_csc_finish(Ap, Ai, Ax, idx, dv, static_cast<const int*>(elem_d.p),
                                  static_cast<const int*>(wf_d.p), 'c', total, s);
        if (e == hipSuccess) e = launch_csc_scale(Ap, Ai, Ax, idx, dv, total, s);
        if (e != hipSuccess) return ctx->hip_fail(e, "batch csc assembly");
        D4_TRY(d2h(vA_h, vA_d, sizeof(double) * T, "batch speed download"));
        D4_TRY(d2h(vB_h, vB_d, sizeof(double) * T, "batch speed download"));
        D4_TRY(d2h(Ap_h, Ap_d, sizeof(int) * (T + 1), "batch Ap download"));
        D4_TRY(d2h(idx_h, idx_d, sizeof(int) * T, "batch indexes download"));
        D4_TRY(d2h(dvec_h, dvec_d, sizeof(double) * T, "batch weights download"));
        D4_TRY(d2h(ncoo_h, ncoo_d, sizeof(int), "batch pair count download"));
        if (with_tol) D4_TRY(d2h(tol_h, tol_d, sizeof(double) * T, "batch tolerance download"));
        return D4_OK;
    }

    // runs `enqueue` directly, or from a graph captured on first use
    template <class F>
    int run(hipGraphExec_t& g, F enqueue)
    {
        if (cfg.use_graph && ctx->stream) {
            if (!g) {
                hipGraph_t graph = nullptr;
                hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return ctx->hip_fail(e, "graph capture begin");
                const int rc = enqueue();
                e = hipStreamEndCapture(ctx->stream, &graph);
                if (rc) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                if (e != hipSuccess) return ctx->hip_fail(e, "graph capture end");
                e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (e != hipSuccess) {
                    g = nullptr;
                    return ctx->hip_fail(e, "graph instantiate");
                }
            }
            hipError_t e = hipGraphLaunch(g, ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "graph launch");
            return D4_OK;
        }
        return enqueue();
    }

    int ensure_coo(size_t want)
    {
        // every pair listed gives at most two CSC entries beside the total diagonals
        const size_t cap = std::max(want, coo_cap), nz = (size_t)total + 2 * cap;
        if (coo_cap == cap && Ai_d.bytes >= sizeof(int) * nz && Ax_d.bytes >= sizeof(double) * nz) return D4_OK;
        drop_graphs();
        coo_cap = cap;
        D4_TRY(ctx->ensure(coo_d, sizeof(CooEntry) * coo_cap, "batch pair list"));
        D4_TRY(ctx->ensure(Ai_d, sizeof(int) * nz, "batch Ai"));
        D4_TRY(ctx->ensure(Ax_d, sizeof(double) * nz, "batch Ax"));
        return D4_OK;
    }

    // speed pass over the staged positions/parameters (xy_h, rp_h, tolin_h) -> vA_h, tol_h
    int speed(bool with_tol)
    {
        finish_params();
        D4_TRY(run(g_speed[with_tol], [&] { return enqueue_speed(with_tol); }));
        return ctx->sync("batch speed pass");
    }

    // Jacobian pass -> vA_h, vB_h, Ap_h, idx_h, dvec_h, tol_h and (fetch_entries) Ai_h and
    // (fetch_values) Ax_h
    int jacobian(bool with_tol, bool fetch_entries, int* nnz_out, bool fetch_values = true)
    {
        ++csc_gen;
        finish_params();
        // the device CSC must hold total + 2 * coo_cap entries for this layout's total, which may have grown
        // since the list was last sized (a later call over more realizations)
        D4_TRY(ensure_coo(std::max(coo_cap, std::max<size_t>((size_t)total * 8, 1 << 16))));
        for (int attempt = 0; attempt < 2; ++attempt) {
            D4_TRY(run(g_jac[with_tol], [&] { return enqueue_jacobian(with_tol); }));
            D4_TRY(ctx->sync("batch jacobian pass"));
            const size_t ncoo = (size_t)ncoo_h.as<int>()[0];
            if (ncoo <= coo_cap) break;
            if (attempt == 1) return fail(D4_ERR_HIP, "batch pair list overflow after resize");
            D4_TRY(ensure_coo(ncoo + ncoo / 4 + 1024));
        }
        const int nnz = Ap_h.as<int>()[total];
        if (nnz_out) *nnz_out = nnz;
        if (fetch_entries) {
            if (!Ai_h.ensure(sizeof(int) * (size_t)nnz) ||
                (fetch_values && !Ax_h.ensure(sizeof(double) * (size_t)nnz)))
                return fail(D4_ERR_HIP, "pinned host allocation failed");
            D4_TRY(d2h(Ai_h, Ai_d, sizeof(int) * (size_t)nnz, "batch Ai download"));
            if (fetch_values) D4_TRY(d2h(Ax_h, Ax_d, sizeof(double) * (size_t)nnz, "batch Ax download"));
            D4_TRY(ctx->sync("batch entries download"));
        }
        return D4_OK;
    }

    // stage-II rescale of the staged block-diagonal CSC (Ap_h, Ai_h, Ax_h global),
    // dvec_h, speeds vA_h, factors fac_h, stress steps dtau_h, weight kinds wf_h;
    // upload_csc = false: the device CSC already holds exactly that matrix
    int rescale(int nnz, bool upload_csc, bool fetch_values)
    {
        ++csc_gen;
        const size_t T = total;
        if (upload_csc) {
            if (!Ai_d.p || Ai_d.bytes < sizeof(int) * (size_t)nnz || Ax_d.bytes < sizeof(double) * (size_t)nnz) {
                D4_TRY(ctx->ensure(Ai_d, sizeof(int) * (size_t)std::max(nnz, 1), "batch Ai"));
                D4_TRY(ctx->ensure(Ax_d, sizeof(double) * (size_t)std::max(nnz, 1), "batch Ax"));
                drop_graphs();
            }
            D4_TRY(h2d(Ap_d, Ap_h, sizeof(int) * (T + 1), "batch Ap upload"));
            D4_TRY(h2d(Ai_d, Ai_h, sizeof(int) * (size_t)nnz, "batch Ai upload"));
            D4_TRY(h2d(Ax_d, Ax_h, sizeof(double) * (size_t)nnz, "batch Ax upload"));
        }
        D4_TRY(h2d(dvec_d, dvec_h, sizeof(double) * T, "batch weights upload"));
        D4_TRY(h2d(vA_d, vA_h, sizeof(double) * T, "batch speed upload"));
        D4_TRY(h2d(fac_d, fac_h, sizeof(double) * R, "batch factor upload"));
        D4_TRY(h2d(dtau_d, dtau_h, sizeof(double) * R, "batch stress upload"));
        D4_TRY(h2d(wf_d, wf_h, sizeof(int) * R, "batch weight kinds upload"));
        RescaleArgs a{};
        a.Ap = static_cast<const int*>(Ap_d.p);
        a.Ai = static_cast<const int*>(Ai_d.p);
        a.Ax = static_cast<double*>(Ax_d.p);
        a.dvec = static_cast<double*>(dvec_d.p);
        a.nw = static_cast<double*>(nw_d.p);
        a.speed = static_cast<double*>(vA_d.p);
        a.elem_real = static_cast<const int*>(elem_d.p);
        a.off = static_cast<const int*>(off_d.p);
        a.factor_arr = static_cast<const double*>(fac_d.p);
        a.dtau_arr = static_cast<const double*>(dtau_d.p);
        a.wf_arr = static_cast<const int*>(wf_d.p);
        a.total = total;
        hipError_t e = launch_rescale(a, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "batch rescale");
        if (fetch_values) D4_TRY(d2h(Ax_h, Ax_d, sizeof(double) * (size_t)nnz, "batch Ax download"));
        D4_TRY(d2h(dvec_h, dvec_d, sizeof(double) * T, "batch weights download"));
        D4_TRY(d2h(vA_h, vA_d, sizeof(double) * T, "batch speed download"));
        return ctx->sync("batch rescale");
    }
};

void batch_passes_free(BatchPasses* b) { delete b; }

static BatchPasses* ctx_batch(d4_ctx* ctx)
{
    if (!ctx->batch) ctx->batch = new BatchPasses(ctx);
    return ctx->batch;
}

// ------------------------------------------------------------ ensemble

namespace sdddstCore {
int apply_sim_params(SimulationData& d, const d4_sim_params* p);  // d4_sim.cpp
}

struct d4_ens {
    d4_ctx* ctx = nullptr;
    std::vector<std::shared_ptr<sdddstCore::SimulationData>> data;
    std::vector<std::unique_ptr<sdddstCore::Simulation>> sims;
    std::vector<int> off;
    std::unique_ptr<BatchPasses> bp;
    std::unique_ptr<ThreadPool> pool;
    std::unique_ptr<d4host::BatchDenseLU> blu;  // the solver-1 realizations' LUs, batched
    // the realizations (in pass order) and column offsets of the last batched pass that left their CSC on
    // the device, and that pass's generation: still there while bp->csc_gen has not moved on
    uint64_t csc_gen = 0;
    std::vector<int> csc_act, csc_off;
    // per realization: its host CSC values (SimulationData::Ax) were not fetched, the device CSC being
    // the copy its GPU solver and the rescale read
    std::vector<char> ax_stale;
    d4_ens_config cfg{0, 0, 0, 0};
    bool started = false;
    std::string err;
    double t_pass = 0, t_solve = 0, t_host = 0, t_total = 0;

    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }
    int check(int rc)
    {
        if (rc && err.empty()) err = d4_last_error(ctx);
        return rc;
    }
};

namespace {

using sdddstCore::PassJob;
using sdddstCore::Simulation;

void pool_for(void* user, int count, void (*fn)(void*, int), void* arg)
{
    static_cast<ThreadPool*>(user)->parallel_for(count, [&](int k) { fn(arg, k); });
}

// Factor / Solve jobs: the realizations with the GPU dense solver through one
// BatchDenseLU call (bitwise what each one's own solver gives), the others
// each with its own solver on the pool
int run_linear(d4_ens* e, const std::vector<int>& act, PassJob::Kind kind)
{
    const double t0 = now_ms();
    std::vector<int> gpu, own;
    for (int r : act) (e->data[r]->solverKind == 1 ? gpu : own).push_back(r);
    std::vector<int> rcs(own.size(), D4_OK);
    e->pool->parallel_for((int)own.size(), [&](int i) {
        Simulation& s = *e->sims[own[i]];
        sdddstCore::PassEngine eng;
        rcs[i] = eng.run(s, s.job());
    });
    for (int rc : rcs)
        if (rc) return e->fail(rc, kind == PassJob::Factor ? "linear solver factorization failed" : "linear solve failed");
    if (!gpu.empty()) {
        if (!e->blu) {
            e->blu.reset(d4host::new_batch_dense_lu(d4_ctx_device(e->ctx)));
            if (!e->blu) return e->fail(D4_ERR_HIP, "batched GPU LU unavailable");
            e->blu->set_parallel(pool_for, e->pool.get());
        }
        const int K = (int)gpu.size();
        if (kind == PassJob::Factor) {
            std::vector<d4host::BatchDenseLU::Matrix> m(K);
            // the matrices are still on the device when this group's Jacobian or rescale pass was the last to
            // write the device CSC (it precedes the Factor job): the solver reads them there
            const bool on_dev = e->bp && e->csc_gen == e->bp->csc_gen && e->csc_act == act;
            if (!on_dev)
                for (int r : gpu)
                    if (e->ax_stale[r]) return e->fail(D4_ERR_HIP, "ensemble: CSC values no longer on the device");
            std::vector<int> pos;
            if (on_dev) {
                pos.assign(e->data.size(), -1);
                for (int k = 0; k < (int)act.size(); ++k) pos[act[k]] = k;
            }
            for (int k = 0; k < K; ++k) {
                const sdddstCore::SimulationData& d = *e->data[gpu[k]];
                m[k] = {(int)d.dc, d.Ap.data(), d.Ai.data(), d.Ax.data()};
                m[k].key = gpu[k];  // the realization: its order follows its own sequence of patterns
                m[k].symmetric = true;  // a step matrix's pattern (V is stored iff V^T is)
                if (on_dev) {
                    const int q = pos[gpu[k]];
                    m[k].dAp = static_cast<const int*>(e->bp->Ap_d.p) + e->csc_off[q];
                    m[k].dAi = static_cast<const int*>(e->bp->Ai_d.p);
                    m[k].dAx = static_cast<const double*>(e->bp->Ax_d.p);
                    m[k].row_base = e->csc_off[q];
                }
            }
            std::vector<unsigned char> sing(K, 0), unpiv(K, 0);
            if (e->blu->factor(K, m.data(), sing.data(), unpiv.data()) < 0)
                return e->fail(D4_ERR_HIP, "batched GPU LU factorization failed");
            for (int k = 0; k < K; ++k) {
                PassJob& j = e->sims[gpu[k]]->job();
                j.singular = sing[k] != 0;
                j.unpivoted = unpiv[k] != 0;
            }
        } else {
            std::vector<const double*> b(K);
            std::vector<double*> x(K);
            for (int k = 0; k < K; ++k) {
                PassJob& j = e->sims[gpu[k]]->job();
                b[k] = j.rhs->data();
                x[k] = j.sol->data();
            }
            if (e->blu->solve(K, b.data(), x.data()) < 0) return e->fail(D4_ERR_HIP, "batched GPU LU solve failed");
        }
    }
    e->t_solve += now_ms() - t0;
    return D4_OK;
}

// stages the jobs of the simulations `act` (all of one kind), runs them as one
// batched pass, and hands each simulation its results
int run_batched(d4_ens* e, const std::vector<int>& act)
{
    if (e->sims[act[0]]->job().kind == PassJob::Factor || e->sims[act[0]]->job().kind == PassJob::Solve)
        return run_linear(e, act, e->sims[act[0]]->job().kind);

    BatchPasses& bp = *e->bp;
    const int R = (int)act.size();
    std::vector<int> off(R + 1, 0);
    for (int k = 0; k < R; ++k) off[k + 1] = off[k] + (int)e->data[act[k]]->dc;
    int rc = bp.set_layout(R, off.data());
    if (rc) return rc;
    const PassJob::Kind kind = e->sims[act[0]]->job().kind;
    bool with_tol = false;
    for (int k = 0; k < R; ++k) with_tol |= e->sims[act[k]]->precision().toleranceData() != nullptr;
    double2* xy = bp.xy_h.as<double2>();
    RealParams* rp = bp.params();
    double* tolin = bp.tolin_h.as<double>();
    const double t0 = now_ms();

    if (kind == PassJob::Speed || kind == PassJob::Jacobian) {
        e->pool->parallel_for(R, [&](int k) {
            Simulation& s = *e->sims[act[k]];
            PassJob& j = s.job();
            const int n = off[k + 1] - off[k];
            std::memcpy(xy + off[k], j.xy->data(), sizeof(double2) * n);
            const double* tol = s.precision().toleranceData();
            if (with_tol) {
                if (tol) std::memcpy(tolin + off[k], tol, sizeof(double) * n);
                else std::fill(tolin + off[k], tolin + off[k + 1], 0.0);
            }
            if (kind == PassJob::Speed) {
                rp[k] = real_params(n, nullptr, j.tau);
                rp[k].tol_cap2 = tol ? s.precision().toleranceCap2() : 0.0;
            } else {
                rp[k] = real_params(n, &j.jp, 0.0);
                if (!tol) rp[k].tol_cap2 = 0.0;
                bp.wf_h.as<int>()[k] = j.jp.wfunc;
            }
        });
        // the host solver's realizations need the values on the host; the GPU solver's read them on the device
        bool host_values = false;
        for (int k = 0; k < R; ++k) host_values |= e->data[act[k]]->solverKind != 1;
        if (kind == PassJob::Speed) {
            rc = bp.speed(with_tol);
        } else {
            int nnz = 0;
            rc = bp.jacobian(with_tol, true, &nnz, host_values);
            e->csc_gen = bp.csc_gen;
            e->csc_act = act;
            e->csc_off = off;
        }
        if (rc) return rc;
        const int* Ap = bp.Ap_h.as<int>();
        e->pool->parallel_for(R, [&](int k) {
            Simulation& s = *e->sims[act[k]];
            PassJob& j = s.job();
            sdddstCore::SimulationData& d = s.data();
            const int n = off[k + 1] - off[k];
            s.pairPasses++;
            double* tol = s.precision().toleranceData();
            if (tol) std::memcpy(tol, bp.tol_h.as<double>() + off[k], sizeof(double) * n);
            if (kind == PassJob::Speed) {
                std::memcpy(j.v->data(), bp.vA_h.as<double>() + off[k], sizeof(double) * n);
                return;
            }
            std::memcpy(j.vA->data(), bp.vA_h.as<double>() + off[k], sizeof(double) * n);
            std::memcpy(j.vB->data(), bp.vB_h.as<double>() + off[k], sizeof(double) * n);
            std::memcpy(d.dVec.data(), bp.dvec_h.as<double>() + off[k], sizeof(double) * n);
            const int base = Ap[off[k]];
            const int nnz = Ap[off[k + 1]] - base;
            if ((int)d.Ai.size() < nnz) {  // grown as the plain path grows it (simulation_data.cpp:157-175)
                d.Ai.resize((size_t)nnz + n);
                d.Ax.resize(d.Ai.size());
            }
            for (int c = 0; c <= n; ++c) d.Ap[c] = Ap[off[k] + c] - base;
            const int* Ai = bp.Ai_h.as<int>() + base;
            for (int m = 0; m < nnz; ++m) d.Ai[m] = Ai[m] - off[k];
            e->ax_stale[act[k]] = d.solverKind == 1;
            if (d.solverKind != 1) std::memcpy(d.Ax.data(), bp.Ax_h.as<double>() + base, sizeof(double) * nnz);
            const int* idx = bp.idx_h.as<int>() + off[k];
            for (int c = 0; c < n; ++c) d.indexes[c] = idx[c] - base;
            j.nnz = nnz;
        });
    } else if (kind == PassJob::Rescale) {
        // the group's Jacobian pass left this very CSC on the device (only speed passes ran since, and they
        // do not touch it): rescale it there instead of staging it from the host again
        const bool on_dev = e->csc_gen == bp.csc_gen && e->csc_act == act;
        bool host_values = false;
        for (int k = 0; k < R; ++k) {
            host_values |= e->data[act[k]]->solverKind != 1;
            if (!on_dev && e->ax_stale[act[k]])
                return e->fail(D4_ERR_HIP, "ensemble: CSC values no longer on the device");
        }
        std::vector<int> base(R + 1, 0);
        for (int k = 0; k < R; ++k) base[k + 1] = base[k] + e->data[act[k]]->Ap[e->data[act[k]]->dc];
        const int nnz = base[R];
        if (!bp.Ai_h.ensure(sizeof(int) * (size_t)std::max(nnz, 1)) ||
            !bp.Ax_h.ensure(sizeof(double) * (size_t)std::max(nnz, 1)))
            return e->fail(D4_ERR_HIP, "pinned host allocation failed");
        int* Ap = bp.Ap_h.as<int>();
        e->pool->parallel_for(R, [&](int k) {
            Simulation& s = *e->sims[act[k]];
            PassJob& j = s.job();
            sdddstCore::SimulationData& d = s.data();
            const int n = off[k + 1] - off[k];
            if (!on_dev) {
                for (int c = 0; c < n; ++c) Ap[off[k] + c] = base[k] + d.Ap[c];
                if (k == R - 1) Ap[off[R]] = nnz;
                int* Ai = bp.Ai_h.as<int>() + base[k];
                const int m = base[k + 1] - base[k];
                for (int q = 0; q < m; ++q) Ai[q] = d.Ai[q] + off[k];
                std::memcpy(bp.Ax_h.as<double>() + base[k], d.Ax.data(), sizeof(double) * m);
            }
            std::memcpy(bp.dvec_h.as<double>() + off[k], d.dVec.data(), sizeof(double) * n);
            std::memcpy(bp.vA_h.as<double>() + off[k], j.speed->data(), sizeof(double) * n);
            bp.fac_h.as<double>()[k] = j.factor;
            bp.dtau_h.as<double>()[k] = j.dtau;
            bp.wf_h.as<int>()[k] = d.weightFunc;
        });
        rc = bp.rescale(nnz, !on_dev, host_values);
        if (rc) return rc;
        e->csc_gen = bp.csc_gen;
        e->csc_act = act;
        e->csc_off = off;
        e->pool->parallel_for(R, [&](int k) {
            Simulation& s = *e->sims[act[k]];
            PassJob& j = s.job();
            sdddstCore::SimulationData& d = s.data();
            const int n = off[k + 1] - off[k];
            e->ax_stale[act[k]] = d.solverKind == 1;
            if (d.solverKind != 1)
                std::memcpy(d.Ax.data(), bp.Ax_h.as<double>() + base[k], sizeof(double) * (base[k + 1] - base[k]));
            std::memcpy(d.dVec.data(), bp.dvec_h.as<double>() + off[k], sizeof(double) * n);
            std::memcpy(j.speed->data(), bp.vA_h.as<double>() + off[k], sizeof(double) * n);
        });
    }
    e->t_pass += now_ms() - t0;
    return D4_OK;
}

// every simulation in `act` has a job of one kind: run them, then the host work
int run_phases(d4_ens* e, std::vector<int>& act)
{
    // largest realizations first, so the pool's last tasks are short ones
    std::stable_sort(act.begin(), act.end(), [&](int a, int b) { return e->data[a]->dc > e->data[b]->dc; });
    while (!act.empty() && e->sims[act[0]]->job().kind != PassJob::None) {
        int rc = run_batched(e, act);
        if (rc) {
            for (int r : act) e->sims[r]->check(rc, "batched pass");
            return e->check(rc);
        }
        const double t0 = now_ms();
        e->pool->parallel_for((int)act.size(), [&](int k) { e->sims[act[k]]->after_pass(); });
        e->t_host += now_ms() - t0;
        for (int r : act)
            if (e->sims[r]->status()) return e->fail(e->sims[r]->status(), e->sims[r]->error());
    }
    return D4_OK;
}

}  // namespace

extern "C" {

int d4_batch_speed_pass(d4_ctx* ctx, int R, const int* off, const double* xy, const double* tau, double* v,
                        double* tol, const double* tol_cap2)
{
    if (!ctx || R <= 0 || !off || !xy || !tau || !v || (tol && !tol_cap2)) return D4_ERR_INVALID;
    BatchPasses& bp = *ctx_batch(ctx);
    D4_TRY(bp.set_layout(R, off));
    const int T = off[R];
    std::memcpy(bp.xy_h.p, xy, sizeof(double) * 2 * (size_t)T);
    bool with_tol = false;
    for (int r = 0; r < R; ++r) {
        bp.params()[r] = real_params(off[r + 1] - off[r], nullptr, tau[r]);
        bp.params()[r].tol_cap2 = tol ? tol_cap2[r] : 0.0;
        with_tol |= tol && tol_cap2[r] > 0;
    }
    if (with_tol) std::memcpy(bp.tolin_h.p, tol, sizeof(double) * (size_t)T);
    D4_TRY(bp.speed(with_tol));
    std::memcpy(v, bp.vA_h.p, sizeof(double) * (size_t)T);
    if (with_tol) std::memcpy(tol, bp.tol_h.p, sizeof(double) * (size_t)T);
    return D4_OK;
}

int d4_batch_jacobian_pass(d4_ctx* ctx, int R, const int* off, const double* xy, const d4_jac_params* p,
                           double* vA, double* vB, int* Ap, int* Ai, double* Ax, int cap, int* nnz, int* indexes,
                           double* dvec, double* tol)
{
    if (!ctx || R <= 0 || !off || !xy || !p || !vA || !vB || !Ap || !nnz || !indexes || !dvec) return D4_ERR_INVALID;
    for (int r = 0; r < R; ++r)
        if (p[r].wfunc != 'c' && p[r].wfunc != 'p' && p[r].wfunc != 'm')
            return ctx->fail(D4_ERR_INVALID, "wfunc must be c, p or m");
    BatchPasses& bp = *ctx_batch(ctx);
    D4_TRY(bp.set_layout(R, off));
    const int T = off[R];
    std::memcpy(bp.xy_h.p, xy, sizeof(double) * 2 * (size_t)T);
    bool with_tol = false;
    for (int r = 0; r < R; ++r) {
        bp.params()[r] = real_params(off[r + 1] - off[r], &p[r], 0.0);
        if (!tol) bp.params()[r].tol_cap2 = 0.0;
        with_tol |= tol && p[r].tol_cap2 > 0;
        bp.wf_h.as<int>()[r] = p[r].wfunc;
    }
    if (with_tol) std::memcpy(bp.tolin_h.p, tol, sizeof(double) * (size_t)T);
    int total_nnz = 0;
    D4_TRY(bp.jacobian(with_tol, false, &total_nnz));
    *nnz = total_nnz;
    if (total_nnz > cap || !Ai || !Ax) return ctx->fail(D4_ERR_CAPACITY, "CSC capacity too small");
    D4_TRY(ctx->d2h(Ai, bp.Ai_d.p, sizeof(int) * (size_t)total_nnz, "batch Ai download"));
    D4_TRY(ctx->d2h(Ax, bp.Ax_d.p, sizeof(double) * (size_t)total_nnz, "batch Ax download"));
    D4_TRY(ctx->sync("batch entries download"));
    std::memcpy(vA, bp.vA_h.p, sizeof(double) * (size_t)T);
    std::memcpy(vB, bp.vB_h.p, sizeof(double) * (size_t)T);
    std::memcpy(Ap, bp.Ap_h.p, sizeof(int) * ((size_t)T + 1));
    std::memcpy(indexes, bp.idx_h.p, sizeof(int) * (size_t)T);
    std::memcpy(dvec, bp.dvec_h.p, sizeof(double) * (size_t)T);
    if (with_tol) std::memcpy(tol, bp.tol_h.p, sizeof(double) * (size_t)T);
    return D4_OK;
}

int d4_ens_create(d4_ctx* ctx, int R, const int* off, const double* x, const double* y, const double* b,
                  const d4_sim_params* p, int params_per_realization, d4_ens** out)
{
    if (!ctx || R <= 0 || !off || !x || !y || !b || !p || !out) return D4_ERR_INVALID;
    *out = nullptr;
    if (off[0] != 0) return ctx->fail(D4_ERR_INVALID, "offsets must start at 0");
    for (int r = 0; r < R; ++r)
        if (off[r + 1] <= off[r]) return ctx->fail(D4_ERR_INVALID, "every realization needs dislocations");
    std::unique_ptr<d4_ens> e(new d4_ens());
    e->ctx = ctx;
    e->off.assign(off, off + R + 1);
    for (int r = 0; r < R; ++r) {
        const int n = off[r + 1] - off[r];
        auto d = std::make_shared<sdddstCore::SimulationData>(n, x + off[r], y + off[r], b + off[r]);
        if (!d->error.empty()) return ctx->fail(D4_ERR_INVALID, "realization " + std::to_string(r) + ": " + d->error);
        const int rc = sdddstCore::apply_sim_params(*d, params_per_realization ? &p[r] : p);
        if (rc) return ctx->fail(rc, "realization " + std::to_string(r) + ": invalid parameters");
        e->sims.emplace_back(new sdddstCore::Simulation(d, ctx));
        e->data.push_back(std::move(d));
    }
    e->bp.reset(new BatchPasses(ctx));
    e->pool.reset(new ThreadPool(default_threads()));
    e->ax_stale.assign(e->data.size(), 0);
    *out = e.release();
    return D4_OK;
}

void d4_ens_destroy(d4_ens* e) { delete e; }

const char* d4_ens_last_error(const d4_ens* e) { return e ? e->err.c_str() : "null ensemble"; }

int d4_ens_step(d4_ens* e, int nsteps, int* nactive)
{
    if (!e || nsteps < 0) return D4_ERR_INVALID;
    const double t0 = now_ms();
    const int R = (int)e->sims.size();
    if (!e->started) {
        std::vector<int> all(R);
        std::iota(all.begin(), all.end(), 0);
        for (int r : all) e->sims[r]->begin_prologue();
        D4_TRY(run_phases(e, all));
        e->started = true;
    }
    int last = 0;
    for (int s = 0; s < nsteps; ++s) {
        std::vector<int> act;
        for (int r = 0; r < R; ++r)
            if (e->sims[r]->begin_step()) act.push_back(r);
        last = (int)act.size();
        if (act.empty()) break;
        D4_TRY(run_phases(e, act));
    }
    if (nactive) *nactive = last;
    e->t_total += now_ms() - t0;
    return D4_OK;
}

int d4_ens_positions(const d4_ens* e, double* xy, int original_order)
{
    if (!e || !xy) return D4_ERR_INVALID;
    for (size_t r = 0; r < e->data.size(); ++r) {
        const auto& d = *e->data[r];
        double* o = xy + 2 * (size_t)e->off[r];
        for (unsigned i = 0; i < d.dc; ++i) {
            const unsigned k = original_order ? d.disl_order[i] : i;
            o[2 * i] = d.disl_sorted[k].x;
            o[2 * i + 1] = d.disl_sorted[k].y;
        }
    }
    return D4_OK;
}

int d4_ens_stats(const d4_ens* e, d4_sim_stats* st)
{
    if (!e || !st) return D4_ERR_INVALID;
    for (size_t r = 0; r < e->data.size(); ++r) {
        const auto& d = *e->data[r];
        const auto& s = *e->sims[r];
        st[r].sim_time = d.simTime;
        st[r].step_size = d.stepSize;
        st[r].strain = d.totalAccumulatedStrainIncrease;
        st[r].last_max_error_ratio_sq = s.lastRatio();
        st[r].successful_steps = (int64_t)d.succesfulSteps;
        st[r].failed_steps = (int64_t)d.failedSteps;
        st[r].pair_passes = s.pairPasses;
        st[r].factorizations = s.factorizations;
        st[r].singular_solves = s.singularSolves;
        st[r].unpivoted_factorizations = s.unpivotedFactorizations;
    }
    return D4_OK;
}

int d4_ens_set_config(d4_ens* e, const d4_ens_config* c)
{
    if (!e || !c || c->host_threads < 0) return D4_ERR_INVALID;
    D4_TRY(e->check(e->bp->set_config(*c)));
    const int want = c->host_threads > 0 ? std::min(c->host_threads, 64) : default_threads();
    if (want != e->pool->size()) {
        e->pool.reset(new ThreadPool(want));
        // the batched LU hands its host loops to the pool: give it the new one (the old one is gone)
        if (e->blu) e->blu->set_parallel(pool_for, e->pool.get());
    }
    e->cfg = *c;
    return D4_OK;
}

int d4_ens_get_config(const d4_ens* e, d4_ens_config* c)
{
    if (!e || !c) return D4_ERR_INVALID;
    *c = e->cfg;
    return D4_OK;
}

int d4_ens_autotune(d4_ens* e, d4_ens_config* best, double* best_ms)
{
    if (!e) return D4_ERR_INVALID;
    const int R = (int)e->sims.size();
    std::vector<int> off(R + 1, 0);
    for (int r = 0; r < R; ++r) off[r + 1] = off[r] + (int)e->data[r]->dc;
    BatchPasses& bp = *e->bp;
    D4_TRY(e->check(bp.set_layout(R, off.data())));
    // one speed pass and one Jacobian pass on the current positions, at the
    // stress and step size each realization is at (results are discarded)
    auto stage = [&](bool jac) {
        for (int r = 0; r < R; ++r) {
            const auto& d = *e->data[r];
            std::memcpy(bp.xy_h.as<double2>() + off[r], d.disl_sorted.data(), sizeof(double2) * d.dc);
            if (jac) {
                d4_jac_params p{};
                p.h = d.stepSize;
                p.tau_a = p.tau_b = d.externalStressProtocol->extStress(d.simTime);
                p.cutoff = d.cutOff;
                p.cut_sq = d.cutOffSqr;
                p.one_per_cut_sq = d.onePerCutOffSqr;
                p.heaviside = d.heavisideCutoff ? 1 : 0;
                p.wfunc = d.weightFunc;
                bp.params()[r] = real_params((int)d.dc, &p, 0.0);
                bp.params()[r].tol_cap2 = 0.0;
                bp.wf_h.as<int>()[r] = d.weightFunc;
            } else {
                bp.params()[r] = real_params((int)d.dc, nullptr, d.externalStressProtocol->extStress(d.simTime));
            }
        }
    };
    const d4_ens_config keep = e->cfg;
    d4_ens_config cand = keep;
    double best_t = 1e300;
    d4_ens_config best_c = keep;
    const int grids[] = {0, 2048, 4096, 8192, 16384};
    for (int order = 0; order < 3; ++order)
        for (int grid : grids)
            for (int graph = 0; graph < 2; ++graph) {
                cand.pass_grid = grid;
                cand.tile_order = order;
                cand.use_graph = graph;
                D4_TRY(e->check(bp.set_config(cand)));
                double t = 0;
                for (int rep = 0; rep < 3; ++rep) {  // first repetition warms up (and captures)
                    stage(false);
                    double t0 = now_ms();
                    D4_TRY(e->check(bp.speed(false)));
                    double ts = now_ms() - t0;
                    stage(true);
                    t0 = now_ms();
                    D4_TRY(e->check(bp.jacobian(false, false, nullptr)));
                    if (rep) t += ts + (now_ms() - t0);
                }
                t /= 2;
                if (t < best_t) {
                    best_t = t;
                    best_c = cand;
                }
            }
    best_c.host_threads = keep.host_threads;
    D4_TRY(d4_ens_set_config(e, &best_c));
    if (best) *best = best_c;
    if (best_ms) *best_ms = best_t;
    return D4_OK;
}

int d4_ens_timing(const d4_ens* e, double* ms4)
{
    if (!e || !ms4) return D4_ERR_INVALID;
    ms4[0] = e->t_pass;
    ms4[1] = e->t_solve;
    ms4[2] = e->t_host;
    ms4[3] = e->t_total;
    return D4_OK;
}

}  // extern "C"
