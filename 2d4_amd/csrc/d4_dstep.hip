// d4_dstep.hip -- the loop of Simulation::run (simulation.cpp:88-276) resident on the device for a batch
// of realizations (include/d4_dstep.h): every stage is a launch, and one iteration is captured once as a
// HIP graph and replayed.
//
// Per realization r (elements off[r] .. off[r+1]-1, n_r of them) the device holds its parameters
// (DsParams), its state (DsState: time, step size, strain, counters, the cutoff, whether interrupt()
// stopped it), four configurations of positions ({x, y} per dislocation: x0, the big step, the first
// and the second half step), the speeds and weights of the stages, the dense step matrix M = V + diag(s)
// (n_r x n_r column-major at moff[r], as the dense pair pass writes it) and the LU store the assembled
// matrix A is factored in.  An iteration enqueues, in the reference's order (simulation.cpp:98-154):
//   begin       interrupt() and the pass parameters of the stages (times, stresses, h, cutoff);
//   stage I     Jacobian pass at (x0, h, t0, t1); A = I + diag((1+d)/2) M; factor; g, solve, big = x0 - x;
//               speed pass at (big, t1); g, solve, big -= x;
//   stage II    d' = w(w^-1(d)/2), speeds moved to t1/2; A from M rescaled (simulation.cpp:546-608);
//               factor; g, solve, s1 = x0 - x; speed pass at (s1, t1/2); g, solve, s1 -= x;
//   stage III   Jacobian pass at (s1, h/2, t1/2, t1); A; factor; g, solve, s2 = s1 - x;
//               speed pass at (s2, t1); g, solve, s2 -= x;
//   finish      error ratio, accept/reject, strain, positions, step size, tolerance reset (stage IV).
// A realization that interrupt() stopped keeps its state; the launches still cover it (fixed shapes).
#include "../../include/d4_dstep.h"
#include "d4_ctx.h"
#include "d4_field.h"
#include "d4_host.h"
#include "d4_solver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace d4;

namespace sdddstCore {
int apply_sim_params(SimulationData& d, const d4_sim_params* p);  // d4_sim.cpp
}

namespace {

// per realization, fixed for the run
struct DsParams {
    double max_dt;        // -M; <= 0: none
    double minprec2;      // minPrecisitySq of PrecisionHandler (precision_handler.cpp:72-89)
    double dip2;          // dipolePrecisitySq (1 without the dipole criterion)
    double cap2;          // the passes' tolerance cap: minprec2 with the dipole criterion, else 0
    double cutoff_mult;
    double time_limit, strain_limit, speed_threshold;  // <= 0: none
    long long step_limit;                              // < 0: none
    int use_dip, heaviside, wfunc, stress_kind;
    double stress_init, stress_rate, stress_period, stress_amp;
};

struct DsState {
    double t, h, strain, last_ratio;
    double t0, t1, th;  // the iteration's times
    double cutoff, cut_sq, one_per;
    long long ok, failed, factorizations, unpivoted, singular_solves, pair_passes;
    int done;  // interrupt() held at the start of the iteration
    int sing;  // the last factorization met a zero pivot
};

// StressProtocol::extStress and its subclasses (stress_protocol.cpp:13-73)
__device__ double ext_stress(const DsParams& p, double t)
{
    if (p.stress_kind == 1) return t * p.stress_rate + p.stress_init;
    if (p.stress_kind == 2 || p.stress_kind == 3) {
        const double pt = t - floor((t / p.stress_period + 1. / 4)) * p.stress_period;
        double s = p.stress_rate * pt + p.stress_init;
        if (pt > p.stress_period / 4) s = p.stress_rate * p.stress_period / 2 - s + p.stress_init;
        return p.stress_kind == 3 ? s * p.stress_amp * t : s;
    }
    return 0;
}

struct DsArgs {
    int R, total;
    const int* off;
    const int* elem;  // realization of each element
    const DsParams* par;
    DsState* st;
    RealParams *rpJ1, *rpS1, *rpS2, *rpJ3, *rpS3;
    const int* slot0;
    double2 *x0, *big, *s1, *s2;
    double *vA, *vB, *vA2, *vB2, *sp, *sp2, *dvec, *dnew, *fac, *tol, *b;
    const double* M;
    const long long* moff;
    double* store;
    const long long* aoff;
    const int* npad;
    const int* flag;
    const int* sing;
    const int* lu_aux;  // the LU's orders and envelopes (ord then inv; last, top, first per block)
    const int* ordoff;  // per realization: its order in lu_aux, or -1 (natural)
    const int* envoff;  // per realization: its envelope in lu_aux, or -1 (whole matrix)
};

__device__ RealParams pass_params(const DsArgs& a, int r, double h, double tA, double tB)
{
    const DsParams& p = a.par[r];
    const DsState& s = a.st[r];
    const int n = a.off[r + 1] - a.off[r];
    RealParams q{};
    q.h = h;
    q.tauA = tA;
    q.tauB = tB;
    q.cut_sq = s.cut_sq;
    q.cutoff = s.cutoff;
    q.one_per = s.one_per;
    q.range_sq = (p.heaviside ? 1.0 : 49.0) * s.cut_sq;  // simulation.cpp:461-465
    q.tol_cap2 = p.cap2;
    q.heaviside = p.heaviside;
    q.wfunc = p.wfunc;
    q.slot0 = a.slot0[r];
    q.nb = (n + kTile - 1) / kTile;
    return q;
}

// interrupt() (simulation.cpp:722-750) and the parameters of the iteration's five pair passes
__global__ void ds_begin_kernel(DsArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    const DsParams& p = a.par[r];
    DsState& s = a.st[r];
    s.done = (p.time_limit > 0 && s.t > p.time_limit) || (p.strain_limit > 0 && s.strain > p.strain_limit) ||
             (p.step_limit >= 0 && s.ok > p.step_limit);
    s.t0 = s.t;
    s.t1 = s.t + s.h;
    s.th = s.t + s.h / 2;
    const double tau0 = ext_stress(p, s.t0), tau1 = ext_stress(p, s.t1), tauh = ext_stress(p, s.th);
    a.rpJ1[r] = pass_params(a, r, s.h, tau0, tau1);
    a.rpS1[r] = pass_params(a, r, 0.0, tau1, tau1);
    a.rpS2[r] = pass_params(a, r, 0.0, tauh, tauh);
    a.rpJ3[r] = pass_params(a, r, s.h / 2, tauh, tau1);
    a.rpS3[r] = pass_params(a, r, 0.0, tau1, tau1);
    if (!s.done) s.pair_passes += 5;
}

// The matrix the LU factors, from the dense pair pass's M (column c of realization r per block, identity
// beyond n_r), unfused as simulation.cpp:516-522 rounds it: A = M (1 + d_row)/2, + 1 on the diagonal;
// rescaled (stage II, simulation.cpp:546-598): A/2 f_row off the diagonal, 1 + (A - 1)/2 f_row on it.
// In the realization's LU order (store row/column i = dislocation ord[i]) and, on the first pass, only
// inside its envelope (blocks top .. last of each column's block: every entry the factorization, the
// check and the solves read).  only_flagged: just the realizations whose unpivoted factorization was
// flagged, whole (the pivoted factorization reads every entry).  M is read only for the pairs inside the
// storage radius at cfg, the configuration of the pass that wrote it (the same test, d^2 of the lower
// index minus the higher, as the pass: the pass leaves the rest unwritten); the others are 0 as the pass
// would have stored them.
// A group of columns of one 64-column block (blk: realization, block; the block split over `split`
// workgroups, group blockIdx.x % split) per workgroup: each of its four wavefronts takes every fourth
// column, lane l the rows r0 + l + 64 k of the block's row range -- the same rows in every column, so a
// row's order, position and factors are read once per k for the wavefront's columns.
__global__ __launch_bounds__(256) void ds_assemble_kernel(DsArgs a, const int2* __restrict__ blk, int split,
                                                          const double2* cfg, int rescaled, int only_flagged)
{
#pragma clang fp contract(off)
    const int2 rb = blk[blockIdx.x / split];
    const int r = rb.x, bc = rb.y, cw = kTile / split, c0 = bc * kTile + (blockIdx.x % split) * cw;
    if (only_flagged && !a.flag[r]) return;
    const int np = a.npad[r];
    const int base = a.off[r], n = a.off[r + 1] - base;
    const int* ord = a.ordoff[r] >= 0 ? a.lu_aux + a.ordoff[r] : nullptr;
    int r0 = 0, r1 = np;
    if (!only_flagged && a.envoff[r] >= 0) {
        const int nb = np / kTile;
        r0 = a.lu_aux[a.envoff[r] + nb + bc] * kTile;
        r1 = min(np, (a.lu_aux[a.envoff[r] + bc] + 1) * kTile);
    }
    double* const store = a.store + a.aoff[r];
    const double* const M = a.M + a.moff[r];
    const double range_sq = (a.par[r].heaviside ? 1.0 : 49.0) * a.st[r].cut_sq;  // the pass's storage test
    // the workgroup's columns: their original indices and positions, read once (the row loop below meets
    // each column once per 64 rows)
    __shared__ int col_o[kTile];
    __shared__ double2 col_p[kTile];
    for (int cc = threadIdx.x; cc < cw; cc += blockDim.x) {
        const int c = c0 + cc;
        const int oc = c < n ? (ord ? ord[c] : c) : -1;
        col_o[cc] = oc;
        col_p[cc] = oc >= 0 ? cfg[base + oc] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int row = r0 + lane; row < r1; row += 64) {  // r0, r1 multiples of 64: every lane takes a row
        const bool real = row < n;
        const int orow = real ? (ord ? ord[row] : row) : 0;
        const double2 pr = cfg[base + orow];
        const double fr = (1 + a.dvec[base + orow]) / 2;
        const double f = rescaled ? a.fac[base + orow] : 0.0;
        for (int cc = w; cc < cw; cc += 4) {
            const int c = c0 + cc;
            double v;
            if (!real || c >= n) {
                v = row == c ? 1.0 : 0.0;
            } else {
                const int oc = col_o[cc];
                double m = 0.0;
                if (orow == oc) {
                    m = M[(size_t)oc * n + orow];  // s, from the pass's reduction
                } else {
                    const double2 pc = col_p[cc];
                    const bool lower = orow < oc;
                    const double dx = lower ? wrap(pr.x - pc.x) : wrap(pc.x - pr.x);
                    const double dy = lower ? wrap(pr.y - pc.y) : wrap(pc.y - pr.y);
                    if (dist2(dx, dy) < range_sq) m = M[(size_t)oc * n + orow];
                }
                v = m * fr;
                if (row == c) v = v + 1;
                if (rescaled) v = row == c ? 1 + (v - 1) / 2 * f : v / 2 * f;
            }
            store[(size_t)c * np + row] = v;
        }
    }
}

// counters of a factorization; the zero-pivot flag for the solves that follow
__global__ void ds_factored_kernel(DsArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    DsState& s = a.st[r];
    s.sing = a.flag[r] ? a.sing[r] : 0;
    if (s.done) return;
    s.factorizations++;
    if (!a.flag[r]) s.unpivoted++;
}

enum GMode : int { kGNone = 0, kGStep1, kGStep2, kGHalf1, kGHalf2, kGStep3a, kGStep3b };

// g of simulation.cpp:100-105, 121-124, 142-145 (no positions) and calcG (:620-624, positions new - old),
// unfused in the reference's order, into the packed right-hand sides
__global__ void ds_g_kernel(DsArgs a, int mode)
{
#pragma clang fp contract(off)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.total) return;
    const int r = a.elem[g];
    const double h = a.st[r].h;
    double hh = h / 2, d, ve, vi, dx = 0.0;
    switch (mode) {
    case kGStep1: hh = h; d = a.dvec[g]; ve = a.vB[g]; vi = a.vA[g]; break;
    case kGStep2: hh = h; d = a.dvec[g]; ve = a.sp2[g]; vi = a.vA[g]; dx = a.big[g].x - a.x0[g].x; break;
    case kGHalf1: d = a.dnew[g]; ve = a.sp[g]; vi = a.vA[g]; break;
    case kGHalf2: d = a.dnew[g]; ve = a.sp2[g]; vi = a.vA[g]; dx = a.s1[g].x - a.x0[g].x; break;
    case kGStep3a: d = a.dvec[g]; ve = a.vB2[g]; vi = a.vA2[g]; break;
    default: d = a.dvec[g]; ve = a.sp2[g]; vi = a.vA2[g]; dx = a.s2[g].x - a.s1[g].x; break;
    }
    const double t = hh * ((1 + d) * ve + (1 - d) * vi) / 2;
    a.b[g] = (mode == kGStep1 || mode == kGHalf1 || mode == kGStep3a) ? -t : dx - t;
}

// dst.x = src.x - x with the solution x (simulation.cpp:106-109 and its like), x zeroed outside (-1, 1)
// after a factorization with a zero pivot (simulation.cpp:651-657); y copied
__global__ void ds_update_kernel(DsArgs a, const double2* src, double2* dst)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.total) return;
    double x = a.b[g];
    if (a.st[a.elem[g]].sing && !((-1 < x) && (x < 1))) x = 0;
    const double2 s = src[g];
    dst[g] = make_double2(s.x - x, s.y);
}

// one singular solve counted per realization (the reference counts the solves after a zero pivot)
__global__ void ds_solved_kernel(DsArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.R) return;
    DsState& s = a.st[r];
    if (!s.done && s.sing) s.singular_solves++;
}

// stage II's weights and speeds (simulation.cpp:546-608): d' = w(w^-1(d) factor), f = (1 + d')/(1 + d),
// the speeds at t1 moved to t1/2 by the stress difference
__global__ void ds_rescale_kernel(DsArgs a)
{
#pragma clang fp contract(off)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.total) return;
    const int r = a.elem[g];
    const DsParams& p = a.par[r];
    const DsState& s = a.st[r];
    const double factor = (s.th - s.t0) / (s.t1 - s.t0);
    const double ow = a.dvec[g];
    const double nw = ow == 0 ? 0. : weight(weight_inv(ow, (char)p.wfunc) * factor, (char)p.wfunc);
    a.dnew[g] = nw;
    a.fac[g] = (1 + nw) / (1 + ow);
    const int n = a.off[r + 1] - a.off[r];
    const double bsg = g - a.off[r] < n / 2 ? 1.0 : -1.0;
    a.sp[g] = a.vB[g] + (ext_stress(p, s.th) - ext_stress(p, s.t1)) * bsg;
}

// dvec := d' for the stage-II solves' g (the reference replaces dVec, simulation.cpp:546-608)
__global__ void ds_take_new_weights_kernel(DsArgs a)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.total) a.dvec[g] = a.dnew[g];
}

constexpr int kFinishThreads = 256;

// stage IV (simulation.cpp:160-276, precision_handler.cpp:41-62): one workgroup per realization
__global__ __launch_bounds__(kFinishThreads) void ds_finish_kernel(DsArgs a)
{
    __shared__ double red[3][kFinishThreads];
    const int r = blockIdx.x, tid = threadIdx.x;
    const DsParams& p = a.par[r];
    DsState& s = a.st[r];
    if (s.done) return;
    const int base = a.off[r], n = a.off[r + 1] - base;
    double mr = 0.0, sa = 0.0, sb = 0.0, av = 0.0;
    for (int i = tid; i < n; i += kFinishThreads) {
        const int g = base + i;
        const double e = fabs(a.big[g].x - a.s2[g].x);
        const double q = p.use_dip ? e * e / a.tol[g] : e * e / p.minprec2;
        mr = fmax(mr, q);
        const double bsg = i < n / 2 ? 1.0 : -1.0;
        sa += bsg * (a.s1[g].x - a.x0[g].x);
        sb += bsg * (a.s2[g].x - a.s1[g].x);
        av += fabs(a.vA2[g]);
    }
    red[0][tid] = mr;
    red[1][tid] = sa;
    red[2][tid] = sb;
    __syncthreads();
    for (int w = kFinishThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + w]);
            red[1][tid] += red[1][tid + w];
            red[2][tid] += red[2][tid + w];
        }
        __syncthreads();
    }
    const double ratio = red[0][0] / p.dip2;
    const bool accept = ratio < 1;
    const double inc1 = red[1][0], inc2 = red[2][0];
    __syncthreads();
    red[0][tid] = av;
    __syncthreads();
    for (int w = kFinishThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[0][tid] += red[0][tid + w];
        __syncthreads();
    }
    const double avg = red[0][0] / n;
    for (int i = tid; i < n; i += kFinishThreads) {
        const int g = base + i;
        if (accept) a.x0[g] = make_double2(wrap(a.s2[g].x), a.x0[g].y);
        if (p.use_dip) a.tol[g] = p.minprec2;  // pH->reset()
    }
    if (tid == 0) {
        if (accept) {
            s.t += s.h;
            s.ok++;
            s.strain += inc1;
            s.strain += inc2;
            if (p.speed_threshold > 0 && p.speed_threshold > avg) {  // simulation.cpp:215-219
                const double m = 1e20;  // SimulationData::updateCutOff with the multiplier 1e20
                s.cutoff = m / sqrt((double)n);
                s.cut_sq = s.cutoff * s.cutoff;
                s.one_per = 1 / s.cut_sq;
            }
        } else {
            s.failed++;
        }
        s.last_ratio = ratio;
        s.h = 0.9 * s.h * fmin(2., pow(ratio, -1. / 6));  // precision_handler.cpp:57-62
        if (p.max_dt > 0 && p.max_dt < s.h) s.h = p.max_dt;
    }
}

}  // namespace

struct d4_dstep {
    d4_ctx* ctx = nullptr;
    int R = 0, total = 0, nslots = 0, ntiles = 0, nblocks = 0, npad_max = 0;
    std::vector<int> off;
    std::vector<std::vector<unsigned>> order;  // per realization: disl_order (input index -> sorted index)
    std::unique_ptr<d4host::DenseLUGraph> lu;
    std::vector<void*> bufs;
    DsArgs a{};
    TileDesc* tiles = nullptr;
    int2* blocks = nullptr;
    double* part = nullptr;
    int* sing_d = nullptr;
    int* flag_d = nullptr;
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    // its own stream: a graph is captured from a stream other than the legacy default one, whatever
    // stream the context works on
    hipStream_t stream = nullptr;
    std::string err;

    ~d4_dstep()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (void* p : bufs) (void)hipFree(p);
    }
    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }
    template <class T>
    int alloc(T*& p, size_t count, const char* what)
    {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess)
            return fail(D4_ERR_HIP, std::string("device allocation failed: ") + what);
        bufs.push_back(q);
        p = static_cast<T*>(q);
        return D4_OK;
    }

    PassArgs pass_args(const double2* xy, const RealParams* rp, int with_m) const
    {
        PassArgs p{};
        p.xy = xy;
        p.off = a.off;
        p.tiles = tiles;
        p.ntiles = ntiles;
        p.rp = rp;
        p.part = part;
        p.tol = a.tol;
        if (with_m) {
            p.M = const_cast<double*>(a.M);
            p.moff = reinterpret_cast<const int64_t*>(a.moff);
            p.skip_empty = 1;  // ds_assemble reads no entry outside the storage radius
        }
        return p;
    }
    ReduceArgs reduce_args(const RealParams* rp, int nq) const
    {
        ReduceArgs q{};
        q.part = part;
        q.nq = nq;
        q.off = a.off;
        q.rp = rp;
        q.R = R;
        q.total = total;
        q.blocks = blocks;
        q.nblocks = nblocks;
        q.add_tau = 1;
        return q;
    }

    // a Jacobian pass with dense storage (the reference's cutoff: V = 0 outside the storage radius)
    int jacobian(const double2* xy, const RealParams* rp, double* vA, double* vB, hipStream_t st)
    {
        hipError_t e = launch_pair_pass(kModeJacDenseCut, pass_args(xy, rp, 1), st);
        ReduceArgs q = reduce_args(rp, 2);
        q.vA = vA;
        q.vB = vB;
        q.dvec = a.dvec;
        q.M = const_cast<double*>(a.M);
        q.moff = reinterpret_cast<const int64_t*>(a.moff);
        if (e == hipSuccess) e = launch_reduce(q, st);
        return e == hipSuccess ? D4_OK : fail(D4_ERR_HIP, std::string("device step: Jacobian pass: ") + hipGetErrorString(e));
    }
    int speed(const double2* xy, const RealParams* rp, double* v, hipStream_t st)
    {
        hipError_t e = launch_pair_pass(kModeSpeed, pass_args(xy, rp, 0), st);
        ReduceArgs q = reduce_args(rp, 1);
        q.vA = v;
        if (e == hipSuccess) e = launch_reduce(q, st);
        return e == hipSuccess ? D4_OK : fail(D4_ERR_HIP, std::string("device step: speed pass: ") + hipGetErrorString(e));
    }
    // cfg: the configuration of the Jacobian pass whose matrix is assembled
    int factor(const double2* cfg, int rescaled, hipStream_t st)
    {
        // a workgroup per 64-column block of a realization, or per part of one when there are few blocks
        int split = 1;
        while (split < 16 && (long long)nblocks * split < 2048) split *= 2;
        const dim3 g(nblocks * split);
        hipLaunchKernelGGL(ds_assemble_kernel, g, dim3(256), 0, st, a, blocks, split, cfg, rescaled, 0);
        if (lu->factor_unpivoted(a.store, flag_d, st)) return fail(D4_ERR_HIP, "device step: factorization");
        hipLaunchKernelGGL(ds_assemble_kernel, g, dim3(256), 0, st, a, blocks, split, cfg, rescaled, 1);
        if (lu->factor_pivoted(a.store, flag_d, sing_d, st)) return fail(D4_ERR_HIP, "device step: factorization");
        hipLaunchKernelGGL(ds_factored_kernel, dim3((R + 255) / 256), dim3(256), 0, st, a);
        return hipGetLastError() == hipSuccess ? D4_OK : fail(D4_ERR_HIP, "device step: factorization");
    }
    // g (mode), the solve, dst = src - x
    int solve(int mode, const double2* src, double2* dst, hipStream_t st)
    {
        const dim3 ge((total + 255) / 256);
        hipLaunchKernelGGL(ds_g_kernel, ge, dim3(256), 0, st, a, mode);
        if (lu->solve(a.store, a.b, st)) return fail(D4_ERR_HIP, "device step: solve");
        hipLaunchKernelGGL(ds_solved_kernel, dim3((R + 255) / 256), dim3(256), 0, st, a);
        hipLaunchKernelGGL(ds_update_kernel, ge, dim3(256), 0, st, a, src, dst);
        return hipGetLastError() == hipSuccess ? D4_OK : fail(D4_ERR_HIP, "device step: update");
    }

#define DS_TRY(x)                \
    do {                         \
        const int rc_ = (x);     \
        if (rc_) return rc_;     \
    } while (0)

    // one loop iteration of every realization (simulation.cpp:88-276)
    int enqueue_iteration(hipStream_t st)
    {
        const dim3 gr((R + 255) / 256), ge((total + 255) / 256);
        hipLaunchKernelGGL(ds_begin_kernel, gr, dim3(256), 0, st, a);
        // stage I
        DS_TRY(jacobian(a.x0, a.rpJ1, a.vA, a.vB, st));
        DS_TRY(factor(a.x0, 0, st));
        DS_TRY(solve(kGStep1, a.x0, a.big, st));
        DS_TRY(speed(a.big, a.rpS1, a.sp2, st));
        DS_TRY(solve(kGStep2, a.big, a.big, st));
        // stage II
        hipLaunchKernelGGL(ds_rescale_kernel, ge, dim3(256), 0, st, a);
        DS_TRY(factor(a.x0, 1, st));
        hipLaunchKernelGGL(ds_take_new_weights_kernel, ge, dim3(256), 0, st, a);
        DS_TRY(solve(kGHalf1, a.x0, a.s1, st));
        DS_TRY(speed(a.s1, a.rpS2, a.sp2, st));
        DS_TRY(solve(kGHalf2, a.s1, a.s1, st));
        // stage III
        DS_TRY(jacobian(a.s1, a.rpJ3, a.vA2, a.vB2, st));
        DS_TRY(factor(a.s1, 0, st));
        DS_TRY(solve(kGStep3a, a.s1, a.s2, st));
        DS_TRY(speed(a.s2, a.rpS3, a.sp2, st));
        DS_TRY(solve(kGStep3b, a.s2, a.s2, st));
        // stage IV
        hipLaunchKernelGGL(ds_finish_kernel, dim3(R), dim3(kFinishThreads), 0, st, a);
        return hipGetLastError() == hipSuccess ? D4_OK : fail(D4_ERR_HIP, "device step: launch");
    }

    int iteration(hipStream_t st)
    {
        if (!use_graph) return enqueue_iteration(st);
        if (!graph) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return fail(D4_ERR_HIP, "device step: graph capture begin");
            const int rc = enqueue_iteration(st);
            e = hipStreamEndCapture(st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(D4_ERR_HIP, "device step: graph capture end");
            e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                graph = nullptr;
                return fail(D4_ERR_HIP, "device step: graph instantiate");
            }
        }
        return hipGraphLaunch(graph, st) == hipSuccess ? D4_OK : fail(D4_ERR_HIP, "device step: graph launch");
    }

    int states(std::vector<DsState>& out) const
    {
        out.resize(R);
        if (hipMemcpyAsync(out.data(), a.st, sizeof(DsState) * R, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return D4_ERR_HIP;
        return D4_OK;
    }
};

extern "C" {

int d4_dstep_create(d4_ctx* ctx, int R, const int* off, const double* x, const double* y, const double* b,
                    const d4_sim_params* p, int params_per_realization, d4_dstep** out)
{
    if (!ctx || R <= 0 || !off || !x || !y || !b || !p || !out) return D4_ERR_INVALID;
    *out = nullptr;
    if (off[0] != 0) return ctx->fail(D4_ERR_INVALID, "offsets must start at 0");
    for (int r = 0; r < R; ++r)
        if (off[r + 1] <= off[r]) return ctx->fail(D4_ERR_INVALID, "every realization needs dislocations");
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(D4_ERR_HIP, "device step: set device");
    std::unique_ptr<d4_dstep> s(new d4_dstep());
    s->ctx = ctx;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        s->stream = nullptr;
        return ctx->fail(D4_ERR_HIP, "device step: stream creation failed");
    }
    s->R = R;
    s->off.assign(off, off + R + 1);
    const int T = s->total = off[R];
    // SimulationData of each realization: sorted positions, numbering, cutoff, parameters
    std::vector<DsParams> par(R);
    std::vector<DsState> st(R);
    std::vector<double2> xy(T);
    std::vector<int> elem(T), npad(R), slot0(R);
    std::vector<long long> moff(R), b_off(R), aoff(R);
    std::vector<double> tol(T);
    s->order.resize(R);
    long long mtot = 0;
    int nslots = 0;
    for (int r = 0; r < R; ++r) {
        const int n = off[r + 1] - off[r];
        sdddstCore::SimulationData d(n, x + off[r], y + off[r], b + off[r]);
        if (!d.error.empty()) return ctx->fail(D4_ERR_INVALID, "realization " + std::to_string(r) + ": " + d.error);
        const d4_sim_params& q = params_per_realization ? p[r] : *p;
        if (sdddstCore::apply_sim_params(d, &q))
            return ctx->fail(D4_ERR_INVALID, "realization " + std::to_string(r) + ": invalid parameters");
        for (int i = 0; i < n; ++i) {
            xy[off[r] + i] = make_double2(d.disl_sorted[i].x, d.disl_sorted[i].y);
            elem[off[r] + i] = r;
        }
        s->order[r] = d.disl_order;
        DsParams& pp = par[r];
        pp.max_dt = d.isMaxStepSizeLimit ? d.maxStepSizeLimit : 0.0;
        pp.use_dip = q.dipole_prec != 0;  // precision_handler.cpp:72-89
        pp.dip2 = pp.use_dip ? q.dipole_prec * q.dipole_prec : 1.0;
        pp.minprec2 = q.prec * q.prec / (pp.use_dip ? pp.dip2 : 1.0);
        pp.cap2 = pp.use_dip ? pp.minprec2 : 0.0;
        pp.cutoff_mult = q.cutoff_mult;
        pp.time_limit = q.time_limit;
        pp.strain_limit = q.strain_limit;
        pp.speed_threshold = q.speed_threshold;
        pp.step_limit = q.step_limit;
        pp.heaviside = q.heaviside != 0;
        pp.wfunc = q.wfunc;
        pp.stress_kind = q.stress_kind;
        pp.stress_init = q.stress_init;
        pp.stress_rate = q.stress_rate;
        pp.stress_period = q.stress_period;
        pp.stress_amp = q.stress_amp;
        DsState& ss = st[r];
        std::memset(&ss, 0, sizeof(ss));
        ss.h = d.stepSize;
        ss.cutoff = d.cutOff;
        ss.cut_sq = d.cutOffSqr;
        ss.one_per = d.onePerCutOffSqr;
        for (int i = 0; i < n; ++i) tol[off[r] + i] = pp.minprec2;
        moff[r] = mtot;
        mtot += (long long)n * n;
        b_off[r] = off[r];
        npad[r] = d4host::gpu_lu_padded(n);
        s->npad_max = std::max(s->npad_max, npad[r]);
        slot0[r] = nslots;
        const int nb = (n + kTile - 1) / kTile;
        nslots += nb * (nb + 1) / 2;
    }
    s->nslots = nslots;
    std::vector<int> nr(R);
    for (int r = 0; r < R; ++r) nr[r] = off[r + 1] - off[r];
    // Each realization's LU order and envelope, fixed for the run.  Rows sorted by |y| (the distance from
    // the line y = 0; y never changes): a pair that can hold a matrix entry is closer than the storage
    // radius, so closer than it in periodic y, so its keys differ by less than the radius -- for any x
    // (key_band_envelope).  Kept where the radius cannot grow during the run (no cutoff change,
    // simulation.cpp:215-219) and the envelope needs at most half the update tiles of the whole matrix.
    std::vector<std::vector<int>> ords(R);
    std::vector<d4host::BlockEnvelope> envs(R);
    for (int r = 0; r < R; ++r) {
        const int n = nr[r], np = npad[r];
        if (par[r].speed_threshold > 0) continue;
        const double reach = std::sqrt((par[r].heaviside ? 1.0 : 49.0) * st[r].cut_sq);
        if (!(reach < 0.5)) continue;  // the whole cell in y
        std::vector<double> key(n);
        for (int i = 0; i < n; ++i) key[i] = std::fabs(xy[off[r] + i].y);
        std::vector<int> o(n);
        for (int i = 0; i < n; ++i) o[i] = i;
        std::stable_sort(o.begin(), o.end(), [&](int u, int v) { return key[u] < key[v]; });
        std::vector<double> ks(n);
        for (int i = 0; i < n; ++i) ks[i] = key[o[i]];
        // the margin covers the rounding of d^2 = dx^2 + dy^2 in the pass's storage test
        d4host::BlockEnvelope e = d4host::key_band_envelope(n, np, kTile, ks.data(), reach * (1 + 1e-12) + 1e-300);
        const long long nb = np / kTile;
        long long dense = 0;
        for (long long q = 0; q < nb; ++q) dense += (nb - 1 - q) * (nb - 1 - q);
        if (2 * e.tiles > dense) continue;
        ords[r] = std::move(o);
        envs[r] = std::move(e);
    }
    s->lu.reset(d4host::new_dense_lu_graph(ctx->device, R, nr.data(), b_off.data(), &ords, &envs));
    if (!s->lu) return ctx->fail(D4_ERR_INVALID, "device step: the LU batch does not fit this device");
    for (int r = 0; r < R; ++r) aoff[r] = s->lu->a_off(r);
    // the tile table (realization-major, each realization's upper triangle row by row) and reduce blocks
    std::vector<TileDesc> tiles;
    std::vector<int2> blocks;
    for (int r = 0; r < R; ++r) {
        const int nb = (nr[r] + kTile - 1) / kTile;
        int slot = slot0[r];
        for (int I = 0; I < nb; ++I) {
            blocks.push_back(make_int2(r, I));
            for (int J = I; J < nb; ++J) tiles.push_back(TileDesc{r, I * kTile, J * kTile, slot++});
        }
    }
    s->ntiles = (int)tiles.size();
    s->nblocks = (int)blocks.size();
    d4_dstep& e = *s;
    DsArgs& a = e.a;
    a.R = R;
    a.total = T;
    int* off_d = nullptr;
    int* elem_d = nullptr;
    DsParams* par_d = nullptr;
    int* slot0_d = nullptr;
    long long* moff_d = nullptr;
    long long* aoff_d = nullptr;
    int* npad_d = nullptr;
    double* M_d = nullptr;
    int rc = e.alloc(off_d, R + 1, "offsets");
    if (!rc) rc = e.alloc(elem_d, T, "element table");
    if (!rc) rc = e.alloc(par_d, R, "parameters");
    if (!rc) rc = e.alloc(a.st, R, "state");
    for (RealParams** q : {&a.rpJ1, &a.rpS1, &a.rpS2, &a.rpJ3, &a.rpS3})
        if (!rc) rc = e.alloc(*q, R, "pass parameters");
    if (!rc) rc = e.alloc(slot0_d, R, "slots");
    for (double2** q : {&a.x0, &a.big, &a.s1, &a.s2})
        if (!rc) rc = e.alloc(*q, T, "positions");
    for (double** q : {&a.vA, &a.vB, &a.vA2, &a.vB2, &a.sp, &a.sp2, &a.dvec, &a.dnew, &a.fac, &a.tol, &a.b})
        if (!rc) rc = e.alloc(*q, T, "vectors");
    if (!rc) rc = e.alloc(M_d, (size_t)mtot, "step matrices");
    if (!rc) rc = e.alloc(moff_d, R, "matrix offsets");
    if (!rc) rc = e.alloc(a.store, (size_t)e.lu->store_size(), "LU store");
    if (!rc) rc = e.alloc(aoff_d, R, "LU offsets");
    if (!rc) rc = e.alloc(npad_d, R, "LU orders");
    int* ordoff_d = nullptr;
    int* envoff_d = nullptr;
    if (!rc) rc = e.alloc(ordoff_d, R, "LU order offsets");
    if (!rc) rc = e.alloc(envoff_d, R, "LU envelope offsets");
    if (!rc) rc = e.alloc(e.flag_d, R, "LU flags");
    if (!rc) rc = e.alloc(e.sing_d, R, "LU zero pivots");
    if (!rc) rc = e.alloc(e.tiles, tiles.size(), "tile table");
    if (!rc) rc = e.alloc(e.blocks, blocks.size(), "reduce blocks");
    if (!rc) rc = e.alloc(e.part, 2 * 128 * (size_t)nslots, "partial sums");
    if (rc) return ctx->fail(rc, e.err);
    a.off = off_d;
    a.elem = elem_d;
    a.par = par_d;
    a.slot0 = slot0_d;
    a.M = M_d;
    a.moff = moff_d;
    a.aoff = aoff_d;
    a.npad = npad_d;
    a.flag = e.flag_d;
    a.sing = e.sing_d;
    a.lu_aux = e.lu->aux();
    a.ordoff = ordoff_d;
    a.envoff = envoff_d;
    std::vector<int> ordoff(R), envoff(R);
    for (int r = 0; r < R; ++r) {
        ordoff[r] = e.lu->ord_off(r);
        envoff[r] = e.lu->env_off(r);
    }
    bool ok = true;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        ok = ok && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    up(off_d, off, sizeof(int) * (R + 1));
    up(elem_d, elem.data(), sizeof(int) * T);
    up(par_d, par.data(), sizeof(DsParams) * R);
    up(a.st, st.data(), sizeof(DsState) * R);
    up(slot0_d, slot0.data(), sizeof(int) * R);
    for (double2* q : {a.x0, a.big, a.s1, a.s2}) up(q, xy.data(), sizeof(double2) * T);
    up(a.tol, tol.data(), sizeof(double) * T);
    up(moff_d, moff.data(), sizeof(long long) * R);
    up(ordoff_d, ordoff.data(), sizeof(int) * R);
    up(envoff_d, envoff.data(), sizeof(int) * R);
    up(aoff_d, aoff.data(), sizeof(long long) * R);
    up(npad_d, npad.data(), sizeof(int) * R);
    up(e.tiles, tiles.data(), sizeof(TileDesc) * tiles.size());
    up(e.blocks, blocks.data(), sizeof(int2) * blocks.size());
    if (!ok) return ctx->fail(D4_ERR_HIP, "device step: upload failed");
    *out = s.release();
    return D4_OK;
}

void d4_dstep_destroy(d4_dstep* s) { delete s; }

const char* d4_dstep_last_error(const d4_dstep* s) { return s ? s->err.c_str() : "null device step"; }

int d4_dstep_set_graph(d4_dstep* s, int use_graph)
{
    if (!s) return D4_ERR_INVALID;
    s->use_graph = use_graph != 0;
    return D4_OK;
}

int d4_dstep_run(d4_dstep* s, int iterations, int* nactive)
{
    if (!s || iterations < 0) return D4_ERR_INVALID;
    hipStream_t st = s->stream;
    std::vector<DsState> ss;
    std::vector<DsParams> par(s->R);
    if (hipMemcpy(par.data(), s->a.par, sizeof(DsParams) * s->R, hipMemcpyDeviceToHost) != hipSuccess)
        return s->fail(D4_ERR_HIP, "device step: parameter download");
    // interrupt() of a realization from its state (simulation.cpp:722-750)
    auto active = [&]() {
        int k = 0;
        for (int r = 0; r < s->R; ++r) {
            const DsParams& p = par[r];
            const DsState& q = ss[r];
            k += !((p.time_limit > 0 && q.t > p.time_limit) || (p.strain_limit > 0 && q.strain > p.strain_limit) ||
                   (p.step_limit >= 0 && q.ok > p.step_limit));
        }
        return k;
    };
    if (s->states(ss)) return s->fail(D4_ERR_HIP, "device step: state download");
    int left = active();
    for (int done = 0; done < iterations && left > 0;) {
        const int chunk = std::min(16, iterations - done);  // then look whether anything still runs
        for (int k = 0; k < chunk; ++k)
            if (const int rc = s->iteration(st)) return rc;
        done += chunk;
        if (s->states(ss)) return s->fail(D4_ERR_HIP, "device step: iteration failed on the device");
        left = active();
    }
    if (nactive) *nactive = left;
    return D4_OK;
}

int d4_dstep_positions(const d4_dstep* s, double* xy, int original_order)
{
    if (!s || !xy) return D4_ERR_INVALID;
    std::vector<double2> h(s->total);
    if (hipMemcpy(h.data(), s->a.x0, sizeof(double2) * s->total, hipMemcpyDeviceToHost) != hipSuccess)
        return D4_ERR_HIP;
    for (int r = 0; r < s->R; ++r) {
        const int n = s->off[r + 1] - s->off[r];
        for (int i = 0; i < n; ++i) {
            const unsigned k = original_order ? s->order[r][i] : (unsigned)i;
            xy[2 * (size_t)(s->off[r] + i)] = h[s->off[r] + k].x;
            xy[2 * (size_t)(s->off[r] + i) + 1] = h[s->off[r] + k].y;
        }
    }
    return D4_OK;
}

int d4_dstep_stats(const d4_dstep* s, d4_sim_stats* out)
{
    if (!s || !out) return D4_ERR_INVALID;
    std::vector<DsState> ss;
    if (s->states(ss)) return D4_ERR_HIP;
    for (int r = 0; r < s->R; ++r) {
        d4_sim_stats& o = out[r];
        std::memset(&o, 0, sizeof(o));
        o.sim_time = ss[r].t;
        o.step_size = ss[r].h;
        o.strain = ss[r].strain;
        o.last_max_error_ratio_sq = ss[r].last_ratio;
        o.successful_steps = ss[r].ok;
        o.failed_steps = ss[r].failed;
        o.pair_passes = ss[r].pair_passes;
        o.factorizations = ss[r].factorizations;
        o.singular_solves = ss[r].singular_solves;
        o.unpivoted_factorizations = ss[r].unpivoted;
    }
    return D4_OK;
}

}  // extern "C"
