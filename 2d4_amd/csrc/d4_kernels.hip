// d4_kernels.hip -- gfx950 kernels of the periodic pair-interaction path.
//
// Pair passes (simulation.cpp:287-333 speed pass, simulation.cpp:345-527
// Jacobian pass) are evaluated per unordered pair (i < j, sorted index order,
// dx = x_i - x_j as simulation.cpp:446-450 orients them) in 64 x 64 tiles,
// one wavefront per tile:
//   * lane l owns target i = I0 + l; at step s it meets source j = J0 + (l ^ s),
//     so within a step every lane touches a different source (no LDS conflicts);
//   * target-side sums stay in registers, source-side sums go to LDS with
//     ds_add_f64, and each tile writes both as partials to its own slot;
//     launch_reduce adds the partials in a fixed order (deterministic, no
//     floating-point atomics in global memory);
//   * in the dense Jacobian mode each 16-step phase is staged in LDS and written
//     twice, V_ij to (row j, column i) and its mirror to (row i, column j), both
//     as 128-byte contiguous column segments.
#include "d4_field.h"
#include "d4_kernels.h"

#include <algorithm>
#include <climits>

namespace d4 {

// minimum of a non-negative double via its bit pattern (ordering of
// non-negative doubles equals that of their unsigned 64-bit patterns)
__device__ __forceinline__ void atomic_min_nonneg(double* p, double v)
{
    atomicMin(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// The wave's buffered COO entries (cbuf[0, fill)) to the global list: one
// atomic reserves their places, the lanes copy them; entries past the
// capacity are counted but not stored (the host grows the list and repeats).
__device__ __forceinline__ void coo_flush(const PassArgs& a, CooEntry* cbuf, int fill, int lane)
{
    __syncthreads();  // the workgroup is this one wave: the buffer's LDS writes are complete
    int b0 = 0;
    if (lane == 0) b0 = atomicAdd(a.coo_count, fill);
    b0 = __shfl(b0, 0);
    for (int k = lane; k < fill; k += 64) {
        const int pos = b0 + k;
        if (pos < a.coo_cap) a.coo[pos] = cbuf[k];
    }
    __syncthreads();  // every lane has read the buffer before it is refilled
}

// v with its sign flipped when neg: the product by b_i b_j = -1 (exact)
__device__ __forceinline__ double flip_if(double v, bool neg)
{
    return __longlong_as_double(__double_as_longlong(v) ^ ((long long)neg << 63));
}

// the value of the neighbouring lane (lanes 2k and 2k + 1 swap), by DPP quad_perm [1, 0, 3, 2]; every
// lane of the wavefront must be active
__device__ __forceinline__ unsigned long long swap_adjacent_u64(unsigned long long v)
{
    constexpr int kQuadSwap = 1 | (0 << 2) | (3 << 4) | (2 << 6);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)v, kQuadSwap, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), kQuadSwap, 0xF, 0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

__device__ __forceinline__ double swap_adjacent(double v)
{
    return __longlong_as_double((long long)swap_adjacent_u64((unsigned long long)__double_as_longlong(v)));
}

__device__ __forceinline__ double2 swap_adjacent(double2 v)
{
    return make_double2(swap_adjacent(v.x), swap_adjacent(v.y));
}

// One 64 x 64 tile of a pass, evaluated by one wavefront (the LDS arrays are
// per-workgroup allocations of this function's instantiation).
template <int MODE, bool EXACT>
__device__ __forceinline__ void pair_tile(const PassArgs& a, const TileDesc td)
{
    constexpr bool kTab = MODE == kModeSpeedTab || MODE == kModeJacDenseTab;
    constexpr bool kDense = MODE == kModeJacDense || MODE == kModeJacDenseCut || MODE == kModeJacDenseTab;
    constexpr bool kCoo = MODE == kModeJacCoo;
    constexpr bool kVjp = MODE == kModeVjp;
    constexpr int NQ = (MODE == kModeSpeed || MODE == kModeJacCoo || MODE == kModeSpeedTab) ? 1 : 2;

    __shared__ double sx[kTile], sy[kTile], sw[kVjp ? kTile : 1];
    __shared__ double jq0[kTile], jq1[NQ == 2 ? kTile : 1];
    __shared__ double stage[kDense ? 4 * 16 * 17 : 1];
    // COO mode: entries gathered in LDS and appended to the global list with
    // one atomic per flush (a full buffer or the tile's end), not per step
    constexpr int kCooBuf = 256;
    __shared__ CooEntry cbuf[kCoo ? kCooBuf : 1];
    int cfill = 0;  // entries in cbuf (the same in every lane)

    const int lane = threadIdx.x;
    const int base = a.off ? a.off[td.r] : 0;
    const int n = a.off ? a.off[td.r + 1] - base : a.n0;
    const RealParams rp = a.rp ? a.rp[td.r] : a.rp0;
    const bool diag = td.I0 == td.J0;
    const int half = n >> 1;
    double* const Mb = (kDense && a.moff) ? a.M + a.moff[td.r] : a.M;
    const int64_t ld = (kDense && a.moff) ? (int64_t)n : a.ld;
    const int i = td.I0 + lane;
    const bool vi = i < n;
    const double2 pi = vi ? a.xy[base + i] : make_double2(0.0, 0.0);
    const double bi = i < half ? 1.0 : -1.0;
    const bool ipos = i < half;
    {
        const int jj = td.J0 + lane;
        const double2 pj = jj < n ? a.xy[base + jj] : make_double2(0.0, 0.0);
        sx[lane] = pj.x;
        sy[lane] = pj.y;
        if (kVjp) sw[lane] = jj < n ? a.w[base + jj] : 0.0;
        jq0[lane] = 0.0;
        if (NQ == 2) jq1[lane] = 0.0;
    }
    const double wi = (kVjp && vi) ? a.w[base + i] : 0.0;
    __syncthreads();

    double q0 = 0.0, q1 = 0.0;
    const bool track_tol = a.tol != nullptr && rp.tol_cap2 > 0.0;
    unsigned long long inr_phase = 0;  // mode 4: the lanes of this phase's steps with a pair inside the radius
    const bool all_in = MODE == kModeJacDenseCut && rp.range_sq > 0.5;  // mode 4: every pair inside it

    for (int ph = 0; ph < 4; ++ph) {
        unsigned pend = 0u;  // mode 4: bit tt set when this lane's pair of step tt is inside the radius
        if constexpr (kTab) {
            // tabulated field, four steps at a time with every record load issued before any is used.
            // A record is S doubles, read as S / 4 pairs of 16-byte chunks; lanes 2k and 2k + 1 read
            // them together -- the even lane the first chunk of a pair, the odd lane the second, first
            // of the even lane's record, then of the odd lane's -- and swap what the other needs, so
            // each load instruction touches 32 cache lines instead of 64 (the vector L1 looks up one
            // line per clock, which bounds these passes).  Every lane loads: addresses of lanes past
            // the tile's edge are cells of the table too (table_cell wraps its indexes).
            constexpr int S = MODE == kModeSpeedTab ? 4 : 8;  // doubles per cell record
            constexpr int NP = S / 4;                          // chunk pairs per record
            const bool odd = lane & 1;
            for (int t0 = 0; t0 < 16; t0 += 4) {
                double2 xa[4][NP], xb[4][NP];
                double axk[4], ayk[4];
                bool actk[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int jl = lane ^ (ph * 16 + t0 + k);
                    const int j = td.J0 + jl;
                    actk[k] = vi && j < n && (!diag || lane < jl);
                    const double dx = wrap(pi.x - sx[jl]);
                    const double dy = wrap(pi.y - sy[jl]);
                    if (actk[k] && track_tol) {
                        const double d2 = dist2(dx, dy);
                        if (d2 < rp.tol_cap2) {
                            atomic_min_nonneg(a.tol + base + i, d2);
                            atomic_min_nonneg(a.tol + base + j, d2);
                        }
                    }
                    const unsigned long long own = reinterpret_cast<unsigned long long>(
                        table_cell(a.tab_s, S, a.tab_n, dx, dy, axk[k], ayk[k]));
                    const unsigned long long other = swap_adjacent_u64(own);
                    const unsigned long long half_off = odd ? 16ull : 0ull;
                    const char* ra = reinterpret_cast<const char*>((odd ? other : own) + half_off);  // even lane's record
                    const char* rb = reinterpret_cast<const char*>((odd ? own : other) + half_off);  // odd lane's record
#pragma unroll
                    for (int m = 0; m < NP; ++m) {
                        xa[k][m] = *reinterpret_cast<const double2*>(ra + 32 * m);
                        xb[k][m] = *reinterpret_cast<const double2*>(rb + 32 * m);
                    }
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // chunks 2m and 2m + 1 of this lane's record: the even lane holds chunk 2m of its own
                    // (xa) and the odd lane chunk 2m + 1 of its own (xb); each sends the other's
                    double2 ch[2 * NP];
#pragma unroll
                    for (int m = 0; m < NP; ++m) {
                        const double2 got = swap_adjacent(odd ? xa[k][m] : xb[k][m]);
                        ch[2 * m] = odd ? got : xa[k][m];
                        ch[2 * m + 1] = odd ? xb[k][m] : got;
                    }
                    const int tt = t0 + k;
                    const int jl = lane ^ (ph * 16 + tt);
                    const int j = td.J0 + jl;
                    double V = 0.0;
                    if (actk[k]) {
                        const double c4[4] = {ch[0].x, ch[0].y, ch[1].x, ch[1].y};
                        const double f = flip_if(bilinear_cell(c4, axk[k], ayk[k]), ipos != (j < half));
                        q0 += f;
                        atomicAdd(&jq0[jl], f);
                        if (MODE == kModeJacDenseTab) {
                            const double d4v[4] = {ch[2 * NP - 2].x, ch[2 * NP - 2].y, ch[2 * NP - 1].x,
                                                   ch[2 * NP - 1].y};
                            V = flip_if(rp.h * bilinear_cell(d4v, axk[k], ayk[k]), ipos != (j < half));
                            q1 += V;
                            atomicAdd(&jq1[jl], V);
                        }
                    }
                    if (kDense) stage[((lane >> 4) * 16 + (lane & 15)) * 17 + ((lane & 15) ^ tt)] = V;
                }
            }
        } else
        for (int tt = 0; tt < 16; ++tt) {
            const int jl = lane ^ (ph * 16 + tt);
            const int j = td.J0 + jl;
            const bool act = vi && j < n && (!diag || lane < jl);
            double V = 0.0;
            bool inr = false;  // COO mode: pair inside the storage radius
            if (act) {
                const double dx = wrap(pi.x - sx[jl]);
                const double dy = wrap(pi.y - sy[jl]);
                const double d2 = dist2(dx, dy);
                if (track_tol && d2 < rp.tol_cap2) {  // precision_handler.cpp:34-38
                    atomic_min_nonneg(a.tol + base + i, d2);
                    atomic_min_nonneg(a.tol + base + j, d2);
                }
                const double bj = j < half ? 1.0 : -1.0;
                const bool neg = ipos != (j < half);  // b_i b_j = -1
                double E = 0.0, R = 0.0, c = 0.0;
                field_inputs(dx, dy, E, R, c);
                double sig = 0.0, dsig = 0.0;
                // source-side sums accumulate with the target's sign and are negated once at the end
                if (MODE == kModeSpeed || kCoo) {
                    // the pair-list pass takes sigma here and d sigma for the listed pairs alone
                    // (launch_coo_derivative), so its tiles do the speed pass's work
                    pair_field<false, false>(dx, dy, E, R, c, sig, dsig);
                    const double f = flip_if(sig, neg);
                    q0 += f;
                    atomicAdd(&jq0[jl], f);
                    if (kCoo) inr = d2 < rp.range_sq;
                } else if (MODE == kModeJacDense) {
                    pair_field<true, EXACT>(dx, dy, E, R, c, sig, dsig);
                    const double f = flip_if(sig, neg);
                    V = flip_if(rp.h * dsig, neg);  // h b_j b_i d sigma: a product by +-1 is exact
                    q0 += f;
                    atomicAdd(&jq0[jl], f);
                    q1 += V;
                    atomicAdd(&jq1[jl], V);
                } else if (MODE == kModeJacDenseCut) {
                    // sigma for every pair here; d sigma of the pairs inside the storage radius after
                    // the phase's 16 steps, one pending pair per lane at a time (below), so a step costs
                    // the derivative only where the lane itself has such a pair -- or here, with sigma,
                    // when the radius holds every pair of the cell (d^2 <= 1/2 for wrapped dx, dy)
                    inr = d2 < rp.range_sq;
                    const bool now = all_in;
                    if (now) pair_field<true, EXACT>(dx, dy, E, R, c, sig, dsig);
                    else pair_field<false, EXACT>(dx, dy, E, R, c, sig, dsig);
                    const double f = flip_if(sig, neg);
                    q0 += f;
                    atomicAdd(&jq0[jl], f);
                    if (now) {
                        V = flip_if(rp.h * dsig, neg);
                        if (!rp.heaviside && d2 > rp.cut_sq) {  // simulation.cpp:469-470
                            const double e = sqrt(d2) - rp.cutoff;
                            V *= exp(-(e * e) * rp.one_per);
                        }
                        q1 += V;
                        atomicAdd(&jq1[jl], V);
                    } else if (inr) {
                        pend |= 1u << tt;
                    }
                } else {  // kVjp
                    pair_field<true, EXACT>(dx, dy, E, R, c, sig, dsig);
                    const double bb = bi * bj;
                    const double coef = (wi - sw[jl]) * bb;
                    const double gx = coef * dsig;
                    q0 += gx;
                    atomicAdd(&jq0[jl], gx);
                    if (a.want_dy) {
                        const double gy = coef * pair_field_dy(dx, E, R, c, sin(kTwoPi * dy));
                        q1 += gy;
                        atomicAdd(&jq1[jl], gy);
                    }
                }
            }
            if (kCoo) {
                // lane offsets from the ballot prefix into the wave's LDS buffer
                const unsigned long long mask = __ballot(inr);
                if (mask) {
                    const int cnt = __popcll(mask);
                    if (cfill + cnt > kCooBuf) {
                        coo_flush(a, cbuf, cfill, lane);
                        cfill = 0;
                    }
                    if (inr) {
                        CooEntry e;
                        e.i = base + i;
                        e.j = base + j;
                        e.v = 0.0;  // d sigma follows in launch_coo_derivative
                        cbuf[cfill + __popcll(mask & ((1ull << lane) - 1ull))] = e;
                    }
                    cfill += cnt;
                }
            }
            if (kDense) stage[((lane >> 4) * 16 + (lane & 15)) * 17 + ((lane & 15) ^ tt)] = V;
            if (MODE == kModeJacDenseCut) inr_phase |= __ballot(inr);
        }
        if constexpr (MODE == kModeJacDenseCut) {
            // the phase's pairs inside the storage radius: each lane takes its own in step order, all
            // lanes with one left together, until none is left.  V = h b_i b_j d sigma (the Gaussian
            // taper beyond rc without -h, simulation.cpp:461-475) into the lane's stage slot of that step
            // (where the step left 0) and into the row and column sums.
            while (__ballot(pend != 0u)) {
                if (pend) {
                    const int tp = __builtin_ctz(pend);
                    pend &= pend - 1u;
                    const int jl = lane ^ (ph * 16 + tp);
                    const int j = td.J0 + jl;
                    const double dx = wrap(pi.x - sx[jl]);
                    const double dy = wrap(pi.y - sy[jl]);
                    const double d2 = dist2(dx, dy);
                    double E = 0.0, R = 0.0, c = 0.0, sig = 0.0, dsig = 0.0;
                    field_inputs(dx, dy, E, R, c);
                    pair_field<true, EXACT>(dx, dy, E, R, c, sig, dsig);
                    double V = flip_if(rp.h * dsig, ipos != (j < half));
                    if (!rp.heaviside && d2 > rp.cut_sq) {  // simulation.cpp:469-470
                        const double e = sqrt(d2) - rp.cutoff;
                        V *= exp(-(e * e) * rp.one_per);
                    }
                    q1 += V;
                    atomicAdd(&jq1[jl], V);
                    stage[((lane >> 4) * 16 + (lane & 15)) * 17 + ((lane & 15) ^ tp)] = V;
                }
            }
        }
        if (kDense) {
            __syncthreads();
            const bool write = !(MODE == kModeJacDenseCut && a.skip_empty && inr_phase == 0);
            inr_phase = 0;
            const int g = lane >> 4, q = lane & 15;
            const int ib = td.I0 + 16 * g;
            const int jb = td.J0 + 16 * (g ^ ph);
            const double* st1 = stage + (g * 16) * 17 + q;  // stage[g][k][q], k = 0..15
            const double* st2 = stage + (g * 16 + q) * 17;  // stage[g][q][k]
            // V(ib+k, jb+q) -> column ib+k, row jb+q;  V(ib+q, jb+k) -> column jb+k, row ib+q
            double* p1 = Mb + (int64_t)ib * ld + jb + q;
            double* p2 = Mb + (int64_t)jb * ld + ib + q;
            if (!write) {
                // no pair of this phase inside the radius: nothing the caller reads
            } else if (!diag && td.I0 + kTile <= n && td.J0 + kTile <= n) {
                // interior tile: every entry valid
#pragma unroll 4
                for (int k = 0; k < 16; ++k) {
                    // the matrix is written once and read by a later launch: streamed past the L2, where
                    // the tabulated field's cells and the positions stay
                    __builtin_nontemporal_store(st1[k * 17], p1);
                    __builtin_nontemporal_store(st2[k], p2);
                    p1 += ld;
                    p2 += ld;
                }
            } else {
                for (int k = 0; k < 16; ++k) {
                    const int i1 = ib + k, j1 = jb + q;
                    if (i1 < n && j1 < n && (!diag || i1 < j1)) __builtin_nontemporal_store(st1[k * 17], p1);
                    const int i2 = ib + q, j2 = jb + k;
                    if (i2 < n && j2 < n && (!diag || i2 < j2)) __builtin_nontemporal_store(st2[k], p2);
                    p1 += ld;
                    p2 += ld;
                }
            }
            __syncthreads();
        }
    }
    if (kCoo && cfill > 0) coo_flush(a, cbuf, cfill, lane);
    __syncthreads();
    // quantity 0 and quantity 1 (the dense modes' -sum V, the VJP's y part) of the targets, then of the
    // sources, whose sums were taken with the target's sign
    double* P = a.part + (size_t)td.slot * NQ * 128;
    P[lane] = q0;
    P[64 + lane] = -jq0[lane];
    if (NQ == 2) {
        P[128 + lane] = kVjp ? q1 : -q1;
        P[192 + lane] = -jq1[lane];
    }
}


// d sigma of the pairs a pair-list pass listed, one thread per entry (the entries past the
// pass's count, or past the list's capacity, are left alone):
// V = h b_i b_j d sigma(x_i - x_j, y_i - y_j), i < j, times exp(-(r - rc)^2 / rc^2) beyond rc when the
// cutoff is not a step (simulation.cpp:461-470) -- the value the reference stores for the pair.
template <bool EXACT>
__global__ void coo_derivative_kernel(CooDerivArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= min(a.cap, *a.count)) return;
    CooEntry& e = a.coo[k];
    const int r = a.elem ? a.elem[e.i] : 0;
    const int base = a.off ? a.off[r] : 0;
    const int half = (a.off ? a.off[r + 1] - base : a.n0) >> 1;
    const RealParams rp = a.rp ? a.rp[r] : a.rp0;
    const double2 pi = a.xy[e.i], pj = a.xy[e.j];
    const double dx = wrap(pi.x - pj.x), dy = wrap(pi.y - pj.y);
    const double d2 = dist2(dx, dy);
    double E, R, c, sig = 0.0, dsig = 0.0;
    field_inputs(dx, dy, E, R, c);
    pair_field<true, EXACT>(dx, dy, E, R, c, sig, dsig);
    const double bi = e.i - base < half ? 1.0 : -1.0, bj = e.j - base < half ? 1.0 : -1.0;
    double V = rp.h * bj * bi * dsig;
    if (!rp.heaviside && d2 > rp.cut_sq) {  // simulation.cpp:469-470
        const double t = sqrt(d2) - rp.cutoff;
        V *= exp(-(t * t) * rp.one_per);
    }
    e.v = V;
}

hipError_t launch_coo_derivative(const CooDerivArgs& a, int exact, hipStream_t s)
{
    if (a.cap <= 0) return hipSuccess;
    const dim3 g((a.cap + 255) / 256);
    if (exact) hipLaunchKernelGGL(coo_derivative_kernel<true>, g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(coo_derivative_kernel<false>, g, dim3(256), 0, s, a);
    return hipGetLastError();
}

// 64 threads (one wavefront) per block; at least 4 waves per SIMD (<= 128
// VGPRs) for the sigma-only pass, 3 (<= 168 VGPRs) for the dense pass, 2 for
// the pair-list, dense-with-cutoff and adjoint passes (their extra live values
// would spill at 168).  PERSIST: the grid strides over the tile table.
template <int MODE, bool EXACT, bool PERSIST>
__global__ __launch_bounds__(64, (MODE == kModeSpeed || MODE == kModeSpeedTab || MODE == kModeJacCoo)
                                     ? 4
                                     : ((MODE == kModeJacDense || MODE == kModeJacDenseTab) ? 3 : 2)) void pair_pass_kernel(
    PassArgs a)
{
    if constexpr (PERSIST) {
        for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
            pair_tile<MODE, EXACT>(a, a.tiles[t]);
            __syncthreads();
        }
    } else {
        pair_tile<MODE, EXACT>(a, a.tiles[blockIdx.x]);
    }
}

template <int MODE>
static hipError_t launch_mode(const PassArgs& a, hipStream_t s, int grid)
{
    if (a.ntiles <= 0) return hipSuccess;
    if (grid > 0) {
        const dim3 g(std::min(grid, a.ntiles));
        if (a.exact) hipLaunchKernelGGL((pair_pass_kernel<MODE, true, true>), g, dim3(kTile), 0, s, a);
        else hipLaunchKernelGGL((pair_pass_kernel<MODE, false, true>), g, dim3(kTile), 0, s, a);
    } else if (a.exact) {
        hipLaunchKernelGGL((pair_pass_kernel<MODE, true, false>), dim3(a.ntiles), dim3(kTile), 0, s, a);
    } else {
        hipLaunchKernelGGL((pair_pass_kernel<MODE, false, false>), dim3(a.ntiles), dim3(kTile), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_pair_pass(int mode, const PassArgs& a, hipStream_t s, int grid)
{
    switch (mode) {
    case kModeSpeed: return launch_mode<kModeSpeed>(a, s, grid);
    case kModeJacDense: return launch_mode<kModeJacDense>(a, s, grid);
    case kModeJacCoo: return launch_mode<kModeJacCoo>(a, s, grid);
    case kModeVjp: return launch_mode<kModeVjp>(a, s, grid);
    case kModeJacDenseCut: return launch_mode<kModeJacDenseCut>(a, s, grid);
    case kModeSpeedTab: return launch_mode<kModeSpeedTab>(a, s, grid);
    case kModeJacDenseTab: return launch_mode<kModeJacDenseTab>(a, s, grid);
    default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------- tabulated field

__global__ void table_build_kernel(int n, double* tab_s, double* tab_d)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)n * n) return;
    const int kx = (int)(k % n), ky = (int)(k / n);
    const double dx = -0.5 + (double)kx / n, dy = -0.5 + (double)ky / n;  // exact for n a power of two
    double s = 0.0, d = 0.0;
    if (dx != 0.0 || dy != 0.0) {  // the field is singular at the origin; its node holds 0
        double E, R, c;
        field_inputs(dx, dy, E, R, c);
        pair_field<true, false>(dx, dy, E, R, c, s, d);
    }
    tab_s[k] = s;
    if (tab_d) tab_d[k] = d;
}

__global__ void table_records_kernel(int n, const double* tab_s, const double* tab_d, double* rec)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)n * n) return;
    const int kx = (int)(k % n), ky = (int)(k / n);
    const int kx1 = (kx + 1) & (n - 1), ky1 = (ky + 1) & (n - 1);
    const int64_t c00 = (int64_t)ky * n + kx, c10 = (int64_t)ky * n + kx1, c01 = (int64_t)ky1 * n + kx,
                  c11 = (int64_t)ky1 * n + kx1;
    double* r = rec + k * (tab_d ? 8 : 4);
    r[0] = tab_s[c00];
    r[1] = tab_s[c10];
    r[2] = tab_s[c01];
    r[3] = tab_s[c11];
    if (tab_d) {
        r[4] = tab_d[c00];
        r[5] = tab_d[c10];
        r[6] = tab_d[c01];
        r[7] = tab_d[c11];
    }
}

hipError_t launch_table_records(int n, const double* tab_s, const double* tab_d, double* rec, hipStream_t s)
{
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(table_records_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, n, tab_s, tab_d,
                       rec);
    return hipGetLastError();
}

hipError_t launch_table_build(int n, double* tab_s, double* tab_d, hipStream_t s)
{
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(table_build_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, n, tab_s, tab_d);
    return hipGetLastError();
}

__global__ void table_eval_kernel(int64_t m, const double* dx, const double* dy, const double* tab_s,
                                  const double* tab_d, int tab_n, double* s_out, double* d_out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double x = wrap(dx[k]), y = wrap(dy[k]);
    if (s_out) s_out[k] = table_bilinear(tab_s, tab_n, x, y);
    if (d_out) d_out[k] = table_bilinear(tab_d, tab_n, x, y);
}

hipError_t launch_table_eval(int64_t m, const double* dx, const double* dy, const double* tab_s, const double* tab_d,
                             int tab_n, double* s_out, double* d_out, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(table_eval_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, dx, dy, tab_s, tab_d,
                       tab_n, s_out, d_out);
    return hipGetLastError();
}

// ------------------------------------------------------------ field eval

__global__ void field_eval_kernel(int64_t n, const double* dx, const double* dy, const double* E, const double* c,
                                  double* sxy, double* dsxy, int exact)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double x = dx[k], y = dy[k];
    double e, r, cc;
    if (E && c) {  // the 4-arg overloads: R as the reference forms it, 1/E
        e = E[k];
        cc = c[k];
        r = 1.0 / e;
    } else {
        field_inputs(x, y, e, r, cc);
    }
    double s = 0.0, ds = 0.0;
    if (dsxy) {
        if (exact) pair_field<true, true>(x, y, e, r, cc, s, ds);
        else pair_field<true, false>(x, y, e, r, cc, s, ds);
        dsxy[k] = ds;
    } else {
        pair_field<false, false>(x, y, e, r, cc, s, ds);
    }
    if (sxy) sxy[k] = s;
}

hipError_t launch_field_eval(int64_t n, const double* dx, const double* dy, const double* E, const double* c,
                             double* sxy, double* dsxy, int exact, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int bs = 256;
    hipLaunchKernelGGL(field_eval_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, n, dx, dy, E, c,
                       sxy, dsxy, exact);
    return hipGetLastError();
}

// ------------------------------------------------------------- reduction

// One block of kReduceWaves wavefronts per 64-element block b of a realization:
// wave w adds the partial items w, w + W, w + 2W, ... (items: target-side sums
// of tiles (b, J), J >= b, then source-side sums of tiles (I, b), I <= b), and
// wave 0 adds the W wave sums in order: a fixed order, so runs are bitwise
// repeatable.
constexpr int kReduceWaves = 8;

__global__ __launch_bounds__(64 * kReduceWaves) void reduce_kernel(ReduceArgs a)
{
#pragma clang fp contract(off)
    __shared__ double sm[kReduceWaves][2][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int r = 0, b = blockIdx.x;
    if (a.blocks) {
        r = a.blocks[blockIdx.x].x;
        b = a.blocks[blockIdx.x].y;
    }
    const int base = a.off ? a.off[r] : 0;
    const int n = a.off ? a.off[r + 1] - base : a.total;
    const RealParams rp = a.rp ? a.rp[r] : a.rp0;
    const int nb = rp.nb;
    const int items = nb + 1;
    const int row0 = rp.slot0 + b * nb - b * (b - 1) / 2;  // slot of tile (b, b)
    double acc[2] = {0.0, 0.0};
    for (int it = w; it < items; it += kReduceWaves) {
        size_t off;
        if (it < nb - b) {  // target side of tile (b, b + it)
            off = (size_t)(row0 + it) * a.nq * 128 + lane;
        } else {           // source side of tile (I, b), I = it - (nb - b)
            const int I = it - (nb - b);
            const int slot = rp.slot0 + I * nb - I * (I - 1) / 2 + (b - I);
            off = (size_t)slot * a.nq * 128 + 64 + lane;
        }
        acc[0] += a.part[off];
        if (a.nq == 2) acc[1] += a.part[off + 128];
    }
    sm[w][0][lane] = acc[0];
    sm[w][1][lane] = acc[1];
    __syncthreads();
    if (w != 0) return;
    const int k = b * kTile + lane;
    if (k >= n) return;
    double sum[2] = {0.0, 0.0};
    for (int q = 0; q < 2; ++q)
        for (int v = 0; v < kReduceWaves; ++v) sum[q] += sm[v][q][lane];
    const int g = base + k;
    const double bsg = k < (n >> 1) ? 1.0 : -1.0;
    if (a.vA) a.vA[g] = a.add_tau ? sum[0] + rp.tauA * bsg : sum[0];
    if (a.vB) a.vB[g] = a.add_tau ? sum[0] + rp.tauB * bsg : sum[0];
    if (a.q1) a.q1[g] = sum[1];
    if (a.dvec) a.dvec[g] = weight(sum[1], (char)rp.wfunc);
    if (a.M) {
        if (a.moff) a.M[a.moff[r] + (int64_t)k * n + k] = sum[1];
        else a.M[(int64_t)k * a.ld + k] = sum[1];
    }
}

hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s)
{
    if (a.total <= 0) return hipSuccess;
    const int nblocks = a.blocks ? a.nblocks : (a.total + kTile - 1) / kTile;
    hipLaunchKernelGGL(reduce_kernel, dim3(nblocks), dim3(64 * kReduceWaves), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ CSC

// ncoo_dev (when not null): the pair count the pass left on the device,
// capped by ncoo (the list capacity); the grid then covers the capacity
__global__ void csc_count_kernel(const CooEntry* coo, int ncoo, const int* ncoo_dev, int* cnt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (ncoo_dev) ncoo = min(ncoo, *ncoo_dev);
    if (k >= ncoo) return;
    const CooEntry e = coo[k];
    atomicAdd(&cnt[e.i], 1);
    if (e.v != 0.0) atomicAdd(&cnt[e.j], 1);  // simulation.cpp:365: exact zeros are not mirrored
}

hipError_t launch_csc_count(const CooEntry* coo, int ncoo, int* cnt, int total, hipStream_t s, const int* ncoo_dev)
{
    (void)total;
    if (ncoo <= 0) return hipSuccess;
    hipLaunchKernelGGL(csc_count_kernel, dim3((ncoo + 255) / 256), dim3(256), 0, s, coo, ncoo, ncoo_dev, cnt);
    return hipGetLastError();
}

// single-workgroup exclusive scan: Ap[0] = 0, Ap[k+1] = Ap[k] + cnt[k]
__global__ __launch_bounds__(1024) void scan_kernel(const int* cnt, int* Ap, int total)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (total + 1023) / 1024;
    const int lo = min(total, t * chunk), hi = min(total, lo + chunk);
    int s = 0;
    for (int k = lo; k < hi; ++k) s += cnt[k];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;  // exclusive
    if (t == 0) Ap[0] = 0;
    for (int k = lo; k < hi; ++k) {
        run += cnt[k];
        Ap[k + 1] = run;
    }
}

// Longer scans in three launches: per block of kScanBlock counts its total,
// the exclusive scan of the totals (one workgroup), then each block's own
// exclusive scan offset by its total's prefix.  Integer sums: any grouping
// gives the same Ap.
constexpr int kScanBlock = 2048;  // 256 threads x 8 counts

__device__ inline int block_exclusive_scan(int v, int* part)  // 256 threads; returns the sum of the values before v
{
    const int t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int u = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    return part[t] - v;
}

__global__ __launch_bounds__(256) void scan_totals_kernel(const int* cnt, int total, int* bsum)
{
    __shared__ int part[256];
    const int k0 = blockIdx.x * kScanBlock + 8 * (int)threadIdx.x;
    int s = 0;
    for (int k = k0; k < min(k0 + 8, total); ++k) s += cnt[k];
    block_exclusive_scan(s, part);
    if (threadIdx.x == 255) bsum[blockIdx.x] = part[255];
}

__global__ __launch_bounds__(256) void scan_apply_kernel(const int* cnt, int total, const int* boff, int* Ap)
{
    __shared__ int part[256];
    const int k0 = blockIdx.x * kScanBlock + 8 * (int)threadIdx.x;
    int v[8], s = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        v[q] = k0 + q < total ? cnt[k0 + q] : 0;
        s += v[q];
    }
    int run = boff[blockIdx.x] + block_exclusive_scan(s, part);
    if (blockIdx.x == 0 && threadIdx.x == 0) Ap[0] = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (k0 + q < total) {
            run += v[q];
            Ap[k0 + q + 1] = run;
        }
}

hipError_t launch_scan(const int* cnt, int* Ap, int total, hipStream_t s, int* scratch)
{
    const int nblk = (total + kScanBlock - 1) / kScanBlock;
    if (total <= 16 * kScanBlock || !scratch) {
        hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, cnt, Ap, total);
        return hipGetLastError();
    }
    // scratch: nblk totals, then their nblk + 1 exclusive prefix sums
    int* bsum = scratch;
    int* boff = scratch + nblk;
    hipLaunchKernelGGL(scan_totals_kernel, dim3(nblk), dim3(256), 0, s, cnt, total, bsum);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, bsum, boff, nblk);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nblk), dim3(256), 0, s, cnt, total, boff, Ap);
    return hipGetLastError();
}

__global__ void csc_diag_kernel(const int* Ap, int* cursor, int* Ai, double* Ax, int total)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int p = Ap[c];
    Ai[p] = c;
    Ax[p] = 0.0;
    cursor[c] = p + 1;
}

__global__ void csc_fill_kernel(const CooEntry* coo, int ncoo, const int* ncoo_dev, int* cursor, int* Ai, double* Ax)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (ncoo_dev) ncoo = min(ncoo, *ncoo_dev);
    if (k >= ncoo) return;
    const CooEntry e = coo[k];
    int p = atomicAdd(&cursor[e.i], 1);
    Ai[p] = e.j;
    Ax[p] = e.v;
    if (e.v != 0.0) {
        p = atomicAdd(&cursor[e.j], 1);
        Ai[p] = e.i;
        Ax[p] = e.v;
    }
}

hipError_t launch_csc_fill(const CooEntry* coo, int ncoo, const int* Ap, int* cursor, int* Ai, double* Ax,
                           int total, hipStream_t s, const int* ncoo_dev)
{
    hipLaunchKernelGGL(csc_diag_kernel, dim3((total + 255) / 256), dim3(256), 0, s, Ap, cursor, Ai, Ax, total);
    if (ncoo > 0)
        hipLaunchKernelGGL(csc_fill_kernel, dim3((ncoo + 255) / 256), dim3(256), 0, s, coo, ncoo, ncoo_dev, cursor,
                           Ai, Ax);
    return hipGetLastError();
}

__global__ void fill_int_kernel(int* p, int v, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// The upper-triangle tile table of one realization of n, row-major (slot = index), as
// d4_ctx::single_tiles builds it on the host: tile t is (I, J) with I the largest row whose first
// index I nb - I (I - 1) / 2 is at most t.
__global__ void tile_table_kernel(TileDesc* t, int nb, int nt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    // first index of row I: I (2 nb - I + 1) / 2; estimate I from the quadratic, then correct it
    const double b = 2.0 * nb + 1.0;
    int I = (int)floor((b - sqrt(b * b - 8.0 * k)) / 2.0);
    I = max(0, min(I, nb - 1));
    while (I > 0 && I * (2 * nb - I + 1) / 2 > k) --I;
    while (I + 1 < nb && (I + 1) * (2 * nb - I) / 2 <= k) ++I;
    const int J = I + (k - I * (2 * nb - I + 1) / 2);
    t[k] = TileDesc{0, I * kTile, J * kTile, k};
}

hipError_t launch_tile_table(TileDesc* t, int n, hipStream_t s)
{
    const int nb = (n + kTile - 1) / kTile;
    const int nt = nb * (nb + 1) / 2;
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(tile_table_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, t, nb, nt);
    return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int v, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// Per column: sort rows ascending (the order getElement + the j > i loop
// produce, simulation.cpp:362-374, 444-478), then the diagonal
// s = -sum of the column in that order (simulation.cpp:370, 475, 512) and
// d = weight(s) (simulation.cpp:513).  One wavefront per column: for up to 64
// entries lane k holds entry k, counts the rows below its own (the rows of a
// column are distinct) and writes the entry to that rank; the sum then runs
// over the sorted entries in ascending order, the same for every lane.
// Longer columns are sorted by lane 0 in place (shell sort).
__device__ inline int csc_lane_int(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__global__ __launch_bounds__(256) void csc_finish_kernel(const int* Ap, int* Ai, double* Ax, int* indexes,
                                                         double* dvec, const int* elem_real, const int* wf_arr,
                                                         int wf0, int total)
{
    const int c = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c >= total) return;  // uniform over the wavefront
    const int lo = Ap[c], hi = Ap[c + 1], len = hi - lo;
    double sub = 0.0;
    int dpos = lo;
    if (len <= 64) {
        const int key = lane < len ? Ai[lo + lane] : INT_MAX;
        const double val = lane < len ? Ax[lo + lane] : 0.0;
        int rank = 0;
        for (int j = 0; j < len; ++j) rank += csc_lane_int(key, j) < key ? 1 : 0;
        // the entries in sorted order, lane r holding rank r
        const int to = 4 * min(rank, 63);
        const int sk = __builtin_amdgcn_ds_permute(to, key);
        const int slo = __builtin_amdgcn_ds_permute(to, __double2loint(val));
        const int shi = __builtin_amdgcn_ds_permute(to, __double2hiint(val));
        for (int r = 0; r < len; ++r) {
            if (csc_lane_int(sk, r) == c) dpos = lo + r;
            else sub -= __hiloint2double(csc_lane_int(shi, r), csc_lane_int(slo, r));
        }
        if (lane < len) {
            Ai[lo + rank] = key;
            Ax[lo + rank] = key == c ? sub : val;
        }
    } else if (lane == 0) {
        // shell sort with a fixed gap sequence ending in an insertion pass
        const int gaps[8] = {701, 301, 132, 57, 23, 10, 4, 1};
        for (int gi = 0; gi < 8; ++gi) {
            const int gap = gaps[gi];
            if (gap >= len) continue;
            for (int k = lo + gap; k < hi; ++k) {
                const int key = Ai[k];
                const double val = Ax[k];
                int m = k;
                while (m - gap >= lo && Ai[m - gap] > key) {
                    Ai[m] = Ai[m - gap];
                    Ax[m] = Ax[m - gap];
                    m -= gap;
                }
                Ai[m] = key;
                Ax[m] = val;
            }
        }
        for (int k = lo; k < hi; ++k) {
            if (Ai[k] == c) dpos = k;
            else sub -= Ax[k];
        }
        Ax[dpos] = sub;
    }
    if (lane == 0) {
        indexes[c] = dpos;
        const int wf = wf_arr ? wf_arr[elem_real ? elem_real[c] : 0] : wf0;
        dvec[c] = weight(sub, (char)wf);
    }
}

hipError_t launch_csc_finish(const int* Ap, int* Ai, double* Ax, int* indexes, double* dvec, const int* elem_real,
                             const int* wf_arr, int wf0, int total, hipStream_t s)
{
    hipLaunchKernelGGL(csc_finish_kernel, dim3((total + 3) / 4), dim3(256), 0, s, Ap, Ai, Ax, indexes, dvec,
                       elem_real, wf_arr, wf0, total);
    return hipGetLastError();
}

// simulation.cpp:516-522
__global__ void csc_scale_kernel(const int* Ap, const int* Ai, double* Ax, const int* indexes, const double* dvec,
                                 int total)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= total) return;
    for (int k = Ap[c]; k < Ap[c + 1]; ++k) Ax[k] *= (1 + dvec[Ai[k]]) / 2;
    Ax[indexes[c]] += 1;
}

hipError_t launch_csc_scale(const int* Ap, const int* Ai, double* Ax, const int* indexes, const double* dvec,
                            int total, hipStream_t s)
{
    hipLaunchKernelGGL(csc_scale_kernel, dim3((total + 255) / 256), dim3(256), 0, s, Ap, Ai, Ax, indexes, dvec,
                       total);
    return hipGetLastError();
}

// ------------------------------------------------------ stage-II rescale

__global__ void rescale_weights_kernel(RescaleArgs a)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.total) return;
    const int r = a.elem_real ? a.elem_real[g] : 0;
    const double ow = a.dvec[g];
    const char T = (char)(a.wf_arr ? a.wf_arr[r] : a.wf0);
    const double factor = a.factor_arr ? a.factor_arr[r] : a.factor0;
    a.nw[g] = ow == 0 ? 0.0 : weight(weight_inv(ow, T) * factor, T);  // simulation.cpp:559-572
}

__global__ void rescale_entries_kernel(RescaleArgs a)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.total) return;
    for (int k = a.Ap[c]; k < a.Ap[c + 1]; ++k) {  // simulation.cpp:575-594
        const int row = a.Ai[k];
        const double f = (1 + a.nw[row]) / (1 + a.dvec[row]);
        if (row == c) a.Ax[k] = 1 + (a.Ax[k] - 1) / 2 * f;
        else a.Ax[k] = a.Ax[k] / 2 * f;
    }
}

__global__ void rescale_commit_kernel(RescaleArgs a)
{
#pragma clang fp contract(off)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.total) return;
    const int r = a.elem_real ? a.elem_real[g] : 0;
    const int base = a.off ? a.off[r] : 0;
    const int n = a.off ? a.off[r + 1] - base : a.total;
    const int k = g - base;
    a.dvec[g] = a.nw[g];
    const double dtau = a.dtau_arr ? a.dtau_arr[r] : a.dtau0;
    if (a.speed) a.speed[g] += dtau * (k < (n >> 1) ? 1.0 : -1.0);  // simulation.cpp:602-607
}

hipError_t launch_rescale(const RescaleArgs& a, hipStream_t s)
{
    if (a.total <= 0) return hipSuccess;
    const dim3 g((a.total + 255) / 256), b(256);
    hipLaunchKernelGGL(rescale_weights_kernel, g, b, 0, s, a);
    hipLaunchKernelGGL(rescale_entries_kernel, g, b, 0, s, a);
    hipLaunchKernelGGL(rescale_commit_kernel, g, b, 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------ dense final scale

// Each thread owns two consecutive rows (one 16-byte access per column when VEC: n and ld even, M
// 16-byte aligned) and keeps their factors (1 + d)/2 for the kScaleCols columns of its block.  The
// interior blocks issue all their columns' loads before the first store (the compiler cannot move a
// load past a store to the same array by itself), and the matrix streams past the caches both ways:
// it is read and written once.
// Unfused, as simulation.cpp:516-522 rounds it: A = M (1 + d_row)/2, then + 1 on the diagonal.
constexpr int kScaleCols = 8;

template <bool VEC>
__global__ __launch_bounds__(256) void dense_scale_kernel(double* M, int64_t ld, int n, const double* dvec)
{
#pragma clang fp contract(off)
    const int row = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (row >= n) return;
    const bool two = row + 1 < n;
    const double f0 = (1 + dvec[row]) / 2, f1 = two ? (1 + dvec[row + 1]) / 2 : 0.0;
    const int c0 = blockIdx.y * kScaleCols, c1 = min(c0 + kScaleCols, n);
    if (VEC && c1 - c0 == kScaleCols) {
        double2* p = reinterpret_cast<double2*>(M + (int64_t)c0 * ld + row);
        const int64_t st = ld / 2;
        double2 v[kScaleCols];
#pragma unroll
        for (int k = 0; k < kScaleCols; ++k) {
            v[k].x = __builtin_nontemporal_load(&p[k * st].x);
            v[k].y = __builtin_nontemporal_load(&p[k * st].y);
        }
#pragma unroll
        for (int k = 0; k < kScaleCols; ++k) {
            double a = v[k].x * f0, b = v[k].y * f1;
            if (row == c0 + k) a = a + 1;
            if (row + 1 == c0 + k) b = b + 1;
            __builtin_nontemporal_store(a, &p[k * st].x);
            __builtin_nontemporal_store(b, &p[k * st].y);
        }
        return;
    }
    for (int c = c0; c < c1; ++c) {
        double* p = M + (int64_t)c * ld + row;
        if (VEC) {
            double2 v = *reinterpret_cast<const double2*>(p);
            v.x = v.x * f0;
            v.y = v.y * f1;
            if (row == c) v.x = v.x + 1;
            if (row + 1 == c) v.y = v.y + 1;
            *reinterpret_cast<double2*>(p) = v;
        } else {
            double v = p[0] * f0;
            if (row == c) v = v + 1;
            p[0] = v;
            if (two) {
                double w = p[1] * f1;
                if (row + 1 == c) w = w + 1;
                p[1] = w;
            }
        }
    }
}

hipError_t launch_dense_scale(double* M, int64_t ld, int n, const double* dvec, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const dim3 g((n + 511) / 512, (n + kScaleCols - 1) / kScaleCols);
    if (n % 2 == 0 && ld % 2 == 0 && reinterpret_cast<uintptr_t>(M) % 16 == 0)
        hipLaunchKernelGGL(dense_scale_kernel<true>, g, dim3(256), 0, s, M, ld, n, dvec);
    else
        hipLaunchKernelGGL(dense_scale_kernel<false>, g, dim3(256), 0, s, M, ld, n, dvec);
    return hipGetLastError();
}

// ------------------------------------------------ batched dense matrices

// x: global column index g (realization r, local column c), y: 256-row chunks
__global__ void dense_scale_batched_kernel(DenseBatch b, const double* M, double* F, const double* dvec)
{
#pragma clang fp contract(off)
    const int g = blockIdx.x;
    const int r = b.elem_real[g];
    const int base = b.off[r], n = b.off[r + 1] - base, c = g - base;
    const int row = blockIdx.y * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int64_t k = b.moff[r] + (int64_t)c * n + row;
    double v = M[k] * ((1 + dvec[base + row]) / 2);  // simulation.cpp:516-522
    if (row == c) v += 1;
    F[k] = v;
}

hipError_t launch_dense_scale_batched(const DenseBatch& b, const double* M, double* F, const double* dvec,
                                      hipStream_t s)
{
    if (b.total <= 0) return hipSuccess;
    hipLaunchKernelGGL(dense_scale_batched_kernel, dim3(b.total, (b.maxn + 255) / 256), dim3(256), 0, s, b, M, F,
                       dvec);
    return hipGetLastError();
}

__global__ void dense_rescale_batched_kernel(DenseBatch b, const double* M, double* F, const double* dold,
                                             const double* dnew)
{
#pragma clang fp contract(off)
    const int g = blockIdx.x;
    const int r = b.elem_real[g];
    const int base = b.off[r], n = b.off[r + 1] - base, c = g - base;
    const int row = blockIdx.y * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int64_t k = b.moff[r] + (int64_t)c * n + row;
    const double d0 = dold[base + row];
    double v = M[k] * ((1 + d0) / 2);  // the stage-I entry, formed as the scale kernel forms it
    if (row == c) v += 1;
    const double f = (1 + dnew[base + row]) / (1 + d0);  // simulation.cpp:575-594
    F[k] = row == c ? 1 + (v - 1) / 2 * f : v / 2 * f;
}

hipError_t launch_dense_rescale_batched(const DenseBatch& b, const double* M, double* F, const double* dold,
                                        const double* dnew, hipStream_t s)
{
    if (b.total <= 0) return hipSuccess;
    hipLaunchKernelGGL(dense_rescale_batched_kernel, dim3(b.total, (b.maxn + 255) / 256), dim3(256), 0, s, b, M, F,
                       dold, dnew);
    return hipGetLastError();
}

}  // namespace d4
