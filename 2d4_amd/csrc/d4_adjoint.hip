// d4_adjoint.hip -- the pair term of the vector-Jacobian product of the reference's implicit step
// (simulation.cpp:98-154): how the step matrix's entries move with the positions.
//
// The step matrix is A = I + diag((1 + d)/2) (V + diag(s)) with V_ij = k b_i b_j m(x_i - x_j, y_i - y_j),
// m = xy_diff_x times the storage weight (simulation.cpp:458-470, zero outside the storage radius),
// s_i = -sum_j V_ij and d_i = weight(s_i).  Given the adjoint of V's off-diagonal entries in the
// symmetric form P_ij = Vbar_ij + Vbar_ji (built from per-dislocation vectors, below), the positions
// receive, for each pair, the derivative of its one evaluation: the reference evaluates the pair i < j
// at (x_i - x_j, y_i - y_j) and stores the value for both V_ij and V_ji (simulation.cpp:362-374,
// 440-478), and m is not exactly even (its truncated, cosh-for-sinh outer walls), so
//     xbar_l = sum_{j > l} k b_l b_j (d m / d dx)(x_l - x_j, y_l - y_j) P_lj
//            - sum_{j < l} k b_l b_j (d m / d dx)(x_j - x_l, y_j - y_l) P_lj,
//     ybar_l = the same with d m / d dy.
//
// P_lj = -(a_l u_j + a_j u_l + c_l w_j + c_j w_l) - t_l - t_j, from five vectors the host side forms
// (a, c: the row-scaled solves' adjoints, u, w: the two Newton corrections, t: the adjoint of s).
//
// One wavefront per target dislocation: lane q takes the sources j = q, q + 64, ..., and the wavefront's
// 64 partial sums are combined by a fixed butterfly, so the result does not depend on scheduling.
#include "d4_dual.h"
#include "d4_kernels.h"

namespace d4 {

namespace {

__global__ __launch_bounds__(256) void theta_pair_vjp_kernel(ThetaVjpArgs a)
{
    const int l = blockIdx.x * 4 + (threadIdx.x >> 6);  // the same in every lane of a wavefront
    const int lane = threadIdx.x & 63;
    if (l >= a.n) return;
    const RealParams& rp = a.rp;
    const int half = a.n >> 1;
    const double2 pl = a.xy[l];
    const double bl = l < half ? 1.0 : -1.0;
    const double al = a.a[l], cl = a.c[l], ul = a.u[l], wl = a.w[l], tl = a.t[l];
    double gx = 0.0, gy = 0.0;
    for (int j = lane; j < a.n; j += 64) {
        if (j == l) continue;
        const double2 pj = a.xy[j];
        const bool lower = l < j;  // the pair's evaluation runs from the lower index to the higher
        const double dx = lower ? wrap(pl.x - pj.x) : wrap(pj.x - pl.x);
        const double dy = lower ? wrap(pl.y - pj.y) : wrap(pj.y - pl.y);
        const double d2 = dist2(dx, dy);
        if (!(d2 < rp.range_sq)) continue;  // no matrix entry (the forward pass's test)
        const Dual2 f = dxy_diff_x(dx, dy);
        const Dual2 wt = dstorage_weight(Dual2{d2, 2.0 * dx, 2.0 * dy}, rp.cutoff, rp.cut_sq, rp.one_per, rp.heaviside);
        const Dual2 m = f * wt;
        const double kbb = (lower ? 1.0 : -1.0) * (bl * (j < half ? 1.0 : -1.0)) * rp.h;
        const double P = -(al * a.u[j] + a.a[j] * ul + cl * a.w[j] + a.c[j] * wl) - tl - a.t[j];
        gx = fma(kbb * m.x, P, gx);
        gy = fma(kbb * m.y, P, gy);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        gx += __shfl_xor(gx, o);
        gy += __shfl_xor(gy, o);
    }
    if (lane == 0) {
        a.gx[l] = gx;
        if (a.gy) a.gy[l] = gy;
    }
}

}  // namespace

hipError_t launch_theta_pair_vjp(const ThetaVjpArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(theta_pair_vjp_kernel, dim3((a.n + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace d4
