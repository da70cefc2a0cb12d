// hostmath.hip -- test-only: the kernels' per-pair math (2d4_amd/csrc/d4_field.h)
// compiled for the host, so its arithmetic can be checked against the oracle on
// a machine without a GPU.  Not part of lib2d4.
#include "../../2d4_amd/csrc/d4_field.h"
#include "../../2d4_amd/csrc/d4_dual.h"

extern "C" {
void hm_field(long n, const double* dx, const double* dy, const double* e, const double* c, double* s, double* ds,
              int exact)
{
    for (long k = 0; k < n; ++k) {
        double E, R, C;
        if (e && c) { E = e[k]; C = c[k]; R = 1.0 / E; }
        else d4::field_inputs(dx[k], dy[k], E, R, C);
        double a = 0, b = 0;
        if (exact) d4::pair_field<true, true>(dx[k], dy[k], E, R, C, a, b);
        else d4::pair_field<true, false>(dx[k], dy[k], E, R, C, a, b);
        s[k] = a;
        ds[k] = b;
    }
}
void hm_field_dy(long n, const double* dx, const double* dy, double* out)
{
    for (long k = 0; k < n; ++k) {
        double E, R, C;
        d4::field_inputs(dx[k], dy[k], E, R, C);
        out[k] = d4::pair_field_dy(dx[k], E, R, C, sin(d4::kTwoPi * dy[k]));
    }
}
double hm_wrap(double v) { return d4::wrap(v); }
// v / u as the triangular-solve kernel forms it off its step chain: q = v r with r = 1/u (correctly
// rounded), then q + (v - q u) r with the residual exact by fma
void hm_div_by_reciprocal(long n, const double* v, const double* u, double* out)
{
    for (long k = 0; k < n; ++k) {
        const double r = 1.0 / u[k], q = v[k] * r;
        out[k] = fma(fma(-q, u[k], v[k]), r, q);
    }
}
// the reference's d sigma / dx and its derivatives in dx and dy (d4_dual.h)
void hm_dsigma_dual(long n, const double* dx, const double* dy, double* v, double* ddx, double* ddy)
{
    for (long k = 0; k < n; ++k) {
        const d4::Dual2 f = d4::dxy_diff_x(dx[k], dy[k]);
        v[k] = f.v;
        ddx[k] = f.x;
        ddy[k] = f.y;
    }
}
// the storage weight and its derivatives at (dx, dy)
void hm_storage_weight(long n, const double* dx, const double* dy, double cutoff, int heaviside, double* v,
                       double* ddx, double* ddy)
{
    for (long k = 0; k < n; ++k) {
        const double d2 = d4::dist2(dx[k], dy[k]);
        const d4::Dual2 w = d4::dstorage_weight(d4::Dual2{d2, 2.0 * dx[k], 2.0 * dy[k]}, cutoff, cutoff * cutoff,
                                                1.0 / (cutoff * cutoff), heaviside);
        v[k] = w.v;
        ddx[k] = w.x;
        ddy[k] = w.y;
    }
}
void hm_inputs(long n, const double* dx, const double* dy, double* e, double* r, double* c)
{
    for (long k = 0; k < n; ++k) d4::field_inputs(dx[k], dy[k], e[k], r[k], c[k]);
}
}
