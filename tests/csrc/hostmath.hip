// hostmath.hip -- test-only: the kernels' per-pair math (2d4_amd/csrc/d4_field.h)
// compiled for the host, so its arithmetic can be checked against the oracle on
// a machine without a GPU.  Not part of lib2d4.
#include "../../2d4_amd/csrc/d4_field.h"

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
void hm_inputs(long n, const double* dx, const double* dy, double* e, double* r, double* c)
{
    for (long k = 0; k < n; ++k) d4::field_inputs(dx[k], dy[k], e[k], r[k], c[k]);
}
}
