// d4_field.h -- the periodic sigma_xy image sum and its x-derivative, per pair,
// in the form the gfx950 kernels evaluate it.
//
// Semantics are those of the reference's production branch
// (USE_IEEE_HYPERBOLIC == 0, include/constants.h:47; src/Fields/Field.cpp:169-347):
//   sigma(dx,dy) = 2 pi^2 [ sum_{k=-3..3} f(dx+k) + f(dx-4s) + (1-|dx|) f(dx+4s) + |dx| f(dx-5s) ]
//   f(u)         = u (C c - 1) / (C - c)^2,   C = cosh 2 pi u,  c = cos 2 pi dy,  s = sign(dx) (+1 at 0)
// with cosh/sinh of the shifted arguments built from E = exp(2 pi dx) through the
// identities of constants.h:63-79 (|k| <= 2) or C = cosh(2 pi k) E^{+-1} (|k| >= 3),
// and the x-derivative of Field.cpp:314-331 including its f_dx_h substitution
// (cosh used as sinh for |k| >= 3, Field.cpp:269-283).  EXACT = true uses the true
// sinh sign there instead (the derivative of sigma, for the adjoint).
// The central wall switches to the reference's 6th-order series when
// dx^2 + dy^2 <= 1e-6 (Field.cpp:176-188, 213-250).
//
// Arithmetic layout (the part that is this build's own): one reciprocal per
// image wall, shared by the wall's sigma and d sigma terms,
//   r = 1/D,  t = N r,  f = u t r,  f' = r^2 (N + 2 pi u S (c - 2 t)),
// and one reciprocal of E shared by all hyperbolics; every division of the
// reference becomes a multiply by one of these eleven reciprocals.
// Where the reference's rounding sequence decides the result -- 1/E, the
// hyperbolic identities, and C c - 1 and C - c of the central wall, which
// cancel for near pairs -- the same sequence is kept (unfused, correctly
// rounded 1/E), so those cancellations reproduce the reference's values.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define D4_HD __host__ __device__ __forceinline__

namespace d4 {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kTwoPiSq = 2.0 * kPi * kPi;
constexpr double kPi2Ln2Inv = 9.0647202836543876192553658914333336203437229354476;  // 2 pi / ln 2

// cosh(2 pi k), sinh(2 pi k) (numerical values as in constants.h:54-61); halves for the identities
constexpr double kCh1 = 267.746761483748222245931879901;
constexpr double kCh2 = 143375.656570070321051450310133;
constexpr double kCh3 = 7.67764676977233502175191858009e7;
constexpr double kCh4 = 4.11131577927974976374895337541e10;
constexpr double kCh5 = 2.20157529303160145057002722946e13;
constexpr double kSh1 = 267.744894041016514257117449688;
constexpr double kSh2 = 143375.656566582978695241314640;
constexpr double kCh1h = 0.5 * kCh1, kSh1h = 0.5 * kSh1, kCh2h = 0.5 * kCh2, kSh2h = 0.5 * kSh2;

// Reciprocal: hardware v_rcp_f64 seed refined by two Newton steps (each
// doubles the correct bits); within ~1 ulp for the positive normal range used here.
D4_HD double rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
#else
    double r = (double)(1.0f / (float)x);
#endif
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// sim_utils.cpp:6-13 semantics: the unique value in [-0.5, 0.5) congruent to v
// mod 1.  v - rint(v) is exact (a multiple of ulp(v) no larger than |v|) and
// lies in [-0.5, 0.5]; only +0.5 needs the final -1, as the reference's
// `while (n >= 0.5) n -= 1` gives.
D4_HD double wrap(double v)
{
    double r = v - rint(v);
    return r >= 0.5 ? r - 1.0 : r;
}

// dx^2 + dy^2 rounded as the reference rounds it (two products, one sum;
// simulation.cpp:302, 452), so cutoff membership and tolerances match exactly.
D4_HD double dist2(double dx, double dy)
{
#pragma clang fp contract(off)
    return dx * dx + dy * dy;
}

// One image wall from C = cosh, S = "sinh", c = cos(2 pi dy): adds w*f to sig
// and w*f' to dsig (both before the common factor 2 pi^2); returns f.
template <bool NEED_D>
D4_HD double wall(double u, double C, double S, double c, double w, double& sig, double& dsig)
{
    double D = C - c;
    double r = rcp(D);
    double N = fma(C, c, -1.0);
    double t = N * r;
    double f = u * t * r;
    sig = fma(w, f, sig);
    if (NEED_D) {
        double q = fma(-2.0, t, c);
        double g = fma(kTwoPi * u, S * q, N);
        dsig = fma(w, g * r * r, dsig);
    }
    return f;
}

// Central wall inside the series radius (Field.cpp:182-187, 224-249): the
// same wall algebra with N, D, S*c and S replaced by their expansions.
template <bool NEED_D>
D4_HD void wall0_series(double dx, double dy, double& sig, double& dsig)
{
    double a = kTwoPi * dx, b = kTwoPi * dy;
    double a2 = a * a, a4 = a2 * a2, a6 = a4 * a2;
    double b2 = b * b, b4 = b2 * b2, b6 = b4 * b2;
    double chm1 = a6 / 720 + a4 / 24 + a2 * 0.5;
    double cm1 = -b6 / 720 + b4 / 24 - b2 * 0.5;
    double D = (b6 + a6) / 720 + (a4 - b4) / 24 + (a2 + b2) * 0.5;
    double N = chm1 * cm1 + cm1 + chm1;
    double r = rcp(D);
    double t = N * r;
    sig = fma(dx, t * r, sig);
    if (NEED_D) {
        double a3 = a2 * a, a5 = a4 * a, a7 = a6 * a;
        double cs = -(a7 * b6) / 3628800 - (a5 * b6) / 86400 - (a3 * b6) / 4320 - (b6 * a) / 720
                  + (a7 * b4) / 120960 + (a5 * b4) / 2880 + (a3 * b4) / 144 + (b4 * a) / 24
                  - (a7 * b2) / 1080 - (a5 * b2) / 240 - (a3 * b2) / 12 - (b2 * a) * 0.5
                  + a7 / 5040 + a5 / 120 + a3 / 6 + a;
        double sh = a + a3 / 6 + a5 / 120 + a7 / 5040;  // sinh(a), |a| <= 2 pi 1e-3
        double g = fma(kTwoPi * dx, fma(-2.0 * t, sh, cs), N);
        dsig += g * r * r;
    }
}

// Hyperbolic values of 2 pi (dx + k) from E (constants.h:63-79), evaluated
// unfused in the reference's order.
struct Hyp {
    double R, P, M;
    double Cp1, Cm1, Sp1, Sm1, Cp2, Cm2, Sp2, Sm2;
};

D4_HD Hyp hyperbolics(double E)
{
#pragma clang fp contract(off)
    Hyp h;
    h.R = 1.0 / E;
    h.P = E + h.R;
    h.M = E - h.R;
    double a = h.P * kCh1h, b = h.M * kSh1h, as = h.M * kCh1h, bs = h.P * kSh1h;
    h.Cp1 = a + b;
    h.Cm1 = a - b;
    h.Sp1 = as + bs;
    h.Sm1 = as - bs;
    a = h.P * kCh2h;
    b = h.M * kSh2h;
    as = h.M * kCh2h;
    bs = h.P * kSh2h;
    h.Cp2 = a + b;
    h.Cm2 = a - b;
    h.Sp2 = as + bs;
    h.Sm2 = as - bs;
    return h;
}

// Central wall outside the series radius: D and N formed as the reference
// forms them (Field.cpp:178-179: C - c, then C*c rounded before - 1).
template <bool NEED_D>
D4_HD void wall0(double dx, double C, double S, double c, double& sig, double& dsig)
{
    double D, N;
    {
#pragma clang fp contract(off)
        D = C - c;
        N = C * c - 1.0;
    }
    double r = rcp(D);
    double t = N * r;
    sig = fma(dx, t * r, sig);
    if (NEED_D) {
        double q = fma(-2.0, t, c);
        double g = fma(kTwoPi * dx, S * q, N);
        dsig = fma(g, r * r, dsig);
    }
}

// sigma (and d sigma/dx if NEED_D) of one pair from (dx, dy) already wrapped
// and the two transcendental inputs E = 2^(kPi2Ln2Inv dx), c = cos(2 pi dy).
template <bool NEED_D, bool EXACT>
D4_HD void pair_field(double dx, double dy, double E, double c, double& sig_out, double& dsig_out)
{
    const Hyp h = hyperbolics(E);
    const double R = h.R;
    const bool pos = dx >= 0.0;
    const double sg = pos ? 1.0 : -1.0;
    const double ad = pos ? dx : -dx;

    // |k| >= 3 walls: C = cosh(2 pi k) E^{+-1}; "sinh" = C (reference) or +-C (exact)
    const double Cp3 = kCh3 * E, Cm3 = kCh3 * R;
    const double Cp4 = kCh4 * E, Cm4 = kCh4 * R;
    const double Cp5 = kCh5 * E, Cm5 = kCh5 * R;
    const double CA = pos ? Cm4 : Cp4;  // dx - 4s, weight 1
    const double CB = pos ? Cp4 : Cm4;  // dx + 4s, weight 1 - |dx|
    const double CC = pos ? Cm5 : Cp5;  // dx - 5s, weight |dx|
    const double neg = EXACT ? -1.0 : 1.0;
    const double SA = EXACT ? -sg * CA : CA;
    const double SB = EXACT ? sg * CB : CB;
    const double SC = EXACT ? -sg * CC : CC;

    double sig = 0.0, dsig = 0.0;
    wall<NEED_D>(dx - 4.0 * sg, CA, SA, c, 1.0, sig, dsig);
    const double fB = wall<NEED_D>(dx + 4.0 * sg, CB, SB, c, 1.0 - ad, sig, dsig);
    const double fC = wall<NEED_D>(dx - 5.0 * sg, CC, SC, c, ad, sig, dsig);
    wall<NEED_D>(dx + 3.0, Cp3, Cp3, c, 1.0, sig, dsig);
    wall<NEED_D>(dx - 3.0, Cm3, neg * Cm3, c, 1.0, sig, dsig);
    wall<NEED_D>(dx + 2.0, h.Cp2, h.Sp2, c, 1.0, sig, dsig);
    wall<NEED_D>(dx - 2.0, h.Cm2, h.Sm2, c, 1.0, sig, dsig);
    wall<NEED_D>(dx + 1.0, h.Cp1, h.Sp1, c, 1.0, sig, dsig);
    wall<NEED_D>(dx - 1.0, h.Cm1, h.Sm1, c, 1.0, sig, dsig);
    // central wall
    if (dist2(dx, dy) > 1e-6)
        wall0<NEED_D>(dx, 0.5 * h.P, 0.5 * h.M, c, sig, dsig);
    else
        wall0_series<NEED_D>(dx, dy, sig, dsig);

    sig_out = kTwoPiSq * sig;
    if (NEED_D) dsig_out = kTwoPiSq * fma(sg, fC - fB, dsig);
}

// The reference's transcendental inputs of a pair (simulation.cpp:456-457,
// Field.cpp:335-336).
D4_HD void field_inputs(double dx, double dy, double& E, double& c)
{
    E = exp2(kPi2Ln2Inv * dx);
    c = cos(kTwoPi * dy);
}

// d sigma / d dy (no reference counterpart; adjoint only): every wall's
// closed form d/dc[(C c - 1)/(C - c)^2] = (C^2 + C c - 2)/(C - c)^3 times
// dc/ddy = -2 pi sin(2 pi dy).
D4_HD double wall_dy(double u, double C, double c, double w, double acc)
{
    double D = C - c;
    double r = rcp(D);
    double num = fma(C, C + c, -2.0);
    return fma(w * u * num, r * r * r, acc);
}

D4_HD double pair_field_dy(double dx, double E, double c, double s2piy)
{
    const Hyp h = hyperbolics(E);
    const double R = h.R;
    const bool pos = dx >= 0.0;
    const double sg = pos ? 1.0 : -1.0;
    const double ad = pos ? dx : -dx;
    const double Cp4 = kCh4 * E, Cm4 = kCh4 * R, Cp5 = kCh5 * E, Cm5 = kCh5 * R;
    double acc = 0.0;
    acc = wall_dy(dx - 4.0 * sg, pos ? Cm4 : Cp4, c, 1.0, acc);
    acc = wall_dy(dx + 4.0 * sg, pos ? Cp4 : Cm4, c, 1.0 - ad, acc);
    acc = wall_dy(dx - 5.0 * sg, pos ? Cm5 : Cp5, c, ad, acc);
    acc = wall_dy(dx + 3.0, kCh3 * E, c, 1.0, acc);
    acc = wall_dy(dx - 3.0, kCh3 * R, c, 1.0, acc);
    acc = wall_dy(dx + 2.0, h.Cp2, c, 1.0, acc);
    acc = wall_dy(dx - 2.0, h.Cm2, c, 1.0, acc);
    acc = wall_dy(dx + 1.0, h.Cp1, c, 1.0, acc);
    acc = wall_dy(dx - 1.0, h.Cm1, c, 1.0, acc);
    acc = wall_dy(dx, 0.5 * h.P, c, 1.0, acc);
    return -kTwoPiSq * kTwoPi * s2piy * acc;
}

// sim_utils.cpp:36-54
D4_HD double weight(double s, char kind)
{
    if (s > 0) {
        if (kind == 'c') { double q = s / (s + 1); return q * q; }
        if (kind == 'p') return s / (s + 1);
        if (kind == 'm') { const double cm = 0.412073; return 1 / (1 + 1.5 / (s + cm * s * s)); }
    }
    return 0;
}

// sim_utils.cpp:57-75
D4_HD double weight_inv(double w, char kind)
{
    if (w > 0) {
        if (kind == 'c') return sqrt(w) / (1 - sqrt(w));
        if (kind == 'p') return w / (1 - w);
        if (kind == 'm') {
            const double cm = 0.412073;
            return (1 - w - sqrt(1 - 2 * w + 6 * cm * w + w * w - 6 * cm * w * w)) / (2 * (-cm + cm * w));
        }
    }
    return 0;
}

}  // namespace d4
