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
// Arithmetic layout (the part that is this library's own): one reciprocal per
// near image wall (|k| <= 2), shared by the wall's sigma and d sigma terms,
//   r = 1/D,  t = N r,  f = u t r,  f' = r^2 (N + 2 pi u S (c - 2 t)),
// the five of them from one batch inversion, one reciprocal of E shared by all
// hyperbolics, and the far walls (|k| >= 3) as series in 1/C, a constant
// times E or 1/E; no division of the reference remains.
// Where the reference's rounding sequence decides the result -- 1/E, the
// hyperbolic identities, and C c - 1 and C - c of the central wall, which
// cancel for near pairs -- the same sequence is kept (unfused, correctly
// rounded 1/E), so those cancellations reproduce the reference's values.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define D4_HD __host__ __device__ __forceinline__

// a * b + c, fused, for a constant c (a polynomial coefficient).  On the
// device as an explicit three-operand v_fma_f64 reading c from a scalar
// register: the compiler's accumulate form (v_fmac_f64) would first copy the
// constant into the destination, one extra VALU slot per step of a
// polynomial.  Arithmetic only.
D4_HD double fma3(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    asm volatile("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
#else
    return fma(a, b, c);
#endif
}

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
// doubles the correct bits).  The last step, y + y(1 - x y) with one FMA
// rounding from a faithful y, returns the correctly rounded 1/x (Markstein's
// theorem) for the positive normal range used here.
D4_HD double rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
#else  // host build (tests): a float-accurate seed over the whole double range
    int ex = 0;
    const double m = frexp(x, &ex);
    double r = ldexp((double)(1.0f / (float)m), -ex);
#endif
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// sim_utils.cpp:6-13 semantics: the unique value in [-0.5, 0.5) congruent to v
// mod 1.  v - rint(v) is exact (a multiple of ulp(v) no larger than |v|) and
// lies in [-0.5, 0.5]; only +0.5 needs the final -1, as the reference's
// `while (n >= 0.5) n -= 1` gives, and -(+0.5) is that value.
D4_HD double wrap(double v)
{
    double r = v - rint(v);
    return r >= 0.5 ? -r : r;  // r = +0.5 exactly: -0.5 = r - 1
}

// dx^2 + dy^2 rounded as the reference rounds it (two products, one sum;
// simulation.cpp:302, 452), so cutoff membership and tolerances match exactly.
D4_HD double dist2(double dx, double dy)
{
#pragma clang fp contract(off)
    return dx * dx + dy * dy;
}

// Central wall inside the series radius (Field.cpp:182-187, 224-249): the
// same wall algebra with N, D, S*c and S replaced by their expansions.
template <bool NEED_D>
D4_HD void wall0_series(double dx, double dy, double& sig, double& dsig)
{
    // 1/n! and the other series denominators as products with their reciprocals
    constexpr double i24 = 1.0 / 24, i720 = 1.0 / 720;
    double a = kTwoPi * dx, b = kTwoPi * dy;
    double a2 = a * a, a4 = a2 * a2, a6 = a4 * a2;
    double b2 = b * b, b4 = b2 * b2, b6 = b4 * b2;
    double chm1 = a6 * i720 + a4 * i24 + a2 * 0.5;
    double cm1 = -b6 * i720 + b4 * i24 - b2 * 0.5;
    double D = (b6 + a6) * i720 + (a4 - b4) * i24 + (a2 + b2) * 0.5;
    double N = chm1 * cm1 + cm1 + chm1;
    double r = rcp(D);
    double t = N * r;
    sig = fma(dx, t * r, sig);
    if (NEED_D) {
        double a3 = a2 * a, a5 = a4 * a, a7 = a6 * a;
        double cs = -(a7 * b6) * (1.0 / 3628800) - (a5 * b6) * (1.0 / 86400) - (a3 * b6) * (1.0 / 4320)
                  - (b6 * a) * i720 + (a7 * b4) * (1.0 / 120960) + (a5 * b4) * (1.0 / 2880)
                  + (a3 * b4) * (1.0 / 144) + (b4 * a) * i24 - (a7 * b2) * (1.0 / 1080)
                  - (a5 * b2) * (1.0 / 240) - (a3 * b2) * (1.0 / 12) - (b2 * a) * 0.5
                  + a7 * (1.0 / 5040) + a5 * (1.0 / 120) + a3 * (1.0 / 6) + a;
        double sh = a + a3 * (1.0 / 6) + a5 * (1.0 / 120) + a7 * (1.0 / 5040);  // sinh(a), |a| <= 2 pi 1e-3
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

D4_HD Hyp hyperbolics(double E, double R)
{
#pragma clang fp contract(off)
    Hyp h;
    h.R = R;
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

// Five reciprocals from one: Montgomery's batch inversion over a product tree
// (4 products, one rcp(), 8 products back).  Each result is within a few ulp
// of 1/x.  The products stay far inside binary64's range here (the five wall
// denominators lie between ~2e-5 and ~3.3e6, their product below ~1e22).
D4_HD void batch_rcp5(const double* x, double* r)
{
    const double p01 = x[0] * x[1], p23 = x[2] * x[3];
    const double p03 = p01 * p23;
    const double iall = rcp(p03 * x[4]);
    const double i03 = iall * x[4];
    r[4] = iall * p03;
    const double i01 = i03 * p23, i23 = i03 * p01;
    r[0] = i01 * x[1];
    r[1] = i01 * x[0];
    r[2] = i23 * x[3];
    r[3] = i23 * x[2];
}

// sigma (and d sigma/dx if NEED_D) of one pair from (dx, dy) already wrapped
// and the transcendental inputs E = 2^(kPi2Ln2Inv dx), R = 1/E, c = cos(2 pi dy).
// Walls (offset u, weight w): dx +- 2, dx +- 1 and dx (the central wall, or the
// series when dx^2 + dy^2 <= 1e-6) through one batch inversion; dx +- 3, dx - 4s
// (1), dx + 4s (1 - |dx|), dx - 5s (|dx|) from series in rho = 1/C.
//
// For |k| >= 3, C = cosh(2 pi k) E^(+-1) >= kCh3 e^-pi ~ 3.3e6, so rho <= 3e-7 is
// a constant times R or E, and the wall terms are the series
//   f  = u rho (c - rho) / (1 - c rho)^2
//      = u rho [c + rho (2c^2 - 1) + O(rho^2)]
//   f' = rho [(c - rho) + 2 pi u s (c - 2t)] / (1 - c rho)^2
//      = rho [c (1 - 2 pi u s) + rho (2c^2 - 1)(1 - 4 pi u s) + O(30 rho^2)]
// (t = (c - rho)/(1 - c rho), s the sign of the "sinh", +1 for the reference's
// cosh-as-sinh).  At +-3 two terms leave ~1e-19 out; at +-4 and +-5
// (rho <= 5.6e-10) the first term alone leaves ~1e-18 out -- both four orders
// under the rounding of the near walls, and no reciprocal is needed.
template <bool NEED_D, bool EXACT>
D4_HD void pair_field(double dx, double dy, double E, double R, double c, double& sig_out, double& dsig_out)
{
    constexpr bool dd = NEED_D;
    const Hyp h = hyperbolics(E, R);
    const bool pos = dx >= 0.0;
    const double sg = pos ? 1.0 : -1.0;
    const double ad = fabs(dx);
    const bool series = !(dist2(dx, dy) > 1e-6);

    // the five reciprocal walls, in batch order: +2, -2, +1, -1, 0
    double C[5], D[5], N[5], r[5];
    C[0] = h.Cp2;
    C[1] = h.Cm2;
    C[2] = h.Cp1;
    C[3] = h.Cm1;
    C[4] = 0.5 * h.P;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        D[k] = C[k] - c;
        N[k] = fma(C[k], c, -1.0);
    }
    {
        // central wall: C - c and C c - 1 rounded as the reference rounds them
        // (Field.cpp:178-179), since for near pairs both cancel
#pragma clang fp contract(off)
        D[4] = C[4] - c;
        N[4] = C[4] * c - 1.0;
    }
    if (series) D[4] = 1.0;  // the series branch brings its own denominator
    batch_rcp5(D, r);

    double sig = 0.0, dsig = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k == 4 && series) break;
        const double uk = k == 0 ? dx + 2.0 : k == 1 ? dx - 2.0 : k == 2 ? dx + 1.0 : k == 3 ? dx - 1.0 : dx;
        const double r2 = r[k] * r[k];
        sig = fma(uk, N[k] * r2, sig);
        if (dd) {
            // sinh of the wall from the identities
            const double Sk = k == 0 ? h.Sp2 : k == 1 ? h.Sm2 : k == 2 ? h.Sp1 : k == 3 ? h.Sm1 : 0.5 * h.M;
            const double q = fma(-2.0, N[k] * r[k], c);
            const double g = fma(kTwoPi * uk, Sk * q, N[k]);
            dsig = fma(g, r2, dsig);
        }
    }
    if (series) {
        if (dd) wall0_series<true>(dx, dy, sig, dsig);
        else wall0_series<false>(dx, dy, sig, dsig);
    }

    constexpr double kIch3 = 1.0 / kCh3, kIch4 = 1.0 / kCh4, kIch5 = 1.0 / kCh5;
    // E f, R f: the factor of rho at the walls -4s, -5s (and -3 for dx >= 0) and at +4s (and +3)
    const double Ef = pos ? E : R, Rf = pos ? R : E;
    if constexpr (EXACT) {
        // the walls one by one (the true sinh signs of the derivative)
        // the walls at +-3: rho = R / cosh(6 pi) at +3, E / cosh(6 pi) at -3; q2 = 2c^2 - 1
        const double q2 = fma(2.0 * c, c, -1.0);
        const double rp3 = kIch3 * R, rm3 = kIch3 * E;
        const double up3 = dx + 3.0, um3 = dx - 3.0;
        const double ap3 = fma(rp3, q2, c), am3 = fma(rm3, q2, c);
        sig = fma(up3 * rp3, ap3, sig);
        sig = fma(um3 * rm3, am3, sig);
        // the walls at -4s (weight 1), +4s (1 - |dx|) and -5s (|dx|): rho c from E c or R c
        const double ec = Ef * c, rc = Rf * c;
        const double u2 = dx - 4.0 * sg, u3 = dx + 4.0 * sg, u4 = dx - 5.0 * sg;
        const double rc2 = kIch4 * ec, rc3 = kIch4 * rc, rc4 = kIch5 * ec;
        const double f3 = u3 * rc3, f4 = u4 * rc4;
        sig = fma(u2, rc2, sig);
        sig = fma(1.0 - ad, f3, sig);
        sig = fma(ad, f4, sig);
        sig_out = kTwoPiSq * sig;
        if (dd) {
            const double bp3 = fma(c, fma(-kTwoPi, up3, 1.0), rp3 * q2 * fma(-2.0 * kTwoPi, up3, 1.0));
            const double bm3 = fma(c, fma(kTwoPi, um3, 1.0), rm3 * q2 * fma(2.0 * kTwoPi, um3, 1.0));
            dsig = fma(rp3, bp3, dsig);
            dsig = fma(rm3, bm3, dsig);
            dsig = fma(rc2, fma(kTwoPi * sg, u2, 1.0), dsig);
            dsig = fma((1.0 - ad) * rc3, fma(-kTwoPi * sg, u3, 1.0), dsig);
            dsig = fma(ad * rc4, fma(kTwoPi * sg, u4, 1.0), dsig);
            dsig_out = kTwoPiSq * fma(sg, f4 - f3, dsig);
        }
    } else {
        // The five far walls gathered by their factor of rho (first order) and rho^2 (the +-3 walls'
        // second order); with a = |dx| and s the sign:
        //   sigma part  s [c (Ef A1(a) + Rf A2(a)) + q2/ch3^2 (Ef^2 (a - 3) + Rf^2 (a + 3))],
        //   A1(a) = (a - 3)/ch3 + (a - 4)/ch4 + a (a - 5)/ch5     (-3 for s = +1, -4s weight 1, -5s weight a),
        //   A2(a) = (a + 3)/ch3 + (1 - a)(a + 4)/ch4              (+3 for s = +1, +4s weight 1 - a);
        //   d/dx part (the reference's cosh for sinh, and the derivatives of the weights 1 - a and a)
        //   c (Ef (P1(a) - 2 pi s A1(a)) + Rf (P2(a) - 2 pi s A2(a)))
        //     + q2/ch3^2 (Ef^2 (1 - 4 pi s (a - 3)) + Rf^2 (1 - 4 pi s (a + 3))),
        //   P1(a) = 1/ch3 + 1/ch4 - 5/ch5 + 2a/ch5,  P2(a) = 1/ch3 - 3/ch4 - 2a/ch4.
        // These terms are below 1e-6 of the near walls', so their rounding stays far under the bound.
        constexpr double a12 = kIch5, a11 = kIch3 + kIch4 - 5.0 * kIch5, a10 = -(3.0 * kIch3 + 4.0 * kIch4);
        constexpr double a22 = -kIch4, a21 = kIch3 - 3.0 * kIch4, a20 = 3.0 * kIch3 + 4.0 * kIch4;
        constexpr double k3sq = kIch3 * kIch3;
        // (fma3 keeps a constant addend in a scalar register instead of copying it into the result first)
        const double A1 = fma3(fma3(ad, a12, a11), ad, a10);
        const double A2 = fma3(fma3(ad, a22, a21), ad, a20);
        const double Ef2 = Ef * Ef, Rf2 = Rf * Rf;
        const double am3 = ad - 3.0, ap3 = ad + 3.0;
        const double q2k = fma3((2.0 * k3sq) * c, c, -k3sq);  // (2 c^2 - 1) / ch3^2
        const double first = fma(Rf, A2, Ef * A1);
        const double second = fma(Rf2, ap3, Ef2 * am3);
        sig = fma(sg, fma(q2k, second, c * first), sig);
        sig_out = kTwoPiSq * sig;
        if (dd) {
            const double s2pi = -kTwoPi * sg, s4pi = 2.0 * s2pi;
            const double P1 = fma3(ad, 2.0 * kIch5, kIch3 + kIch4 - 5.0 * kIch5);
            const double P2 = fma3(ad, -2.0 * kIch4, kIch3 - 3.0 * kIch4);
            const double first_d = fma(Rf, fma(s2pi, A2, P2), Ef * fma(s2pi, A1, P1));
            const double second_d = fma(Rf2, fma(s4pi, ap3, 1.0), Ef2 * fma(s4pi, am3, 1.0));
            dsig_out = kTwoPiSq * (dsig + fma(q2k, second_d, c * first_d));
        }
    }
}

// The reference's transcendental inputs of a pair (simulation.cpp:456-457,
// Field.cpp:335-336): E = exp2(PI2LN2INV * dx), c = cos(2 pi dy), here with
// R = 1/E as well.  Own implementations, each within ~1 ulp:
//   exp2_pair: y = k + f, |f| <= 1/2, 2^f = 1 + f g(f) with g a degree-10
//     polynomial near the minimax one for (2^f - 1)/f on [-1/2, 1/2] (Chebyshev
//     fit, coefficients rounded to double: 2.0e-17 relative), scaled by 2^k;
//     R = 1/E by rcp();
//   cos2pi_reduced: half-angle form, sin r = r + r^3 h(r^2) on |r| <= pi/4 with h
//     of degree 5 fitted the same way (2.1e-17 relative).
namespace tc {
// g(f) = (2^f - 1)/f, highest degree first
constexpr double g10 = 0x1.e9d3fe3952179p-32, g9 = 0x1.e6063f7217bc6p-28, g8 = 0x1.b524fae627834p-24,
                 g7 = 0x1.62bfd47773353p-20, g6 = 0x1.ffcbfc670dcd4p-17, g5 = 0x1.430913096fd9fp-13,
                 g4 = 0x1.5d87fe78a5276p-10, g3 = 0x1.3b2ab6fba1ddap-7, g2 = 0x1.c6b08d704a0c2p-5,
                 g1 = 0x1.ebfbdff82c598p-3, g0 = 0x1.62e42fefa39efp-1;
// h(u) = (sin sqrt(u) - sqrt(u)) / u^(3/2), highest degree first
constexpr double h5 = 0x1.5e0b19f8b1451p-33, h4 = -0x1.ae600b02b6262p-26, h3 = 0x1.71de37968a100p-19,
                 h2 = -0x1.a01a019e83aaep-13, h1 = 0x1.1111111110bb2p-7, h0 = -0x1.5555555555555p-3;
constexpr double pio2_1 = 0x1.921fb54442d18p+0, pio2_2 = 0x1.1a62633145c07p-54, two_over_pi = 0x1.45f306dc9c883p-1;
}  // namespace tc

D4_HD void exp2_pair(double y, double& E, double& R)
{
    const double k = rint(y);
    const double f = y - k;
    double q = fma3(f, tc::g10, tc::g9);
    q = fma3(f, q, tc::g8);
    q = fma3(f, q, tc::g7);
    q = fma3(f, q, tc::g6);
    q = fma3(f, q, tc::g5);
    q = fma3(f, q, tc::g4);
    q = fma3(f, q, tc::g3);
    q = fma3(f, q, tc::g2);
    q = fma3(f, q, tc::g1);
    q = fma3(f, q, tc::g0);
    const double p = fma3(f, q, 1.0);
    E = ldexp(p, (int)k);
    // 1/E correctly rounded, as the reference's identities take it (constants.h:63-79):
    // the last Newton step of rcp() runs from a faithful estimate, which rounds correctly
    R = rcp(E);
}

// cos(2 pi dy) for |dy| <= 1/2 from one sine polynomial: with theta = pi dy,
// cos 2 theta = 1 - 2 sin^2 theta for |theta| <= pi/4 and
// -(1 - 2 sin^2 phi), phi = pi/2 - |theta|, above; |rho| <= pi/4 either way, and
// 1 - 2 s^2 keeps c accurate to rounding where c -> 1 (near pairs).
D4_HD double cos2pi_reduced(double dy)
{
    const double th = kPi * dy;  // = fl(2 pi dy) / 2 exactly
    const double at = fabs(th);
    const bool hi = at > 0.5 * tc::pio2_1;
    const double rho = hi ? (tc::pio2_1 - at) + tc::pio2_2 : th;
    const double r2 = rho * rho;
    double sp = fma3(r2, tc::h5, tc::h4);
    sp = fma3(r2, sp, tc::h3);
    sp = fma3(r2, sp, tc::h2);
    sp = fma3(r2, sp, tc::h1);
    sp = fma3(r2, sp, tc::h0);
    const double sn = fma(rho * r2, sp, rho);
    const double v = fma(-2.0 * sn, sn, 1.0);
    return hi ? -v : v;
}

D4_HD void field_inputs(double dx, double dy, double& E, double& R, double& c)
{
    exp2_pair(kPi2Ln2Inv * dx, E, R);
    c = cos2pi_reduced(dy);
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

D4_HD double pair_field_dy(double dx, double E, double R, double c, double s2piy)
{
    const Hyp h = hyperbolics(E, R);
    const bool pos = dx >= 0.0;
    const double sg = pos ? 1.0 : -1.0;
    const double ad = fabs(dx);
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

// sim_utils.cpp:36-54 (unfused, as the reference rounds it)
D4_HD double weight(double s, char kind)
{
#pragma clang fp contract(off)
    if (s > 0) {
        if (kind == 'c') { double q = s / (s + 1); return q * q; }
        if (kind == 'p') return s / (s + 1);
        if (kind == 'm') { const double cm = 0.412073; return 1 / (1 + 1.5 / (s + cm * s * s)); }
    }
    return 0;
}

// sim_utils.cpp:57-75 (unfused)
D4_HD double weight_inv(double w, char kind)
{
#pragma clang fp contract(off)
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

// Bilinear interpolation in an n x n table of node values over the periodic
// cell [-1/2, 1/2)^2 (node (kx, ky) at (-1/2 + kx/n, -1/2 + ky/n), row ky,
// column kx; n a power of two).  dx, dy are wrapped differences; the same
// operations in the same order on host and device, without contraction, so a
// host build of this function gives bit-identical values.
D4_HD double table_bilinear(const double* tab, int n, double dx, double dy)
{
#pragma clang fp contract(off)
    const double u = (dx + 0.5) * n;
    const double w = (dy + 0.5) * n;
    const double fu = floor(u), fw = floor(w);
    const double ax = u - fu, ay = w - fw;
    const int i0 = ((int)fu) & (n - 1), j0 = ((int)fw) & (n - 1);
    const int i1 = (i0 + 1) & (n - 1), j1 = (j0 + 1) & (n - 1);
    const double t00 = tab[(size_t)j0 * n + i0], t10 = tab[(size_t)j0 * n + i1];
    const double t01 = tab[(size_t)j1 * n + i0], t11 = tab[(size_t)j1 * n + i1];
    return (1.0 - ax) * ((1.0 - ay) * t00 + ay * t01) + ax * ((1.0 - ay) * t10 + ay * t11);
}

// The same interpolation from cell records: record (ky * n + kx) holds the cell's corner values
// {t(kx, ky), t(kx+1, ky), t(kx, ky+1), t(kx+1, ky+1)} (wrapped), so one pair reads one 32-byte record
// (two loads in one cache line) instead of four scattered nodes; the arithmetic is table_bilinear's, so
// the value is the same bits.  The cell of (dx, dy) and its weights are returned for a second quantity
// stored in the same record (t = rec + 4).
D4_HD double bilinear_cell(const double* t, double ax, double ay)
{
#pragma clang fp contract(off)
    return (1.0 - ax) * ((1.0 - ay) * t[0] + ay * t[2]) + ax * ((1.0 - ay) * t[1] + ay * t[3]);
}

D4_HD const double* table_cell(const double* rec, int stride, int n, double dx, double dy, double& ax, double& ay)
{
#pragma clang fp contract(off)
    const double u = (dx + 0.5) * n;
    const double w = (dy + 0.5) * n;
    const double fu = floor(u), fw = floor(w);
    ax = u - fu;
    ay = w - fw;
    const int i0 = ((int)fu) & (n - 1), j0 = ((int)fw) & (n - 1);
    return rec + ((size_t)j0 * n + i0) * stride;
}

}  // namespace d4
