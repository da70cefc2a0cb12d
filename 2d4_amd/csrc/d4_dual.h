// d4_dual.h -- the reference's x-derivative of the field, Field::xy_diff_x (Field.cpp:169-334, the
// branch the reference builds: the walls at +-3, +-4, +-5 take cosh in place of sinh), together with
// its own first derivatives in dx and in dy, by forward-mode dual numbers.
//
// The implicit step's matrix holds h b_i b_j xy_diff_x(x_i - x_j, y_i - y_j) for the pairs inside the
// storage radius (simulation.cpp:458-470); the vector-Jacobian product of the step needs how those
// entries move with the positions.  Every expression below follows the reference's formula for its wall
// term by term, so the derivative is that of the function the step uses (the quirky walls included), not
// of the true d sigma / dx.  Value, d/d(dx) and d/d(dy) are carried together; the close-pair Taylor branch
// (dx^2 + dy^2 <= 1e-6) is differentiated as written too.
//
// Not tuned for speed: the adjoint evaluates it once per pair inside the storage radius per step.
#ifndef D4_DUAL_H
#define D4_DUAL_H

#include "d4_field.h"

namespace d4 {

// a value with its derivatives in dx and dy
struct Dual2 {
    double v, x, y;
};

D4_HD Dual2 dconst(double v) { return Dual2{v, 0.0, 0.0}; }
D4_HD Dual2 operator+(Dual2 a, Dual2 b) { return Dual2{a.v + b.v, a.x + b.x, a.y + b.y}; }
D4_HD Dual2 operator-(Dual2 a, Dual2 b) { return Dual2{a.v - b.v, a.x - b.x, a.y - b.y}; }
D4_HD Dual2 operator-(Dual2 a) { return Dual2{-a.v, -a.x, -a.y}; }
D4_HD Dual2 operator*(Dual2 a, Dual2 b) { return Dual2{a.v * b.v, a.x * b.v + a.v * b.x, a.y * b.v + a.v * b.y}; }
D4_HD Dual2 operator*(double s, Dual2 a) { return Dual2{s * a.v, s * a.x, s * a.y}; }
D4_HD Dual2 operator*(Dual2 a, double s) { return s * a; }
D4_HD Dual2 operator+(Dual2 a, double s) { return Dual2{a.v + s, a.x, a.y}; }
D4_HD Dual2 operator+(double s, Dual2 a) { return Dual2{s + a.v, a.x, a.y}; }
D4_HD Dual2 operator-(Dual2 a, double s) { return Dual2{a.v - s, a.x, a.y}; }
D4_HD Dual2 operator-(double s, Dual2 a) { return Dual2{s - a.v, -a.x, -a.y}; }
D4_HD Dual2 operator/(Dual2 a, Dual2 b)
{
    const double q = a.v / b.v;
    return Dual2{q, (a.x - q * b.x) / b.v, (a.y - q * b.y) / b.v};
}
D4_HD Dual2 operator/(Dual2 a, double s) { return Dual2{a.v / s, a.x / s, a.y / s}; }
D4_HD Dual2 dpow(Dual2 a, int k)  // a^k, k >= 1
{
    double p = 1.0;
    for (int i = 0; i < k - 1; ++i) p *= a.v;  // a^(k-1)
    const double dk = k * p;
    return Dual2{p * a.v, dk * a.x, dk * a.y};
}
D4_HD Dual2 dsinh(Dual2 a)
{
    const double c = cosh(a.v);
    return Dual2{sinh(a.v), c * a.x, c * a.y};
}
D4_HD Dual2 dcosh(Dual2 a)
{
    const double s = sinh(a.v);
    return Dual2{cosh(a.v), s * a.x, s * a.y};
}
D4_HD Dual2 dsqrt(Dual2 a)
{
    const double r = sqrt(a.v);
    return Dual2{r, a.x / (2.0 * r), a.y / (2.0 * r)};
}
D4_HD Dual2 dexp(Dual2 a)
{
    const double e = exp(a.v);
    return Dual2{e, e * a.x, e * a.y};
}

// one wall of the field at distance u (Field.cpp:191-196): 2 pi u (C c - 1) / (C - c)^2 pi,
// C = cosh 2 pi u, c = cos 2 pi dy
D4_HD Dual2 dwall(Dual2 u, Dual2 c, Dual2 C)
{
    const Dual2 t = C - c;
    return (kTwoPi * u) * (C * c - 1.0) / (t * t) * kPi;
}

// the x-derivative of one wall (Field.cpp:253-266); S = sinh 2 pi u, or C itself for the walls that take
// cosh in place of sinh (Field.cpp:269-282)
D4_HD Dual2 dwall_dx(Dual2 u, Dual2 c, Dual2 C, Dual2 S)
{
    const Dual2 t = C - c, n = C * c - 1.0;
    const Dual2 t2 = t * t;
    return (n / t2 + (kTwoPi * u) * (S * c / t2 - n / (t2 * t) * 2.0 * S)) * kTwoPiSq;
}

// the central wall's x-derivative (Field.cpp:209-250): the wall form beyond dx^2 + dy^2 = 1e-6, the
// reference's Taylor form inside
D4_HD Dual2 dwall0_dx(Dual2 dx, Dual2 dy, Dual2 c)
{
    const Dual2 X = kTwoPi * dx, Y = kTwoPi * dy;
    if (dx.v * dx.v + dy.v * dy.v > 1e-6) return dwall_dx(dx, c, dcosh(X), dsinh(X));
    const Dual2 X2 = dpow(X, 2), X3 = dpow(X, 3), X4 = dpow(X, 4), X5 = dpow(X, 5), X6 = dpow(X, 6), X7 = dpow(X, 7);
    const Dual2 Y2 = dpow(Y, 2), Y4 = dpow(Y, 4), Y6 = dpow(Y, 6);
    const Dual2 chm1 = X6 / 720.0 + X4 / 24.0 + X2 * 0.5;     // cosh X - 1
    const Dual2 cm1 = -Y6 / 720.0 + Y4 / 24.0 - Y2 * 0.5;     // cos Y - 1
    const Dual2 cmc = (Y6 + X6) / 720.0 + (X4 - Y4) / 24.0 + (X2 + Y2) * 0.5;  // cosh X - cos Y
    // cos Y sinh X to the reference's order
    const Dual2 cs = -(X7 * Y6) / 3628800.0 - (X5 * Y6) / 86400.0 - (X3 * Y6) / 4320.0 - (Y6 * X) / 720.0 +
                     (X7 * Y4) / 120960.0 + (X5 * Y4) / 2880.0 + (X3 * Y4) / 144.0 + (Y4 * X) / 24.0 -
                     (X7 * Y2) / 1080.0 - (X5 * Y2) / 240.0 - (X3 * Y2) / 12.0 - (Y2 * X) * 0.5 + X7 / 5040.0 +
                     X5 / 120.0 + X3 / 6.0 + X;
    const Dual2 num = chm1 * cm1 + cm1 + chm1;
    const Dual2 cmc2 = cmc * cmc;
    return (num / cmc2 + ((cs * dx) / cmc2 - num / (cmc2 * cmc) * dx * dsinh(X) * 2.0) * kTwoPi) * kTwoPiSq;
}

// Field::xy_diff_x(dx, dy) (Field.cpp:314-331) with its derivatives in dx and dy; dx, dy wrapped
D4_HD Dual2 dxy_diff_x(double dxv, double dyv)
{
    const Dual2 dx{dxv, 1.0, 0.0}, dy{dyv, 0.0, 1.0};
    const Dual2 c{cos(kTwoPi * dyv), 0.0, -kTwoPi * sin(kTwoPi * dyv)};
    // cosh and sinh of 2 pi (dx + m): exact for |m| <= 2 (constants.h:63-67, 75-79); for |m| >= 3 the
    // reference takes cosh(2 pi |m|) e^(+-2 pi dx) (constants.h:68-73), whose derivative is +-2 pi times itself
    auto C = [&](double m) { return dcosh(kTwoPi * (dx + m)); };
    auto S = [&](double m) { return dsinh(kTwoPi * (dx + m)); };
    auto Cfar = [&](double m) {
        const double sg = m > 0 ? 1.0 : -1.0;
        const double v = cosh(kTwoPi * (sg * m)) * exp(sg * kTwoPi * dxv);
        return Dual2{v, sg * kTwoPi * v, 0.0};
    };
    auto hi = [&](double m) { const Dual2 Cm = Cfar(m); return dwall_dx(dx + m, c, Cm, Cm); };  // cosh for sinh
    auto lo = [&](double m) { return dwall_dx(dx + m, c, C(m), S(m)); };
    auto wl = [&](double m) { return dwall(dx + m, c, Cfar(m)); };
    // the three wall pairs and the central wall (Field.cpp:285-292)
    const Dual2 g3 = dwall0_dx(dx, dy, c) + lo(1.0) + lo(-1.0) + lo(2.0) + lo(-2.0) + hi(3.0) + hi(-3.0);
    if (dxv < 0)
        return hi(5.0) * (0.0 - dx) + wl(5.0) * (-1.0) + hi(-4.0) * (1.0 + dx) + wl(-4.0) + hi(4.0) + g3;
    return hi(-5.0) * dx + wl(-5.0) + hi(4.0) * (1.0 - dx) + wl(4.0) * (-1.0) + hi(-4.0) + g3;
}

// the storage weight of a pair at squared distance d2 (simulation.cpp:462-470): 1 within the cutoff, the
// Gaussian taper exp(-(r - rc)^2 / rc^2) beyond it (not with the heaviside cutoff), with its derivatives;
// the caller has checked that the pair is inside the storage radius
D4_HD Dual2 dstorage_weight(Dual2 d2, double cutoff, double cut_sq, double one_per_cut_sq, int heaviside)
{
    if (heaviside || d2.v <= cut_sq) return dconst(1.0);
    const Dual2 e = dsqrt(d2) - cutoff;
    return dexp(-(e * e) * one_per_cut_sq);
}

}  // namespace d4

#endif
