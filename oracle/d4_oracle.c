/*
 * d4_oracle.c -- CPU restatement of 2D4's periodic pair-interaction path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product never
 * does.  It restates, in plain C, the production branch of the reference
 * (USE_IEEE_HYPERBOLIC == 0, include/constants.h:47) and the pair loops of
 * src/simulation.cpp, keeping the reference's operation order so that results
 * agree with the reference's own compiled Field.cpp to rounding.
 *
 * Pinning (what checks each function, see tests/):
 *   orc_xy4 / orc_dxy4 / orc_xy / orc_dxy   golden vectors from the reference's
 *       Field.cpp compiled under oracle/_ref (both -O2 and the reference flags),
 *       an independent many-image lattice sum, odd/even symmetries, continuity.
 *   orc_dxy4(exact=1), orc_dyxy4            central finite differences of orc_xy4.
 *   orc_normalize / orc_weight / orc_weight_inv   reference sim_utils.cpp under
 *       oracle/_ref on edge tables, plus closed-form round trips.
 *   orc_speed_pass / orc_jacobian_pass / orc_jacobian_dense
 *       brute force at small n over the reference Field (oracle/_ref), exact
 *       invariants (antisymmetric pair sums, V symmetric, row sums, A*1 = 1);
 *       the weights d_i equal the reference's own weight (oracle/_ref) of the
 *       unscaled diagonal s_i, bit for bit.
 *   orc_jacobian_from_prev                  V and s are linear in h, so the pass
 *       at h rescaled by 1/2 must equal the pass at h/2 entry for entry (all
 *       weight functions, both cutoff modes).
 *   orc_read_dconf / orc_cutoff             the reference SimulationData under
 *       oracle/_ref on init_config_gen/README_files/1000_64.dconf.
 *   orc_table_bilinear / orc_table_build / orc_speed_pass_table   the build's own
 *       table format: exact at the nodes (the node value itself), exact on
 *       bilinear functions (a table of a + b x + c y + d x y reproduces it),
 *       periodic wrap; the nodes themselves are orc_xy / orc_dxy.
 *   orc_run                                 not run against the reference's own
 *       loop (simulation.cpp needs UMFPACK, absent from the image).  Pinned by
 *       what the mathematics fixes: accept-all runs converge at second order to
 *       a DOP853 integration of dx/dt = v(x; tau(t)) built on the pinned speed
 *       pass, under a stress that changes in time; copies with a planted
 *       mistake (ORC_MUTANT below: theta = 1, a wrong stage time, the sign of
 *       g) fail that check; plus invariants (strain = sum b dx, step counts).
 */
#include "d4_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846
/* 2*pi/ln 2 (constants.h:45) */
#define ORC_PI2LN2INV 9.0647202836543876192553658914333336203437229354476

/* cosh(2*pi*k) and sinh(2*pi*k): numerical constants of constants.h:54-61 */
static const double CH1 = 267.746761483748222245931879901;
static const double CH2 = 143375.656570070321051450310133;
static const double CH3 = 7.67764676977233502175191858009e7;
static const double CH4 = 4.11131577927974976374895337541e10;
static const double CH5 = 2.20157529303160145057002722946e13;
static const double SH1 = 267.744894041016514257117449688;
static const double SH2 = 143375.656566582978695241314640;

/* ------------------------------------------------------------------ utils */

/* sim_utils.cpp:6-13: repeated +-1 until the value lies in [-0.5, 0.5) */
double orc_normalize(double v)
{
    while (v < -0.5) v += 1;
    while (v >= 0.5) v -= 1;
    return v;
}

/* sim_utils.cpp:36-54 */
double orc_weight(double s, char kind)
{
    if (s > 0) {
        if (kind == 'c') { double q = s / (s + 1); return q * q; }
        if (kind == 'p') return s / (s + 1);
        if (kind == 'm') { double c = 0.412073; return 1 / (1 + 1.5 / (s + c * s * s)); }
    }
    return 0;
}

/* sim_utils.cpp:57-75 */
double orc_weight_inv(double w, char kind)
{
    if (w > 0) {
        if (kind == 'c') return sqrt(w) / (1 - sqrt(w));
        if (kind == 'p') return w / (1 - w);
        if (kind == 'm') {
            double c = 0.412073;
            return (1 - w - sqrt(1 - 2 * w + 6 * c * w + w * w - 6 * c * w * w)) / (2 * (-c + c * w));
        }
    }
    return 0;
}

/* ------------------------------------------------------------ field math */

/* Hyperbolic values of 2*pi*(dx+k) for k = -5..5 (index k+5), built from
 * E = exp(2*pi*dx) with the identities of constants.h:63-79.  For |k| >= 3
 * only the cosh value exists; the reference reuses it as "sinh". */
typedef struct { double ch[11]; double sh[11]; } hyp_t;

static void hyp_fill(double E, hyp_t* h)
{
    double P = E + 1 / E, M = E - 1 / E;
    h->ch[5] = P / 2;
    h->sh[5] = M / 2;
    h->ch[6] = (P * CH1 + M * SH1) / 2;
    h->ch[4] = (P * CH1 - M * SH1) / 2;
    h->sh[6] = (M * CH1 + P * SH1) / 2;
    h->sh[4] = (M * CH1 - P * SH1) / 2;
    h->ch[7] = (P * CH2 + M * SH2) / 2;
    h->ch[3] = (P * CH2 - M * SH2) / 2;
    h->sh[7] = (M * CH2 + P * SH2) / 2;
    h->sh[3] = (M * CH2 - P * SH2) / 2;
    h->ch[8] = CH3 * E;  h->ch[2] = CH3 / E;
    h->ch[9] = CH4 * E;  h->ch[1] = CH4 / E;
    h->ch[10] = CH5 * E; h->ch[0] = CH5 / E;
    h->sh[8] = h->ch[8]; h->sh[2] = h->ch[2];
    h->sh[9] = h->ch[9]; h->sh[1] = h->ch[1];
    h->sh[10] = h->ch[10]; h->sh[0] = h->ch[0];
}

/* Stress of one image wall at x-offset u (Field.cpp:191-197):
 * 2*pi*u * (C c - 1) / (C - c)^2 * pi */
static double wall(double u, double c, double C)
{
    double a = ORC_PI * 2 * u;
    double D = C - c;
    return a * (C * c - 1) / (D * D) * ORC_PI;
}

/* Series of the central wall used when dx^2+dy^2 <= 1e-6 (Field.cpp:182-187):
 * cosh-1, cos-1 and cosh-cos expanded to 6th order. */
static void near_series(double a, double b, double* chm1, double* cm1, double* cmc)
{
    *chm1 = pow(a, 6) / 720 + pow(a, 4) / 24 + pow(a, 2) * 0.5;
    *cm1 = -pow(b, 6) / 720 + pow(b, 4) / 24 - pow(b, 2) * 0.5;
    *cmc = (pow(b, 6) + pow(a, 6)) / 720 + (pow(a, 4) - pow(b, 4)) / 24 + (pow(a, 2) + pow(b, 2)) * 0.5;
}

/* Central wall (Field.cpp:172-188) */
static double wall0(double dx, double c, double dy, double C)
{
    double a = ORC_PI * 2 * dx;
    if (dx * dx + dy * dy > 1e-6) {
        double D = C - c;
        return a * (C * c - 1) / (D * D) * ORC_PI;
    }
    double b = ORC_PI * 2 * dy, chm1, cm1, cmc;
    near_series(a, b, &chm1, &cm1, &cmc);
    return ((chm1 * cm1 + cm1 + chm1) * a) / (cmc * cmc) * ORC_PI;
}

/* x-derivative of one wall given cosh and sinh of 2*pi*u (Field.cpp:253-266) */
static double wall_dx(double u, double c, double C, double S)
{
    double a = u * 2 * ORC_PI;
    double D = C - c;
    double Nn = C * c - 1;
    return (Nn / (D * D) + a * (S * c / (D * D) - Nn / pow(D, 3) * 2 * S)) * ORC_PI * 2 * ORC_PI;
}

/* x-derivative of the central wall (Field.cpp:209-250) */
static double wall0_dx(double dx, double c, double dy, double C, double S)
{
    double a = dx * 2 * ORC_PI;
    if (dx * dx + dy * dy > 1e-6) {
        double D = C - c;
        double Nn = C * c - 1;
        return (Nn / (D * D) + a * (S * c / (D * D) - Nn / pow(D, 3) * 2 * S)) * ORC_PI * 2 * ORC_PI;
    }
    double b = dy * ORC_PI * 2, chm1, cm1, cmc;
    near_series(a, b, &chm1, &cm1, &cmc);
    /* cos(b)*sinh(a) expanded (Field.cpp:229-244) */
    double a3 = pow(a, 3), a5 = pow(a, 5), a7 = pow(a, 7);
    double b2 = pow(b, 2), b4 = pow(b, 4), b6 = pow(b, 6);
    double cs = -(a7 * b6) / 3628800 - (a5 * b6) / 86400 - (a3 * b6) / 4320 - (b6 * a) / 720
              + (a7 * b4) / 120960 + (a5 * b4) / 2880 + (a3 * b4) / 144 + (b4 * a) / 24
              - (a7 * b2) / 1080 - (a5 * b2) / 240 - (a3 * b2) / 12 - (b2 * a) * 0.5
              + a7 / 5040 + a5 / 120 + a3 / 6 + a;
    double num = chm1 * cm1 + cm1 + chm1;
    return (num / (cmc * cmc) + ((cs * dx) / (cmc * cmc) - num / pow(cmc, 3) * dx * sinh(a) * 2) * 2 * ORC_PI)
           * ORC_PI * 2 * ORC_PI;
}

#define CH(k) (h.ch[(k) + 5])
#define SH(k) (h.sh[(k) + 5])

/* Field::xy 4-arg (Field.cpp:295-312) with g3 (Field.cpp:199-206) */
double orc_xy4(double dx, double dy, double E, double c)
{
    hyp_t h;
    hyp_fill(E, &h);
    double g3 = wall(dx + 3, c, CH(3)) + wall(dx - 3, c, CH(-3)) + wall(dx + 2, c, CH(2)) + wall(dx - 2, c, CH(-2))
              + wall(dx + 1, c, CH(1)) + wall(dx - 1, c, CH(-1)) + wall0(dx + 0, c, dy, CH(0));
    if (dx < 0)
        return wall(dx + 5, c, CH(5)) * (0 - dx) + wall(dx - 4, c, CH(-4)) * (1 + dx) + wall(dx + 4, c, CH(4)) + g3;
    return wall(dx - 5, c, CH(-5)) * (0 + dx) + wall(dx + 4, c, CH(4)) * (1 - dx) + wall(dx - 4, c, CH(-4)) + g3;
}

/* Field::xy_diff_x 4-arg (Field.cpp:314-331) with g3_dx (Field.cpp:285-292).
 * exact == 0 keeps the reference's f_dx_h (cosh passed as sinh for |k| >= 3);
 * exact != 0 passes the true sinh, whose sign follows the shift. */
double orc_dxy4(double dx, double dy, double E, double c, int exact)
{
    hyp_t h;
    hyp_fill(E, &h);
    double sg = exact ? -1.0 : 1.0; /* sign of "sinh" for negative shifts */
    double g3 = wall0_dx(dx + 0, c, dy, CH(0), SH(0))
              + wall_dx(dx + 1, c, CH(1), SH(1)) + wall_dx(dx - 1, c, CH(-1), SH(-1))
              + wall_dx(dx + 2, c, CH(2), SH(2)) + wall_dx(dx - 2, c, CH(-2), SH(-2))
              + wall_dx(dx + 3, c, CH(3), CH(3)) + wall_dx(dx - 3, c, CH(-3), sg * CH(-3));
    if (dx < 0)
        return wall_dx(dx + 5, c, CH(5), CH(5)) * (0 - dx) + wall(dx + 5, c, CH(5)) * (-1)
             + wall_dx(dx - 4, c, CH(-4), sg * CH(-4)) * (1 + dx) + wall(dx - 4, c, CH(-4)) * (+1)
             + wall_dx(dx + 4, c, CH(4), CH(4)) + g3;
    return wall_dx(dx - 5, c, CH(-5), sg * CH(-5)) * (0 + dx) + wall(dx - 5, c, CH(-5)) * (+1)
         + wall_dx(dx + 4, c, CH(4), CH(4)) * (1 - dx) + wall(dx + 4, c, CH(4)) * (-1)
         + wall_dx(dx - 4, c, CH(-4), sg * CH(-4)) + g3;
}

/* d/d(dy) of one wall through c = cos(2 pi dy): dc/ddy = -2 pi sin(2 pi dy);
 * d/dc [(C c - 1)/(C - c)^2] = (C^2 + C c - 2)/(C - c)^3 (closed form). */
static double wall_dy(double u, double c, double s2piy, double C)
{
    double D = C - c;
    return 2 * ORC_PI * ORC_PI * u * (C * C + C * c - 2) / (D * D * D) * (-2 * ORC_PI * s2piy);
}

/* d sigma / d dy (no reference counterpart): the closed form of every wall,
 * the central one included (its cancellation costs ~1e-11 relative at
 * r^2 = 1e-6, far below the 6th-order series' own truncation there). */
double orc_dyxy4(double dx, double dy, double E, double c)
{
    hyp_t h;
    hyp_fill(E, &h);
    double s = sin(2 * ORC_PI * dy);
    double g3 = wall_dy(dx + 3, c, s, CH(3)) + wall_dy(dx - 3, c, s, CH(-3)) + wall_dy(dx + 2, c, s, CH(2))
              + wall_dy(dx - 2, c, s, CH(-2)) + wall_dy(dx + 1, c, s, CH(1)) + wall_dy(dx - 1, c, s, CH(-1))
              + wall_dy(dx, c, s, CH(0));
    if (dx < 0)
        return wall_dy(dx + 5, c, s, CH(5)) * (0 - dx) + wall_dy(dx - 4, c, s, CH(-4)) * (1 + dx)
             + wall_dy(dx + 4, c, s, CH(4)) + g3;
    return wall_dy(dx - 5, c, s, CH(-5)) * (0 + dx) + wall_dy(dx + 4, c, s, CH(4)) * (1 - dx)
         + wall_dy(dx - 4, c, s, CH(-4)) + g3;
}

/* Field.cpp:333-338 */
double orc_xy(double dx, double dy)
{
    return orc_xy4(dx, dy, exp2(ORC_PI2LN2INV * dx), cos(ORC_PI * 2 * dy));
}

/* Field.cpp:341-346 */
double orc_dxy(double dx, double dy)
{
    return orc_dxy4(dx, dy, exp(ORC_PI * 2 * dx), cos(ORC_PI * 2 * dy), 0);
}

/* simulation.cpp:456-457 */
void orc_field_inputs(double dx, double dy, double* e, double* c)
{
    *e = exp2(ORC_PI2LN2INV * dx);
    *c = cos(2 * ORC_PI * dy);
}

void orc_field_batch(int64_t n, const double* dx, const double* dy, const double* e, const double* c,
                     double* sxy, double* dsxy, int exact)
{
    for (int64_t k = 0; k < n; ++k) {
        double E, C;
        if (e && c) { E = e[k]; C = c[k]; }
        else orc_field_inputs(dx[k], dy[k], &E, &C);
        if (sxy) sxy[k] = orc_xy4(dx[k], dy[k], E, C);
        if (dsxy) dsxy[k] = orc_dxy4(dx[k], dy[k], E, C, exact);
    }
}

/* ---------------------------------------------------------------- passes */

static int bsign(int i, int n) { return i < n / 2 ? 1 : -1; } /* simulation_data.cpp:265-270 */

/* precision_handler.cpp:34-38 */
static void upd_tol(double* tol, double cap2, double d2, int id)
{
    if (tol && cap2 > 0 && d2 < cap2 && d2 < tol[id]) tol[id] = d2;
}

/* simulation.cpp:287-333 */
void orc_speed_pass(int n, const double* xy, double tau, double* v, double* tol, double cap2)
{
    for (int i = 0; i < n; ++i) v[i] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dx = orc_normalize(xy[2 * i] - xy[2 * j]);
            double dy = orc_normalize(xy[2 * i + 1] - xy[2 * j + 1]);
            double d2 = dx * dx + dy * dy;
            upd_tol(tol, cap2, d2, i);
            upd_tol(tol, cap2, d2, j);
            double f = orc_xy(dx, dy) * bsign(i, n) * bsign(j, n);
            v[i] += f;
            v[j] -= f;
        }
        v[i] += tau * bsign(i, n);
    }
}

/* ---------------------------------------------------- tabulated field
 * The build's own format (no reference counterpart): n x n node values over
 * the periodic cell, tab[ky*n + kx] at (-1/2 + kx/n, -1/2 + ky/n), bilinear
 * in between.  Written out independently of the kernel's table_bilinear, with
 * the same order of operations (the library is built without contraction). */
double orc_table_bilinear(const double* tab, int n, double dx, double dy)
{
    double u = (dx + 0.5) * n, w = (dy + 0.5) * n;
    double fu = floor(u), fw = floor(w);
    double ax = u - fu, ay = w - fw;
    int i0 = ((int)fu) & (n - 1), j0 = ((int)fw) & (n - 1);
    int i1 = (i0 + 1) & (n - 1), j1 = (j0 + 1) & (n - 1);
    double t00 = tab[(size_t)j0 * n + i0], t10 = tab[(size_t)j0 * n + i1];
    double t01 = tab[(size_t)j1 * n + i0], t11 = tab[(size_t)j1 * n + i1];
    return (1.0 - ax) * ((1.0 - ay) * t00 + ay * t01) + ax * ((1.0 - ay) * t10 + ay * t11);
}

void orc_table_build(int n, double* tab_s, double* tab_d)
{
    for (int ky = 0; ky < n; ++ky)
        for (int kx = 0; kx < n; ++kx) {
            double dx = -0.5 + (double)kx / n, dy = -0.5 + (double)ky / n;
            size_t k = (size_t)ky * n + kx;
            int origin = dx == 0.0 && dy == 0.0;
            tab_s[k] = origin ? 0.0 : orc_xy(dx, dy);
            if (tab_d) tab_d[k] = origin ? 0.0 : orc_dxy(dx, dy);
        }
}

/* orc_speed_pass with sigma from the table */
void orc_speed_pass_table(int n, const double* xy, const double* tab, int tn, double tau, double* v)
{
    for (int i = 0; i < n; ++i) v[i] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dx = orc_normalize(xy[2 * i] - xy[2 * j]);
            double dy = orc_normalize(xy[2 * i + 1] - xy[2 * j + 1]);
            double f = orc_table_bilinear(tab, tn, dx, dy) * bsign(i, n) * bsign(j, n);
            v[i] += f;
            v[j] -= f;
        }
        v[i] += tau * bsign(i, n);
    }
}

/* Column lookup replacing getElement (simulation.cpp:662-685): binary search of
 * row r in column col's ascending rows. */
static double csc_get(const int* Ap, const int* Ai, const double* Ax, int r, int col)
{
    int lo = Ap[col], hi = Ap[col + 1];
    while (hi - lo > 1) {
        int mid = lo + (hi - lo) / 2;
        if (Ai[mid] > r) hi = mid; else lo = mid;
    }
    if (lo < Ap[col + 1] && Ai[lo] == r) return Ax[lo];
    return 0;
}

/* simulation.cpp:345-527 */
int orc_jacobian_pass(int n, const double* xy, double h, double tauA, double tauB,
                      double cutoff, double cut_sq, double one_per, int heaviside, char wfunc,
                      double* vA, double* vB, int* Ap, int* Ai, double* Ax, int cap,
                      int* indexes, double* dvec, double* tol, double cap2)
{
    double rng2 = heaviside ? 1 : 49; /* simulation.cpp:461-463 */
    /* count first so that a short capacity is reported before writing */
    long need = n;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = orc_normalize(xy[2 * i] - xy[2 * j]);
            double dy = orc_normalize(xy[2 * i + 1] - xy[2 * j + 1]);
            if (dx * dx + dy * dy < rng2 * cut_sq) need += 2;
        }
    if (need > cap) return (int)-need;

    for (int i = 0; i < n; ++i) { vA[i] = 0; vB[i] = 0; }
    int nz = 0;
    Ap[0] = 0;
    for (int i = 0; i < n; ++i) {
        double sub = 0;
        for (int j = 0; j < i; ++j) { /* previously computed column entries (simulation.cpp:362-374) */
            double v = csc_get(Ap, Ai, Ax, i, j);
            if (v != 0) { Ai[nz] = j; Ax[nz] = v; sub -= v; nz++; }
        }
        Ai[nz] = i;
        indexes[i] = nz;
        Ax[nz] = 0; /* -tmp * h with no point defects (simulation.cpp:441) */
        nz++;
        for (int j = i + 1; j < n; ++j) { /* simulation.cpp:444-482 */
            double dx = orc_normalize(xy[2 * i] - xy[2 * j]);
            double dy = orc_normalize(xy[2 * i + 1] - xy[2 * j + 1]);
            double d2 = dx * dx + dy * dy;
            upd_tol(tol, cap2, d2, i);
            upd_tol(tol, cap2, d2, j);
            double E, c;
            orc_field_inputs(dx, dy, &E, &c);
            double f = orc_xy4(dx, dy, E, c) * bsign(i, n) * bsign(j, n);
            if (d2 < rng2 * cut_sq) {
                double v = h * bsign(j, n) * bsign(i, n) * orc_dxy4(dx, dy, E, c, 0);
                if (!heaviside && d2 > cut_sq) v *= exp(-pow(sqrt(d2) - cutoff, 2) * one_per);
                Ai[nz] = j; Ax[nz] = v; sub -= v; nz++;
            }
            vA[i] += f;
            vA[j] -= f;
        }
        Ap[i + 1] = nz;
        vB[i] = vA[i] + tauB * bsign(i, n);
        vA[i] += tauA * bsign(i, n);
        Ax[indexes[i]] = sub;
        dvec[i] = orc_weight(sub, wfunc);
    }
    for (int i = 0; i < n; ++i) { /* simulation.cpp:516-522 */
        for (int k = Ap[i]; k < Ap[i + 1]; ++k) Ax[k] *= (1 + dvec[Ai[k]]) / 2;
        Ax[indexes[i]] += 1;
    }
    return nz;
}

/* simulation.cpp:546-608 */
void orc_jacobian_from_prev(int n, int* Ap, int* Ai, double* Ax, double* dvec, char wfunc,
                            double base_t, double new_t, double old_t,
                            double tau_old, double tau_new, double* speed)
{
    double factor = (new_t - base_t) / (old_t - base_t);
    double* nw = (double*)malloc(sizeof(double) * (size_t)n);
    double* wf = (double*)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        double ow = dvec[i];
        nw[i] = (ow == 0) ? 0. : orc_weight(orc_weight_inv(ow, wfunc) * factor, wfunc);
        wf[i] = (1 + nw[i]) / (1 + ow);
    }
    for (int j = 0; j < n; ++j)
        for (int k = Ap[j]; k < Ap[j + 1]; ++k) {
            double f = wf[Ai[k]];
            if (Ai[k] == j) Ax[k] = 1 + (Ax[k] - 1) / 2 * f;
            else Ax[k] = Ax[k] / 2 * f;
        }
    memcpy(dvec, nw, sizeof(double) * (size_t)n);
    double dif = tau_new - tau_old;
    for (int i = 0; i < n; ++i) speed[i] += dif * bsign(i, n);
    free(nw);
    free(wf);
}

/* Dense counterpart of orc_jacobian_pass with an infinite cutoff: same pair
 * orientation (lower index minus higher), M = V + diag(s) unscaled. */
void orc_jacobian_dense(int n, const double* xy, double h, double tauA, double tauB, char wfunc,
                        double* vA, double* vB, double* M, double* dvec, double* tol, double cap2)
{
    for (int i = 0; i < n; ++i) { vA[i] = 0; vB[i] = 0; }
    for (int i = 0; i < n; ++i) M[(size_t)i * n + i] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = orc_normalize(xy[2 * i] - xy[2 * j]);
            double dy = orc_normalize(xy[2 * i + 1] - xy[2 * j + 1]);
            double d2 = dx * dx + dy * dy;
            upd_tol(tol, cap2, d2, i);
            upd_tol(tol, cap2, d2, j);
            double E, c;
            orc_field_inputs(dx, dy, &E, &c);
            double f = orc_xy4(dx, dy, E, c) * bsign(i, n) * bsign(j, n);
            double v = h * bsign(j, n) * bsign(i, n) * orc_dxy4(dx, dy, E, c, 0);
            M[(size_t)i * n + j] = v;
            M[(size_t)j * n + i] = v;
            vA[i] += f;
            vA[j] -= f;
        }
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j)
            if (j != i) s -= M[(size_t)i * n + j];
        M[(size_t)i * n + i] = s;
        dvec[i] = orc_weight(s, wfunc);
        vB[i] = vA[i] + tauB * bsign(i, n);
        vA[i] += tauA * bsign(i, n);
    }
}

/* simulation_data.cpp:253-262 */
void orc_cutoff(int n, double mult, int heaviside, double* cutoff, double* cut_sq, double* one_per)
{
    double m = mult;
    if (!heaviside && mult >= sqrt(2.0 * n) / 12) m = 1e20;
    *cutoff = m / sqrt((double)n);
    *cut_sq = *cutoff * *cutoff;
    *one_per = 1 / *cut_sq;
}

/* stress_protocol.cpp:13-73 */
double orc_ext_stress(int kind, double init, double rate, double period, double amp, double t)
{
    if (kind == 1) return t * rate + init;
    if (kind == 2 || kind == 3) {
        double pt = t - floor((t / period + 1. / 4)) * period;
        double s = rate * pt + init;
        if (pt > period / 4) s = rate * period / 2 - s + init;
        if (kind == 3) return s * amp * t;
        return s;
    }
    return 0;
}

/* ----------------------------------------------------------- dense solve */

/* In-place LU with partial pivoting of column-major a (n x n); returns 0 or
 * 1 + index of a zero pivot. */
static int lu_factor(int n, double* a, int* piv)
{
    int sing = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[(size_t)k * n + k]);
        for (int r = k + 1; r < n; ++r) {
            double v = fabs(a[(size_t)k * n + r]);
            if (v > best) { best = v; p = r; }
        }
        piv[k] = p;
        if (p != k)
            for (int c = 0; c < n; ++c) {
                double t = a[(size_t)c * n + k]; a[(size_t)c * n + k] = a[(size_t)c * n + p]; a[(size_t)c * n + p] = t;
            }
        double d = a[(size_t)k * n + k];
        if (d == 0) { if (!sing) sing = k + 1; continue; }
        for (int r = k + 1; r < n; ++r) a[(size_t)k * n + r] /= d;
        /* columns are independent: threads change which core updates a column, not its operations */
#pragma omp parallel for schedule(static) if (n - k > 256)
        for (int c = k + 1; c < n; ++c) {
            double m = a[(size_t)c * n + k];
            if (m == 0) continue;
            double* col = a + (size_t)c * n;
            const double* lk = a + (size_t)k * n;
            for (int r = k + 1; r < n; ++r) col[r] -= lk[r] * m;
        }
    }
    return sing;
}

static void lu_solve(int n, const double* a, const int* piv, double* b)
{
    for (int k = 0; k < n; ++k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int c = 0; c < n; ++c) {
        double v = b[c];
        if (v != 0) for (int r = c + 1; r < n; ++r) b[r] -= a[(size_t)c * n + r] * v;
    }
    for (int c = n - 1; c >= 0; --c) {
        b[c] /= a[(size_t)c * n + c];
        double v = b[c];
        if (v != 0) for (int r = 0; r < c; ++r) b[r] -= a[(size_t)c * n + r] * v;
    }
}

/* ------------------------------------------------------------- whole run */

typedef struct {
    int n;
    int *Ap, *Ai, *idx, *piv;
    double *Ax, *dense, *dvec, *x, *g;
    int cap, sing;
} orc_sys;

static void sys_factor(orc_sys* s)
{
    int n = s->n;
    memset(s->dense, 0, sizeof(double) * (size_t)n * n);
    for (int c = 0; c < n; ++c)
        for (int k = s->Ap[c]; k < s->Ap[c + 1]; ++k) s->dense[(size_t)c * n + s->Ai[k]] = s->Ax[k];
    s->sing = lu_factor(n, s->dense, s->piv);
}

/* simulation.cpp:643-659 (solve, singular handling) */
static void sys_solve(orc_sys* s)
{
    memcpy(s->x, s->g, sizeof(double) * (size_t)s->n);
    lu_solve(s->n, s->dense, s->piv, s->x);
    if (s->sing)
        for (int i = 0; i < s->n; ++i)
            if (!((-1 < s->x[i]) && (s->x[i] < 1))) s->x[i] = 0;
}

/* Planted mistakes, for the tests that show the run loop's checks can fail (never defined in the normal
 * builds; tests/test_oracle.py compiles a copy with -DORC_MUTANT=k):
 *   1  theta_i = 1 in every g (the implicit end point alone: backward Euler);
 *   2  stage III's second speed pass at t1/2 instead of t1 (a wrong stage time);
 *   3  the sign of g in every solve. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define THP(d) (ORC_MUTANT == 1 ? 2.0 : 1 + (d))
#define THM(d) (ORC_MUTANT == 1 ? 0.0 : 1 - (d))
#define GSIGN (ORC_MUTANT == 3 ? -1.0 : 1.0)

int orc_run(int n, double* xy, const orc_run_params* p, orc_run_stats* st)
{
    /* PrecisionHandler setup (simulation.cpp:24-26, precision_handler.cpp:75-89) */
    double minprec2 = p->prec * p->prec, dip2 = 1;
    int use_dip = p->dipole_prec != 0;
    if (use_dip) { dip2 = p->dipole_prec * p->dipole_prec; minprec2 /= dip2; }
    double cap2 = use_dip ? minprec2 : 0;

    double cutoff, cut_sq, one_per, mult = p->cutoff_mult;
    orc_cutoff(n, mult, p->heaviside, &cutoff, &cut_sq, &one_per);

    size_t nd = (size_t)n;
    double* buf = (double*)calloc(nd * 2 * 4 + nd * 8, sizeof(double));
    double *disl = xy, *big = buf, *s1 = buf + 2 * nd, *s2 = buf + 4 * nd, *tmpxy = buf + 6 * nd;
    double *init = buf + 8 * nd, *init2 = init + nd, *sp = init2 + nd, *sp2 = sp + nd, *tol = sp2 + nd;
    (void)tmpxy;
    orc_sys S;
    S.n = n;
    S.cap = n * n;
    S.Ap = (int*)calloc(nd + 1, sizeof(int));
    S.Ai = (int*)calloc((size_t)S.cap, sizeof(int));
    S.Ax = (double*)calloc((size_t)S.cap, sizeof(double));
    S.idx = (int*)calloc(nd, sizeof(int));
    S.piv = (int*)calloc(nd, sizeof(int));
    S.dense = (double*)calloc(nd * nd, sizeof(double));
    S.dvec = (double*)calloc(nd, sizeof(double));
    S.x = (double*)calloc(nd, sizeof(double));
    S.g = (double*)calloc(nd, sizeof(double));
    for (int i = 0; i < n; ++i) tol[i] = minprec2;

    double t = 0, h = p->dt0, strain = 0, maxratio = 0;
    long ok = 0, failed = 0;
#define TAU(tt) orc_ext_stress(p->stress_kind, p->stress_init, p->stress_rate, p->stress_period, p->stress_amp, (tt))
    orc_speed_pass(n, disl, TAU(t), init, tol, cap2); /* simulation.cpp:46 */

    for (;;) {
        /* interrupt() (simulation.cpp:722-750) */
        if (p->time_limit > 0 && t > p->time_limit) break;
        if (p->strain_limit > 0 && strain > p->strain_limit) break;
        if (p->step_limit >= 0 && ok > p->step_limit) break;

        double t0 = t, t1 = t + h, th = t + h / 2;
        /* stage I */
        orc_jacobian_pass(n, disl, h, TAU(t0), TAU(t1), cutoff, cut_sq, one_per, p->heaviside, p->wfunc,
                          init, sp, S.Ap, S.Ai, S.Ax, S.cap, S.idx, S.dvec, tol, cap2);
        sys_factor(&S);
        for (int i = 0; i < n; ++i) S.g[i] = -GSIGN * h * (THP(S.dvec[i]) * sp[i] + THM(S.dvec[i]) * init[i]) / 2;
        sys_solve(&S);
        for (int i = 0; i < n; ++i) { big[2 * i] = disl[2 * i] - S.x[i]; big[2 * i + 1] = disl[2 * i + 1]; }
        orc_speed_pass(n, big, TAU(t1), sp2, tol, cap2);
        for (int i = 0; i < n; ++i)
            S.g[i] = GSIGN * (big[2 * i] - disl[2 * i] - h * (THP(S.dvec[i]) * sp2[i] + THM(S.dvec[i]) * init[i]) / 2);
        sys_solve(&S);
        for (int i = 0; i < n; ++i) big[2 * i] -= S.x[i];
        /* stage II */
        orc_jacobian_from_prev(n, S.Ap, S.Ai, S.Ax, S.dvec, p->wfunc, t0, th, t1, TAU(t1), TAU(th), sp);
        sys_factor(&S);
        for (int i = 0; i < n; ++i) S.g[i] = -GSIGN * h / 2 * (THP(S.dvec[i]) * sp[i] + THM(S.dvec[i]) * init[i]) / 2;
        sys_solve(&S);
        for (int i = 0; i < n; ++i) { s1[2 * i] = disl[2 * i] - S.x[i]; s1[2 * i + 1] = disl[2 * i + 1]; }
        orc_speed_pass(n, s1, TAU(th), sp2, tol, cap2);
        for (int i = 0; i < n; ++i)
            S.g[i] = GSIGN * (s1[2 * i] - disl[2 * i] - h / 2 * (THP(S.dvec[i]) * sp2[i] + THM(S.dvec[i]) * init[i]) / 2);
        sys_solve(&S);
        for (int i = 0; i < n; ++i) s1[2 * i] -= S.x[i];
        /* stage III */
        orc_jacobian_pass(n, s1, h / 2, TAU(th), TAU(t1), cutoff, cut_sq, one_per, p->heaviside, p->wfunc,
                          init2, sp2, S.Ap, S.Ai, S.Ax, S.cap, S.idx, S.dvec, tol, cap2);
        sys_factor(&S);
        for (int i = 0; i < n; ++i) S.g[i] = -GSIGN * h / 2 * (THP(S.dvec[i]) * sp2[i] + THM(S.dvec[i]) * init2[i]) / 2;
        sys_solve(&S);
        for (int i = 0; i < n; ++i) { s2[2 * i] = s1[2 * i] - S.x[i]; s2[2 * i + 1] = s1[2 * i + 1]; }
        orc_speed_pass(n, s2, TAU(ORC_MUTANT == 2 ? th : t1), sp2, tol, cap2);
        for (int i = 0; i < n; ++i)
            S.g[i] = GSIGN * (s2[2 * i] - s1[2 * i] - h / 2 * (THP(S.dvec[i]) * sp2[i] + THM(S.dvec[i]) * init2[i]) / 2);
        sys_solve(&S);
        for (int i = 0; i < n; ++i) s2[2 * i] -= S.x[i];
        /* stage IV: error (simulation.cpp:692-700, precision_handler.cpp:41-54) */
        double mr = 0;
        for (int i = 0; i < n; ++i) {
            double e = fabs(big[2 * i] - s2[2 * i]);
            double r = use_dip ? e * e / tol[i] : e * e / minprec2;
            if (r > mr) mr = r;
        }
        double ratio = mr / dip2; /* precision_handler.cpp:92-95 */
        maxratio = ratio;
        if (ratio < 1) {
            t += h;
            ok++;
            for (int i = 0; i < n; ++i) {
                double b = bsign(i, n);
                strain += b * (s1[2 * i] - disl[2 * i]);
            }
            for (int i = 0; i < n; ++i) {
                double b = bsign(i, n);
                strain += b * (s2[2 * i] - s1[2 * i]);
            }
            for (int i = 0; i < 2 * n; ++i) disl[i] = s2[i];
            for (int i = 0; i < n; ++i) disl[2 * i] = orc_normalize(disl[2 * i]);
            if (p->speed_threshold > 0) { /* simulation.cpp:215-219 */
                double avg = 0;
                for (int i = 0; i < n; ++i) avg += fabs(init2[i]);
                avg /= n;
                if (p->speed_threshold > avg) { mult = 1e20; orc_cutoff(n, mult, p->heaviside, &cutoff, &cut_sq, &one_per); }
            }
        } else failed++;
        /* precision_handler.cpp:57-62 */
        double fac = fmin(2., pow(ratio, -1. / 6));
        h = 0.9 * h * fac;
        for (int i = 0; i < n; ++i) tol[i] = minprec2; /* pH->reset() */
        if (p->max_dt > 0 && p->max_dt < h) h = p->max_dt;
    }
    if (st) {
        st->sim_time = t; st->step_size = h; st->strain = strain;
        st->ok_steps = ok; st->failed_steps = failed; st->last_max_ratio = maxratio;
    }
    free(buf); free(S.Ap); free(S.Ai); free(S.Ax); free(S.idx); free(S.piv); free(S.dense);
    free(S.dvec); free(S.x); free(S.g);
    return 0;
#undef TAU
}
#undef THP
#undef THM
#undef GSIGN

/* ------------------------------------------------------------ .dconf read */

typedef struct { double x, y, b; unsigned id; } rec_t;

static int rec_cmp(const void* pa, const void* pb)
{
    const rec_t* a = (const rec_t*)pa;
    const rec_t* b = (const rec_t*)pb;
    double ka = a->y + a->b, kb = b->y + b->b; /* simulation_data.cpp:116: descending y+b */
    if (ka > kb) return -1;
    if (ka < kb) return 1;
    return (a->id < b->id) ? -1 : (a->id > b->id);
}

/* simulation_data.cpp:76-154 (parse, sum-b check, sort, disl_order) */
int orc_read_dconf(const char* path, double* xy, unsigned* order, int cap)
{
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    rec_t* r = (rec_t*)malloc(sizeof(rec_t) * (size_t)(cap > 0 ? cap : 1));
    int n = 0;
    double sb = 0, x, y, b;
    while (fscanf(f, "%lf %lf %lf", &x, &y, &b) == 3) {
        if (n >= cap) { fclose(f); free(r); return -2; }
        if (fabs(b - rint(b)) > 1e-5) { fclose(f); free(r); return -3; }
        r[n].x = x; r[n].y = y; r[n].b = b; r[n].id = (unsigned)n;
        sb += b;
        n++;
    }
    fclose(f);
    if (sb != 0) { free(r); return -4; }
    qsort(r, (size_t)n, sizeof(rec_t), rec_cmp);
    for (int i = 0; i < n; ++i) {
        xy[2 * i] = r[i].x;
        xy[2 * i + 1] = r[i].y;
        order[r[i].id] = (unsigned)i;
    }
    free(r);
    return n;
}

/* ------------------------------------------------- multi-threaded variants */

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_speed_pass_omp(int n, const double* xy, double tau, double* v)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        double acc = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            int lo = i < j ? i : j, hi = i < j ? j : i;
            double dx = orc_normalize(xy[2 * lo] - xy[2 * hi]);
            double dy = orc_normalize(xy[2 * lo + 1] - xy[2 * hi + 1]);
            double f = orc_xy(dx, dy) * bsign(i, n) * bsign(j, n);
            acc += (i == lo) ? f : -f;
        }
        v[i] = acc + tau * bsign(i, n);
    }
}

void orc_jacobian_dense_omp(int n, const double* xy, double h, double tauA, double tauB, char wfunc,
                            double* vA, double* vB, double* M, double* dvec)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        double acc = 0, s = 0;
        double* col = M + (size_t)i * n;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            int lo = i < j ? i : j, hi = i < j ? j : i;
            double dx = orc_normalize(xy[2 * lo] - xy[2 * hi]);
            double dy = orc_normalize(xy[2 * lo + 1] - xy[2 * hi + 1]);
            double E, c;
            orc_field_inputs(dx, dy, &E, &c);
            double bb = (double)(bsign(i, n) * bsign(j, n));
            double f = orc_xy4(dx, dy, E, c) * bb;
            double v = h * bb * orc_dxy4(dx, dy, E, c, 0);
            col[j] = v;
            s -= v;
            acc += (i == lo) ? f : -f;
        }
        col[i] = s;
        dvec[i] = orc_weight(s, wfunc);
        vB[i] = acc + tauB * bsign(i, n);
        vA[i] = acc + tauA * bsign(i, n);
    }
}

void orc_inputs_batch(int64_t n, const double* dx, const double* dy, double* e, double* c)
{
    for (int64_t k = 0; k < n; ++k) orc_field_inputs(dx[k], dy[k], e + k, c + k);
}
