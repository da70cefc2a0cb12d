"""Pins the oracle (oracle/d4_oracle.c) before anything is checked against it.

Each function of the oracle is checked against something that a plausible
mistake (dropped term, wrong sign, transposed index, wrong orientation) fails:
the reference's own compiled sources (golden fixtures from oracle/_ref), an
independent many-image lattice sum, finite differences, closed forms and
brute force at small n.
"""
import os

import numpy as np
import pytest

from helpers import (TOL, all_pairs, bsign, c5_config, random_config, scale_dsigma, scale_sigma, wrap)

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ field

def test_field_matches_reference_o2_bitwise(orc, golden):
    """Field.cpp:295-331 at -O2: same inputs -> identical doubles (NaN where the reference gives NaN:
    (dx, dy) so close to 0 that the series branch divides 0 by 0)."""
    g = golden["field"]
    s, ds = orc.field(g["dx"], g["dy"], g["e"], g["c"])
    assert np.array_equal(s, g["xy4"], equal_nan=True)
    assert np.array_equal(ds, g["dxy4"], equal_nan=True)


def test_field_two_arg_matches_reference(orc, golden):
    """Field.cpp:333-346: the 2-arg overloads compute their own exp2/exp and cos."""
    g = golden["field"]
    e, c = orc.inputs(g["dx"], g["dy"])
    same = (e == g["e"]) & (c == g["c"])
    assert same.mean() > 0.99  # same libm here; differing inputs are checked by bound below
    s = np.array([orc.L.orc_xy(a, b) for a, b in zip(g["dx"], g["dy"])])
    ds = np.array([orc.L.orc_dxy(a, b) for a, b in zip(g["dx"], g["dy"])])
    assert np.array_equal(s[same], g["xy2"][same], equal_nan=True)
    assert np.array_equal(ds[same], g["dxy2"][same], equal_nan=True)


def test_field_matches_reference_fast_flags(orc, golden):
    """Against the reference built with -Ofast -ffast-math: reassociation only (S-scaled bound)."""
    g = golden["field"]
    s, ds = orc.field(g["dx"], g["dy"], g["e_fast"], g["c_fast"])
    S = scale_sigma(g["dx"], g["dy"], g["e_fast"], g["c_fast"])
    Sd = scale_dsigma(g["dx"], g["dy"], g["e_fast"], g["c_fast"])
    # C c - 1 and C - c of the central wall cancel for near pairs; fast-math fuses them, which moves
    # the result by ~1e-16 / (2 pi^2 r^2) relative: beyond r = 0.01 that stays below 1e-12
    far = np.hypot(g["dx"], g["dy"]) > 1e-2
    assert np.all(np.abs(s - g["xy4_fast"])[far] <= TOL * np.maximum(np.abs(s), S)[far])
    assert np.all(np.abs(ds - g["dxy4_fast"])[far] <= TOL * np.maximum(np.abs(ds), Sd)[far])


def test_quirk_known_answers(orc):
    """xy_diff_x is not d/dx of xy: f_dx_h passes cosh as sinh (Field.cpp:269-283, 291, 319-328).
    Relative gaps to a central difference measured on the reference: 2.71e-6 and -2.10e-5."""
    for (dx, dy), want in (((0.1, 0.2), 2.71e-6), ((0.45, -0.4), -2.10e-5)):
        h = 1e-5
        fd = (orc.L.orc_xy(dx + h, dy) - orc.L.orc_xy(dx - h, dy)) / (2 * h)
        rel = (orc.L.orc_dxy(dx, dy) - fd) / fd
        assert abs(rel - want) < 0.02 * abs(want)


def _lattice(dx, dy, K=40, deriv=False):
    """Independent periodic image sum with exact cosh/sinh: sum_k f(dx + k)."""
    c = np.cos(2 * np.pi * dy)
    tot = np.zeros_like(dx)
    for k in range(-K, K + 1):
        u = dx + k
        C = np.cosh(2 * np.pi * u)
        D = C - c
        N = C * c - 1
        if deriv:
            S = np.sinh(2 * np.pi * u)
            tot += N / D ** 2 + 2 * np.pi * u * (S * c / D ** 2 - 2 * N * S / D ** 3)
        else:
            tot += u * N / D ** 2
    return 2 * np.pi ** 2 * tot


def test_sigma_is_the_periodic_image_sum(orc):
    """The blended 10-wall sum approximates the infinite wall sum to ~1e-10 (image truncation),
    which a dropped or mis-signed wall would exceed by orders of magnitude."""
    rng = np.random.default_rng(11)
    dx = rng.uniform(-0.5, 0.5, 3000)
    dy = rng.uniform(-0.5, 0.5, 3000)
    keep = np.hypot(dx, dy) > 0.02
    dx, dy = dx[keep], dy[keep]
    e, c = orc.inputs(dx, dy)
    s, ds_exact = orc.field(dx, dy, e, c, exact=True)
    S = scale_sigma(dx, dy, e, c)
    assert np.all(np.abs(s - _lattice(dx, dy)) <= 1e-9 * np.maximum(np.abs(s), S))
    Sd = scale_dsigma(dx, dy, e, c)
    assert np.all(np.abs(ds_exact - _lattice(dx, dy, deriv=True)) <= 1e-9 * np.maximum(np.abs(ds_exact), Sd))


def test_exact_derivative_matches_finite_differences(orc):
    """orc_dxy4(exact=1): Richardson-extrapolated central differences of sigma."""
    rng = np.random.default_rng(12)
    dx = rng.uniform(-0.45, 0.45, 200)
    dy = rng.uniform(-0.5, 0.5, 200)
    keep = np.hypot(dx, dy) > 0.05
    dx, dy = dx[keep], dy[keep]

    def sig(x):
        e, c = orc.inputs(x, dy)
        return orc.field(x, dy, e, c)[0]

    def fd(h):
        return (sig(dx + h) - sig(dx - h)) / (2 * h)

    h = 1e-3
    r1 = (4 * fd(h / 2) - fd(h)) / 3
    r2 = (4 * fd(h / 4) - fd(h / 2)) / 3
    rich = (16 * r2 - r1) / 15
    e, c = orc.inputs(dx, dy)
    _, d = orc.field(dx, dy, e, c, exact=True)
    Sd = scale_dsigma(dx, dy, e, c)
    assert np.all(np.abs(d - rich) <= 1e-8 * np.maximum(np.abs(d), Sd))


def test_dy_derivative_matches_finite_differences(orc):
    rng = np.random.default_rng(13)
    dx = rng.uniform(-0.45, 0.45, 100)
    dy = rng.uniform(-0.45, 0.45, 100)
    keep = np.hypot(dx, dy) > 0.05
    dx, dy = dx[keep], dy[keep]

    def sig(y):
        e, c = orc.inputs(dx, y)
        return orc.field(dx, y, e, c)[0]

    h = 1e-4
    fd1 = (sig(dy + h) - sig(dy - h)) / (2 * h)
    fd2 = (sig(dy + h / 2) - sig(dy - h / 2)) / h
    rich = (4 * fd2 - fd1) / 3
    e, c = orc.inputs(dx, dy)
    got = np.array([orc.L.orc_dyxy4(a, b, ee, cc) for a, b, ee, cc in zip(dx, dy, e, c)])
    assert np.all(np.abs(got - rich) <= 1e-7 * (np.abs(rich) + 1.0))


def test_sigma_symmetries(orc):
    """sigma is odd under (dx, dy) -> (-dx, -dy) up to the image truncation, and even in dy exactly."""
    rng = np.random.default_rng(14)
    dx = rng.uniform(-0.5, 0.5, 2000)
    dy = rng.uniform(-0.5, 0.5, 2000)
    e, c = orc.inputs(dx, dy)
    s, _ = orc.field(dx, dy, e, c)
    e2, c2 = orc.inputs(-dx, -dy)
    s2, _ = orc.field(-dx, -dy, e2, c2)
    S = scale_sigma(dx, dy, e, c)
    assert np.all(np.abs(s + s2) <= 1e-9 * np.maximum(np.abs(s), S))
    e3, c3 = orc.inputs(dx, -dy)
    s3, ds3 = orc.field(dx, -dy, e3, c3)
    _, ds = orc.field(dx, dy, e, c)
    assert np.array_equal(s3, s) and np.array_equal(ds3, ds)


def test_sigma_continuity(orc):
    """Continuous across dx = 0 (blend switch) and across the periodic wrap dx = +-0.5."""
    for dy in (0.07, 0.31, -0.44):
        lo, hi = orc.L.orc_xy(-1e-12, dy), orc.L.orc_xy(1e-12, dy)
        assert abs(lo - hi) <= 1e-9 * (abs(lo) + 1)
        a, b = orc.L.orc_xy(np.nextafter(0.5, 0), dy), orc.L.orc_xy(-0.5, dy)
        assert abs(a - b) <= 1e-9 * (abs(a) + 1)


# ------------------------------------------------------------------ utils

def test_normalize_matches_reference(orc, golden):
    g = golden["utils"]
    out = np.array([orc.L.orc_normalize(v) for v in g["norm_in"]])
    assert np.array_equal(out, g["norm_out"])
    assert np.array_equal(wrap(g["norm_in"]), g["norm_out"])  # the exact closed form used by the kernels


def test_weights_match_reference(orc, golden):
    g = golden["utils"]
    for k in "cpm":
        assert np.array_equal(np.array([orc.L.orc_weight(v, k.encode()) for v in g["s_in"]]), g[f"w_{k}"])
        assert np.array_equal(np.array([orc.L.orc_weight_inv(v, k.encode()) for v in g["w_in"]]), g[f"wi_{k}"])


def test_weight_closed_forms(orc):
    """sim_utils.cpp:36-75: d(s) in [0, 1), 0 for s <= 0; weightInv inverts weight."""
    s = 10 ** np.linspace(-4, 3, 50)
    for k in "cpm":
        w = np.array([orc.L.orc_weight(v, k.encode()) for v in s])
        assert np.all((w > 0) & (w < 1))
        back = np.array([orc.L.orc_weight_inv(v, k.encode()) for v in w])
        ok = w < 0.999  # the inverse is ill-conditioned as d -> 1
        assert np.allclose(back[ok], s[ok], rtol=1e-8 if k != "m" else 1e-6)
    assert np.allclose([orc.L.orc_weight(v, b"p") for v in s], s / (1 + s), rtol=1e-15)
    assert orc.L.orc_weight(-1.0, b"c") == 0.0 and orc.L.orc_weight(0.0, b"m") == 0.0


def test_stress_protocols_match_reference(orc, golden):
    g = golden["utils"]
    for k in range(4):
        got = np.array([orc.L.orc_ext_stress(k, 0.1, 0.3, 1.3, 0.7, t) for t in g["t_in"]])
        assert np.array_equal(got, g[f"stress_k{k}"])


def test_dconf_read_and_cutoff_match_reference(orc, golden):
    g = golden["dconf"]
    xy, order = orc.read_dconf(os.path.join(GOLDEN_DIR, "1000_64.dconf"))
    assert np.array_equal(xy, g["xy"])
    assert np.array_equal(order, g["order"])
    assert np.array_equal(bsign(64).astype(np.int32), g["b"])
    for mult, hv in ((1e20, 0), (0.05 * 8, 1), (0.05 * 8, 0), (2.0, 0)):
        assert np.array_equal(np.array(orc.cutoff(64, mult, bool(hv))), g[f"cut_{mult}_{hv}"])


# ------------------------------------------------------------------ passes

def _speed_brute(ref, xy, tau):
    """The i < j loop of simulation.cpp:292-332 over the reference's compiled Field::xy."""
    n = xy.shape[0]
    b = bsign(n)
    v = np.zeros(n)
    for i in range(n):
        for j in range(i + 1, n):
            dx = wrap(xy[i, 0] - xy[j, 0])
            dy = wrap(xy[i, 1] - xy[j, 1])
            f = ref.L.ref_xy(float(dx), float(dy)) * b[i] * b[j]
            v[i] += f
            v[j] -= f
        v[i] += tau * b[i]
    return v


def test_speed_pass_matches_reference_field_loop(orc, ref, golden):
    xy = golden["dconf"]["xy"]
    for tau in (0.0, 0.1):
        assert np.array_equal(orc.speed_pass(xy, tau), _speed_brute(ref, xy, tau))


def test_speed_pass_small_cases(orc):
    """n = 2: v = (+f, -f) + tau b with f = b0 b1 sigma(x0 - x1); pair sums cancel exactly."""
    xy = np.array([[0.1, 0.2], [-0.3, 0.05]])
    s = orc.L.orc_xy(float(wrap(xy[0, 0] - xy[1, 0])), float(wrap(xy[0, 1] - xy[1, 1])))
    v = orc.speed_pass(xy, 0.25)
    assert v[0] == -s + 0.25 and v[1] == s - 0.25
    xy3 = random_config(3, 1)
    v3 = orc.speed_pass(xy3, 0.0)
    assert abs(v3.sum()) <= 1e-15 * np.abs(v3).sum()


def test_speed_pass_tolerance_is_pair_minimum(orc):
    xy = random_config(50, 2)
    n = xy.shape[0]
    cap2 = 4e-3
    tol = np.full(n, cap2)
    orc.speed_pass(xy, 0.0, tol, cap2)
    i, j, dx, dy = all_pairs(xy)
    d2 = dx * dx + dy * dy
    want = np.full(n, cap2)
    np.minimum.at(want, i, np.where(d2 < cap2, d2, cap2))
    np.minimum.at(want, j, np.where(d2 < cap2, d2, cap2))
    assert np.array_equal(tol, want)


def _dense_from_csc(R, n):
    A = np.zeros((n, n))
    for col in range(n):
        for k in range(R["Ap"][col], R["Ap"][col + 1]):
            A[R["Ai"][k], col] = R["Ax"][k]
    return A


def test_jacobian_pass_structure_and_invariants(orc):
    """simulation.cpp:345-527: V symmetric, rows ascending per column, diagonal present,
    A = I + diag((1+d)/2)(V + diag(s)), s = -row sums, so A 1 = 1 to rounding."""
    xy = random_config(40, 3)
    n = 40
    h, ta, tb = 1e-3, 0.05, 0.07
    cut = orc.cutoff(n, 1e20, False)
    R = orc.jacobian_pass(xy, h, ta, tb, *cut, False, "c")
    assert R["nnz"] == n * n
    for col in range(n):
        rows = R["Ai"][R["Ap"][col]:R["Ap"][col + 1]]
        assert np.all(np.diff(rows) > 0)
        assert R["Ai"][R["indexes"][col]] == col
    A = _dense_from_csc(R, n)
    vA, vB, M, d = orc.jacobian_dense(xy, h, ta, tb, "c")
    assert np.array_equal(M, M.T)
    assert np.allclose(A, np.eye(n) + ((1 + d) / 2)[:, None] * M, rtol=0, atol=1e-15 * np.abs(M).max())
    assert np.all(np.abs(A.sum(axis=1) - 1) <= 1e-13 * np.abs(A).sum(axis=1))
    assert np.array_equal(R["dvec"], d)
    assert np.array_equal(R["vA"], vA) and np.array_equal(R["vB"], vB)
    b = bsign(n)
    assert np.all(np.abs(vB - vA - (tb - ta) * b) <= 4e-16 * (np.abs(vA) + 1))
    # the speeds of the Jacobian pass are the speed pass's
    assert np.allclose(vA, orc.speed_pass(xy, ta), rtol=0, atol=1e-13 * np.abs(vA).max())
    # d sigma per entry against the reference formula with the pair's own orientation
    i, j = 3, 17
    dx, dy = wrap(xy[i] - xy[j])
    e, c = orc.inputs(np.array([dx]), np.array([dy]))
    want = h * b[j] * b[i] * orc.field(np.array([dx]), np.array([dy]), e, c)[1][0]
    assert M[i, j] == want and M[j, i] == want


def test_jacobian_pass_cutoff_modes(orc):
    """Storage radius rc (heaviside) or 7 rc with the taper exp(-(r - rc)^2/rc^2) (simulation.cpp:461-470)."""
    n = 200
    xy = random_config(n, 4)
    i, j, dx, dy = all_pairs(xy)
    d2 = dx * dx + dy * dy
    for hv in (True, False):
        mult = 0.05 * np.sqrt(n)
        cutoff, cut_sq, one_per = orc.cutoff(n, mult, hv)
        R = orc.jacobian_pass(xy, 1e-3, 0.0, 0.0, cutoff, cut_sq, one_per, hv, "p")
        inside = d2 < (1 if hv else 49) * cut_sq
        assert R["nnz"] == n + 2 * inside.sum()
        A = _dense_from_csc(R, n)
        k = np.flatnonzero(inside & (d2 > cut_sq))
        if not hv and k.size:
            kk = k[0]
            e, c = orc.inputs(dx[kk:kk + 1], dy[kk:kk + 1])
            raw = 1e-3 * bsign(n)[j[kk]] * bsign(n)[i[kk]] * orc.field(dx[kk:kk + 1], dy[kk:kk + 1], e, c)[1][0]
            raw *= np.exp(-(np.sqrt(d2[kk]) - cutoff) ** 2 * one_per)
            d = R["dvec"]
            assert A[j[kk], i[kk]] == raw * ((1 + d[j[kk]]) / 2)


def test_rescale_equals_the_half_step_pass(orc):
    """simulation.cpp:546-608 against what the mathematics fixes: V = h b b dsigma and s = -sum V are
    linear in h, so a pass at h rescaled with factor 1/2 (d' = w(w^-1(d)/2), entries scaled by
    (1 + d'_row)/(1 + d_row), speeds moved from the stress at t1 to the one at t1/2) must be the pass at
    h/2 itself, entry for entry, to rounding -- for every weight function and cutoff mode."""
    h, ta, t1, th = 2e-3, 0.01, 0.05, 0.03
    for wf in "cpm":
        for n, mult, hv in ((40, 1e20, False), (300, 0.05 * np.sqrt(300), False), (300, 0.05 * np.sqrt(300), True)):
            xy = random_config(n, 50 + n)
            cut = orc.cutoff(n, mult, hv)
            R = orc.jacobian_pass(xy, h, ta, t1, *cut, hv, wf)
            H = orc.jacobian_pass(xy, h / 2, ta, th, *cut, hv, wf)
            assert (R["dvec"] > 0).sum() > n // 4  # the weights are active on many rows
            Ax, dv, sp = R["Ax"].copy(), R["dvec"].copy(), R["vB"].copy()
            orc.jacobian_from_prev(R["Ap"], R["Ai"], Ax, dv, wf, 0.0, 0.5, 1.0, t1, th, sp)
            assert np.array_equal(R["Ap"], H["Ap"]) and np.array_equal(R["Ai"], H["Ai"])
            assert np.all(np.abs(Ax - H["Ax"]) <= 2e-15 * np.abs(H["Ax"])), (wf, n, hv)
            assert np.all(np.abs(dv - H["dvec"]) <= 1e-13 * H["dvec"]), (wf, n, hv)
            assert np.all(np.abs(sp - H["vB"]) <= 1e-14 * (np.abs(H["vB"]) + 1)), (wf, n, hv)


def test_weights_are_the_reference_weight_of_the_diagonal(orc, ref):
    """d_i = w(s_i) (simulation.cpp:500-510) with s_i the unscaled diagonal of M = V + diag(s), through
    the reference's own sim_utils.cpp weight (oracle/_ref), bit for bit; and s_i = -sum_j V_ij."""
    n = 200
    xy = random_config(n, 61)
    for wf in "cpm":
        vA, vB, M, d = orc.jacobian_dense(xy, 3e-3, 0.0, 0.1, wf)
        s = np.diag(M).copy()
        w = np.array([ref.L.ref_weight(x, wf.encode()) for x in s])
        assert (w > 0).sum() > n // 4 and (w == 0).sum() > 0
        assert np.array_equal(d, w)
        off = M - np.diag(s)
        assert np.all(np.abs(s + off.sum(axis=1)) <= 1e-13 * np.abs(off).sum(axis=1))
        cut = orc.cutoff(n, 1e20, False)
        R = orc.jacobian_pass(xy, 3e-3, 0.0, 0.1, *cut, False, wf)
        assert np.array_equal(R["dvec"], w)


# ------------------------------------------------------------------ the run loop (simulation.cpp:88-276)

def _convergence(o, tau0=0.3, rate=40.0, T=0.02, levels=(10, 20, 40, 80)):
    """Accept-all fixed steps (huge -P, -p 0, -M = the step) of the loop at h = T/k against a DOP853
    integration (rtol = atol = 1e-13) of dx_i/dt = v_i(x; tau(t)) built on the pinned speed pass, from a
    well-separated 16-dislocation configuration under a stress that grows in time (-r rate -s tau0)."""
    from scipy.integrate import solve_ivp
    from helpers import sort_reference
    rng = np.random.default_rng(3)
    g = np.arange(4) * 0.25 - 0.375
    X, Y = np.meshgrid(g, g)
    xy = sort_reference(np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.03, 0.03, (16, 2)))
    ode = oracle_mod().Oracle()
    sol = solve_ivp(lambda t, x: ode.speed_pass(np.stack([x, xy[:, 1]], 1), tau0 + rate * t), (0, T), xy[:, 0],
                    method="DOP853", rtol=1e-13, atol=1e-13)
    xe = sol.y[:, -1]
    errs = []
    for k in levels:
        h = T / k
        out, st = o.run(xy, dt0=h, max_dt=h, prec=1e10, dipole_prec=0.0, step_limit=k - 1, stress_kind=1,
                        stress_init=tau0, stress_rate=rate)
        assert st.ok_steps == k and st.failed_steps == 0
        errs.append(np.abs(wrap(out[:, 0] - xe)).max())
    return np.array(errs)


def oracle_mod():
    import oracle
    return oracle


def test_run_loop_converges_at_second_order(orc):
    """The theta-step (theta_i = (1 + d_i)/2, d_i = w(s_i) = O(h^2) on these non-stiff rows) with step
    doubling is a second-order method: halving h divides the end-point error by 4. A wrong theta weight,
    a wrong stage time or a wrong sign in g makes it first order or no method at all (next test)."""
    e = _convergence(orc)
    ratios = e[:-1] / e[1:]
    print("errors", e, "ratios", ratios)
    assert np.all((ratios > 3.8) & (ratios < 4.2)) and e[-1] < 1e-6


@pytest.mark.parametrize("mutant", [1, 2, 3])
def test_run_loop_check_catches_planted_mistakes(orc, tmp_path, mutant):
    """The same check against copies of the oracle with a planted mistake (ORC_MUTANT in d4_oracle.c):
    theta = 1, a stage-III speed at the wrong time, the sign of g. Each one loses the second order."""
    import subprocess
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "d4_oracle.c")
    lib = str(tmp_path / f"liboracle_m{mutant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", f"-DORC_MUTANT={mutant}", "-o", lib, src, "-lm"],
                   check=True)
    e = _convergence(oracle_mod().Oracle(path=lib))
    ratios = e[:-1] / e[1:]
    good = _convergence(orc)
    print("mutant", mutant, "errors", e, "ratios", ratios)
    assert not np.all((ratios > 3.8) & (ratios < 4.2)) and e[-1] > 10 * good[-1]


def test_c1_run_is_chaotic_under_the_reference_flags(orc):
    """C1 (1000_64.dconf, 100 accept-all steps of 1e-3) with the oracle at -O2 against the same source at
    the reference's flags (-Ofast -ffast-math, CMakeLists.txt:6): the two agree to rounding after one step
    and separate beyond 1e-9 within a few steps, so no implementation that rounds differently from the
    reference's build can meet 1e-9 over the whole run; the GPU test of C1 is held to 1e-9 for as long
    as these two builds stay within it."""
    fast = oracle_mod().Oracle(fast=True)
    xy0 = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dconf_golden.npz"))["xy"]
    devs = {}
    for k in (1, 3, 5, 10, 100):
        kw = dict(dt0=1e-3, max_dt=1e-3, prec=1e10, dipole_prec=0.0, step_limit=k - 1)
        a, _ = orc.run(xy0, **kw)
        b, _ = fast.run(xy0, **kw)
        devs[k] = np.abs(wrap(a[:, 0] - b[:, 0])).max()
    print("O2 vs reference flags, max |dx| after k steps:", devs)
    assert devs[1] <= 1e-12
    assert devs[10] > 1e-6 and devs[100] > 1e-3


def test_run_invariants(orc, golden):
    """A fixed-step run (every step accepted): strain equals sum_i b_i (x_end - x_0) unwrapped,
    y is untouched, and positions stay normalised (simulation.cpp:163-175, 713-720)."""
    xy0 = golden["dconf"]["xy"]
    xy, st = orc.run(xy0, dt0=1e-3, max_dt=1e-3, prec=1e10, dipole_prec=0.0, step_limit=9)
    assert st.ok_steps == 10 and st.failed_steps == 0
    assert abs(st.sim_time - 10e-3) < 1e-15
    assert np.array_equal(xy[:, 1], xy0[:, 1])
    assert np.all((xy[:, 0] >= -0.5) & (xy[:, 0] < 0.5))
    disp = wrap(xy[:, 0] - xy0[:, 0])  # displacement modulo whole cells
    gap = st.strain - (bsign(64) * disp).sum()
    assert abs(gap - round(gap)) < 1e-12
    xy, st = orc.run(xy0, dt0=1e-6, max_dt=1e-6, prec=1e10, dipole_prec=0.0, step_limit=4)
    disp = wrap(xy[:, 0] - xy0[:, 0])
    assert np.abs(disp).max() < 0.4 and abs(st.strain - (bsign(64) * disp).sum()) < 1e-13



def test_c5_fixed_steps_from_uniform_positions_are_not_differentiable(orc):
    """C5 as stated (N = 2048 uniform, tau = 0.2, fixed dt = 1e-3: -P 1e10 -p 0 -M 1e-3 --initial-stepsize 1e-3)
    through the oracle's run loop (simulation.cpp:88-276): the very first step moves dislocations by up to
    half the cell and accumulates a strain of thousands of cell lengths, and a 1e-9 change of one coordinate
    changes the end positions by order one.  A finite-difference check of a strain gradient through 256 such
    steps has nothing to compare with; the adjoint's C5 test takes the reference's own accepted step sizes."""
    xy0 = c5_config()
    kw = dict(dt0=1e-3, max_dt=1e-3, prec=1e10, dipole_prec=0.0, step_limit=0, stress_kind=1, stress_init=0.2,
              stress_rate=0.0)
    a, st = orc.run(xy0, **kw)
    xp = xy0.copy()
    xp[7, 0] += 1e-9
    b, stp = orc.run(xp, **kw)
    step = np.abs(wrap(a[:, 0] - xy0[:, 0])).max()
    dev = np.abs(wrap(b[:, 0] - a[:, 0])).max()
    slope = (stp.strain - st.strain) / 1e-9
    print(f"one step of 1e-3: strain {st.strain:.4e}, largest move {step:.4f}, "
          f"x_7 + 1e-9 moves the end positions by {dev:.4f}, strain slope {slope:.3e}")
    assert st.ok_steps == 1
    assert step > 0.4 and abs(st.strain) > 1e3
    assert dev > 0.1 and abs(slope) > 1e10
