"""GPU parity of lib2d4 (through the C ABI) against the oracle.

Bounds (helpers.py): (A) per pair 1e-12 * max(|ref|, S) on shared inputs;
(B) with the device's own exp2/cos, (A) + 4x the reference's one-ulp input
sensitivity.  Pass-level results are bounded by the sum of their pairs'
bounds plus 1e-13 of the absolute sum (different, fixed summation order).
Integer results (CSC structure, counts, tolerance minima) are bit-exact.
"""
import os

import numpy as np
import pytest

from helpers import (TOL, all_pairs, bsign, pair_allowance, random_config, scale_dsigma, scale_sigma,
                     sort_reference, speed_allowance, wrap)

pytestmark = pytest.mark.gpu
GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _dconf64():
    return np.load(os.path.join(GOLDEN_DIR, "dconf_golden.npz"))["xy"]


# ------------------------------------------------------------------ field

def test_field_eval_same_inputs(ctx, orc, golden):
    """(A): device field from the reference's own E, c (golden fixture inputs)."""
    g = golden["field"]
    ok = np.hypot(g["dx"], g["dy"]) > 1e-200
    dx, dy, e, c = g["dx"][ok], g["dy"][ok], g["e"][ok], g["c"][ok]
    print(f"[field A] n={dx.size}", flush=True)
    for exact in (False, True):
        s, d = ctx.field_eval(dx, dy, e, c, exact=exact)
        s0, d0 = orc.field(dx, dy, e, c, exact=exact)
        assert np.all(np.abs(s - s0) <= TOL * np.maximum(np.abs(s0), scale_sigma(dx, dy, e, c)))
        assert np.all(np.abs(d - d0) <= TOL * np.maximum(np.abs(d0), scale_dsigma(dx, dy, e, c)))
    # against the reference's compiled Field directly (golden values)
    s, d = ctx.field_eval(dx, dy, e, c)
    assert np.all(np.abs(s - g["xy4"][ok]) <= TOL * np.maximum(np.abs(s), scale_sigma(dx, dy, e, c)))
    assert np.all(np.abs(d - g["dxy4"][ok]) <= TOL * np.maximum(np.abs(d), scale_dsigma(dx, dy, e, c)))


def test_field_eval_device_inputs(ctx, orc):
    """(B): the device evaluates exp2 and cos itself; report the pure-relative pass fraction."""
    rng = np.random.default_rng(31)
    dx = np.concatenate([rng.uniform(-0.5, 0.5, 200000), rng.uniform(-2e-3, 2e-3, 4000)])
    dy = np.concatenate([rng.uniform(-0.5, 0.5, 200000), rng.uniform(-2e-3, 2e-3, 4000)])
    print(f"[field B] n={dx.size}", flush=True)
    s, d = ctx.field_eval(dx, dy)
    e, c = orc.inputs(dx, dy)
    a, s0 = pair_allowance(orc, dx, dy, e, c)
    ad, d0 = pair_allowance(orc, dx, dy, e, c, deriv=True)
    assert np.all(np.abs(s - s0) <= a)
    assert np.all(np.abs(d - d0) <= ad)
    rel = np.abs(s - s0) / np.abs(s0)
    print(f"[field B] sigma pure-relative <= 1e-12 for {np.mean(rel <= 1e-12):.5f} of pairs, max {rel.max():.2e}")


# ------------------------------------------------------------------ speed pass

@pytest.mark.parametrize("n", [2, 3, 64, 100, 1000, 1024])
def test_speed_pass_vs_oracle(ctx, orc, n):
    xy = _dconf64() if n == 64 else random_config(n, 100 + n)
    for tau in (0.0, 0.1):
        print(f"[speed] n={n} tau={tau}", flush=True)
        v = ctx.speed_pass(xy, tau)
        v0 = orc.speed_pass(xy, tau)
        assert np.all(np.abs(v - v0) <= speed_allowance(orc, xy) + 1e-15 * abs(tau))


def test_speed_pass_tolerance_exact(ctx, orc):
    """PrecisionHandler::updateTolerance minima are bit-exact (same d^2 rounding)."""
    xy = random_config(1000, 7)
    cap2 = (1e-5 / 0.05) ** 2 * 400
    t_gpu = np.full(1000, cap2)
    t_orc = t_gpu.copy()
    ctx.speed_pass(xy, 0.0, t_gpu, cap2)
    orc.speed_pass(xy, 0.0, t_orc, cap2)
    assert (t_orc < cap2).sum() > 0
    assert np.array_equal(t_gpu, t_orc)


def test_speed_pass_deterministic(ctx):
    xy = random_config(3000, 8)
    a = ctx.speed_pass(xy, 0.05)
    b = ctx.speed_pass(xy, 0.05)
    assert np.array_equal(a, b)


def test_single_dislocation(pkg, ctx, orc):
    """n = 1: no pair; its b is -1 (the first n/2 = 0 have b = +1, simulation_data.cpp:265-270), so its speed is
    -tau, its tolerance is untouched, and the Jacobian is the 1 x 1 identity with d = 0."""
    xy = np.array([[0.1, -0.2]])
    tol = np.array([0.5])
    v = ctx.speed_pass(xy, 0.3, tol, 1.0)
    assert np.array_equal(v, orc.speed_pass(xy, 0.3)) and v[0] == -0.3 and tol[0] == 0.5
    p = pkg.jac_params(h=1e-3, tau_a=0.1, tau_b=0.2, n=1)
    G = ctx.jacobian_pass(xy, p)
    assert G["nnz"] == 1 and list(G["Ap"]) == [0, 1] and list(G["Ai"]) == [0] and G["Ax"][0] == 1.0
    assert G["vA"][0] == -0.1 and G["vB"][0] == -0.2 and G["dvec"][0] == 0.0


def test_coincident_pair_is_nan_as_in_the_reference(ctx, orc):
    """Two dislocations at the same point: the reference's near-field series divides 0 by 0 (Field.cpp:176-188),
    so both their speeds are NaN; the other rows stay finite and within the bound, and the tolerance of the two
    is d^2 = 0, as the oracle gives."""
    xy = random_config(40, 21)
    xy[7] = xy[30]
    n = xy.shape[0]
    cap2 = 1e-2
    t_gpu, t_orc = np.full(n, cap2), np.full(n, cap2)
    v = ctx.speed_pass(xy, 0.05, t_gpu, cap2)
    v0 = orc.speed_pass(xy, 0.05, t_orc, cap2)
    assert np.array_equal(np.isnan(v), np.isnan(v0)) and set(np.flatnonzero(np.isnan(v))) == {7, 30}
    assert np.array_equal(t_gpu, t_orc) and t_gpu[7] == 0.0 and t_gpu[30] == 0.0
    ok = ~np.isnan(v0)
    assert np.all(np.isfinite(v[ok]))
    with np.errstate(invalid="ignore"):
        allow = speed_allowance(orc, xy)
    assert np.all(np.abs(v[ok] - v0[ok]) <= allow[ok] + 1e-15)


def test_speed_pass_at_n65536(ctx):
    """The largest speed pass of the suite (65 536 dislocations, 2.1e9 pairs in 524 800 tiles): the pair sums
    cancel (sum_i v_i = tau sum_i b_i = 0), a permutation-free property that holds at any size, and the pass is
    bitwise repeatable."""
    import torch
    n = 65536
    xy = random_config(n, 22)
    t = torch.from_numpy(xy).to("cuda:0")
    print(f"[max] speed pass n={n}", flush=True)
    v = ctx.speed_pass_t(t, tau=0.1)
    w = ctx.speed_pass_t(t, tau=0.1)
    torch.cuda.synchronize()
    assert torch.equal(v, w) and bool(torch.isfinite(v).all())
    assert abs(v.sum().item()) <= 1e-12 * v.abs().sum().item()
    b = torch.from_numpy(bsign(n)).to("cuda:0")
    assert abs((v - 0.1 * b).sum().item()) <= 1e-12 * v.abs().sum().item()


# ------------------------------------------------------------------ Jacobian (CSC)

def _jac_case(name):
    if name == "dense64":
        xy = _dconf64()
        return xy, dict(h=1e-3, ta=0.0, tb=0.0, mult=1e20, hv=False, wf="c")
    if name == "heaviside1024":
        xy = random_config(1024, 9)
        return xy, dict(h=1e-4, ta=0.1, tb=0.1, mult=0.05 * 32, hv=True, wf="c")
    if name == "taper300":
        xy = random_config(300, 10)
        return xy, dict(h=2e-4, ta=0.02, tb=0.05, mult=0.05 * np.sqrt(300), hv=False, wf="m")
    if name == "pair2":  # the smallest configuration: one pair, one dislocation of each sign
        xy = sort_reference(np.array([[0.1, 0.2], [-0.3, -0.1]]))
        return xy, dict(h=1e-3, ta=0.05, tb=0.07, mult=1e20, hv=False, wf="p")
    if name == "nopairs16":  # a cutoff no pair is inside: the matrix is its diagonal alone
        xy = random_config(16, 12)
        return xy, dict(h=1e-3, ta=0.0, tb=0.0, mult=1e-4 * 4, hv=True, wf="c")
    if name == "close8":
        # a dipole 5e-4 apart (dx^2 + dy^2 < 1e-6, the near-coincident branch of Field.cpp:176-188)
        # and a pair straddling the periodic boundary in x
        rng = np.random.default_rng(13)
        xy = rng.uniform(-0.4, 0.4, (8, 2))
        xy[2] = xy[1] + (3e-4, 4e-4)
        xy[3] = (0.4999, 0.05)
        xy[4] = (-0.4998, 0.0502)
        return sort_reference(xy), dict(h=1e-4, ta=0.01, tb=0.01, mult=0.3 * np.sqrt(8), hv=True, wf="c")
    raise KeyError(name)


@pytest.mark.parametrize("case", ["dense64", "heaviside1024", "taper300", "pair2", "nopairs16", "close8"])
def test_jacobian_pass_vs_oracle(pkg, ctx, orc, case):
    xy, k = _jac_case(case)
    n = xy.shape[0]
    print(f"[jacobian] {case} n={n}", flush=True)
    cutoff, cut_sq, one_per = orc.cutoff(n, k["mult"], k["hv"])
    p = pkg.jac_params(h=k["h"], tau_a=k["ta"], tau_b=k["tb"], cutoff=cutoff, heaviside=k["hv"], wfunc=k["wf"])
    assert p.cut_sq == cut_sq
    G = ctx.jacobian_pass(xy, p)
    R = orc.jacobian_pass(xy, k["h"], k["ta"], k["tb"], cutoff, cut_sq, one_per, k["hv"], k["wf"])
    # structure: bit-exact
    assert G["nnz"] == R["nnz"]
    if case == "nopairs16":
        assert G["nnz"] == n
    if case == "close8":
        assert G["nnz"] > n + 2  # the dipole and the boundary pair are both inside the cutoff
    assert np.array_equal(G["Ap"], R["Ap"]) and np.array_equal(G["Ai"], R["Ai"])
    assert np.array_equal(G["indexes"], R["indexes"])
    # values: per-entry bound from the pair bounds on d sigma
    i, j, dx, dy = all_pairs(xy)
    e, c = orc.inputs(dx, dy)
    ad, _ = pair_allowance(orc, dx, dy, e, c, deriv=True)
    aV = np.zeros((n, n))
    aV[i, j] = aV[j, i] = k["h"] * ad
    col_allow = aV.sum(axis=0)
    d = R["dvec"]
    rows = R["Ai"]
    cols = np.repeat(np.arange(n), np.diff(R["Ap"]))
    Vabs = np.abs(R["Ax"])
    bound = (1 + d[rows]) / 2 * aV[rows, cols] + Vabs * 0.5 * col_allow[rows] + 1e-13 * Vabs
    diag = rows == cols
    bound[diag] = col_allow[cols[diag]] + np.abs(R["Ax"][diag]) * (1e-13 + 0.5 * col_allow[cols[diag]])
    assert np.all(np.abs(G["Ax"] - R["Ax"]) <= bound)
    assert np.all(np.abs(G["dvec"] - d) <= 0.5 * col_allow + 1e-13)
    allow = speed_allowance(orc, xy)
    assert np.all(np.abs(G["vA"] - R["vA"]) <= allow + 1e-15)
    assert np.all(np.abs(G["vB"] - R["vB"]) <= allow + 1e-15)


def test_jacobian_pass_list_overflow_on_a_fresh_context(pkg, ctx, orc):
    # a new context's pair list starts from the caller's capacity (here none):
    # the tapered 300-point case lists far more pairs, so the pass runs again
    # with a longer list, then reports the CSC size; the wrapper retries with it
    xy, k = _jac_case("taper300")
    n = xy.shape[0]
    cutoff, cut_sq, one_per = orc.cutoff(n, k["mult"], k["hv"])
    p = pkg.jac_params(h=k["h"], tau_a=k["ta"], tau_b=k["tb"], cutoff=cutoff, heaviside=k["hv"], wfunc=k["wf"])
    R = orc.jacobian_pass(xy, k["h"], k["ta"], k["tb"], cutoff, cut_sq, one_per, k["hv"], k["wf"])
    assert (R["nnz"] - n) // 2 > n // 2 + 1024  # more pairs than the first list holds
    fresh = pkg.Context(0)
    try:
        G = fresh.jacobian_pass(xy, p, cap=0)
    finally:
        fresh.close()
    assert G["nnz"] == R["nnz"]
    assert np.array_equal(G["Ap"], R["Ap"]) and np.array_equal(G["Ai"], R["Ai"])
    assert np.array_equal(G["indexes"], R["indexes"])
    # the sorted CSC does not depend on the order the list filled in: the same bits as a context whose
    # list was long enough from the start
    S = ctx.jacobian_pass(xy, p, cap=G["nnz"])
    for key in ("vA", "vB", "Ap", "Ai", "Ax", "indexes", "dvec"):
        np.testing.assert_array_equal(G[key], S[key])


def test_jacobian_pass_capacity_and_tolerance(pkg, ctx, orc):
    xy, k = _jac_case("heaviside1024")
    n = xy.shape[0]
    cutoff, cut_sq, one_per = orc.cutoff(n, k["mult"], True)
    cap2 = 4e-6
    p = pkg.jac_params(h=1e-4, cutoff=cutoff, heaviside=True, tol_cap2=cap2)
    tg = np.full(n, cap2)
    G = ctx.jacobian_pass(xy, p, tol=tg, cap=16)  # too small: the wrapper retries with the reported size
    to = np.full(n, cap2)
    R = orc.jacobian_pass(xy, 1e-4, 0, 0, cutoff, cut_sq, one_per, True, "c", tol=to, cap2=cap2)
    assert G["nnz"] == R["nnz"]
    assert np.array_equal(tg, to)


def test_jacobian_rescale_vs_oracle(pkg, ctx, orc):
    xy, k = _jac_case("taper300")
    n = xy.shape[0]
    cutoff, cut_sq, one_per = orc.cutoff(n, k["mult"], False)
    R = orc.jacobian_pass(xy, k["h"], k["ta"], k["tb"], cutoff, cut_sq, one_per, False, "m")
    Ax1, d1, sp1 = R["Ax"].copy(), R["dvec"].copy(), R["vB"].copy()
    Ax2, d2, sp2 = R["Ax"].copy(), R["dvec"].copy(), R["vB"].copy()
    orc.jacobian_from_prev(R["Ap"], R["Ai"], Ax1, d1, "m", 0.0, 0.5e-3, 1e-3, 0.05, 0.035, sp1)
    ctx.jacobian_rescale(R["Ap"], R["Ai"], Ax2, d2, "m", 0.5, 0.035 - 0.05, sp2)
    # a few ulp: the same formula per entry, weights through sqrt/divisions of the device libm
    assert np.allclose(Ax2, Ax1, rtol=2e-15, atol=2e-16)
    assert np.allclose(d2, d1, rtol=1e-14, atol=0)
    assert np.allclose(sp2, sp1, rtol=0, atol=4e-16 * (np.abs(sp1).max() + 1))


# ------------------------------------------------------------------ Jacobian (dense)

@pytest.mark.parametrize("n", [64, 300, 301])
def test_jacobian_dense_vs_oracle(pkg, ctx, orc, n):
    xy = _dconf64() if n == 64 else random_config(n, 11)
    h = 1e-3
    print(f"[dense] n={n}", flush=True)
    p = pkg.jac_params(h=h, tau_a=0.1, tau_b=0.2, n=n, wfunc="p")
    vA, vB, M, d = ctx.jacobian_dense(xy, p)
    vA0, vB0, M0, d0 = orc.jacobian_dense(xy, h, 0.1, 0.2, "p")
    assert np.array_equal(M, M.T)
    i, j, dx, dy = all_pairs(xy)
    e, c = orc.inputs(dx, dy)
    ad, _ = pair_allowance(orc, dx, dy, e, c, deriv=True)
    aV = np.zeros((n, n))
    aV[i, j] = aV[j, i] = h * ad
    np.fill_diagonal(aV, aV.sum(axis=1) + 1e-13 * np.abs(M0).sum(axis=1))
    assert np.all(np.abs(M - M0) <= aV)
    allow = speed_allowance(orc, xy)
    assert np.all(np.abs(vA - vA0) <= allow + 1e-15) and np.all(np.abs(vB - vB0) <= allow + 1e-15)
    assert np.all(np.abs(d - d0) <= np.diag(aV))
    # into the caller's buffer: column-major, the same bits
    out = np.full((n, n), np.nan)
    vA2, _, Mo, _ = ctx.jacobian_dense(xy, p, out=out)
    assert Mo is out and np.array_equal(out.T, M) and np.array_equal(vA2, vA)
    # scaled form equals the CSC pass's A (dense cutoff) to rounding
    _, _, A, _ = ctx.jacobian_dense(xy, p, scale=True)
    G = ctx.jacobian_pass(xy, p)
    Acsc = np.zeros((n, n))
    cols = np.repeat(np.arange(n), np.diff(G["Ap"]))
    Acsc[G["Ai"], cols] = G["Ax"]
    assert np.allclose(A, Acsc, rtol=1e-13, atol=1e-13 * np.abs(Acsc).max())


@pytest.mark.parametrize("case", ["heaviside1024", "taper300", "close8", "nopairs16", "dense64"])
def test_jacobian_dense_cut_vs_csc_pass(pkg, ctx, orc, case):
    """The dense pass with the reference's storage cutoff (d4_jacobian_dense_cut_dev: sigma of every pair in
    the tile steps, d sigma of the pairs inside the radius taken after each 16-step phase): its scaled matrix is
    the CSC pass's A (pinned to the oracle above) entry for entry, zero outside the stored pattern, and its
    speeds and weights are the CSC pass's."""
    import torch
    xy, k = _jac_case(case)
    n = xy.shape[0]
    print(f"[dense cut] {case} n={n}", flush=True)
    cutoff, cut_sq, one_per = orc.cutoff(n, k["mult"], k["hv"])
    p = pkg.jac_params(h=k["h"], tau_a=k["ta"], tau_b=k["tb"], cutoff=cutoff, heaviside=k["hv"], wfunc=k["wf"])
    t = torch.tensor(xy, dtype=torch.float64, device="cuda:0")
    vA, vB, At, d = ctx.jacobian_dense_t(t, p, cutoff=True, scale=True)  # At[col, row]
    A = At.cpu().numpy().T
    G = ctx.jacobian_pass(xy, p)
    Acsc = np.zeros((n, n))
    cols = np.repeat(np.arange(n), np.diff(G["Ap"]))
    Acsc[G["Ai"], cols] = G["Ax"]
    stored = np.zeros((n, n), dtype=bool)
    stored[G["Ai"], cols] = True
    assert np.all(A[~stored] == 0.0)
    assert np.allclose(A, Acsc, rtol=1e-13, atol=1e-13 * np.abs(Acsc).max())
    assert np.allclose(vA.cpu().numpy(), G["vA"], rtol=0, atol=1e-13 * np.abs(G["vA"]).max())
    assert np.allclose(vB.cpu().numpy(), G["vB"], rtol=0, atol=1e-13 * np.abs(G["vB"]).max())
    assert np.allclose(d.cpu().numpy(), G["dvec"], rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------------ full size, device-resident

def test_full_size_speed_and_dense_invariants(pkg, ctx, orc):
    """N = 16384 (the headline size): exact symmetry of the dense Jacobian, zero row sums of
    V + diag(s), cancelling pair sums of the speeds, and sampled rows against the oracle."""
    import torch
    n = 16384
    xy = random_config(n, 12)
    dev = torch.device("cuda:0")
    t = torch.from_numpy(xy).to(dev)
    print(f"[full] speed pass n={n}", flush=True)
    v = ctx.speed_pass_t(t, tau=0.0)
    torch.cuda.synchronize()
    vs = v.double()
    assert abs(vs.sum().item()) <= 1e-12 * vs.abs().sum().item()
    print(f"[full] dense jacobian n={n}", flush=True)
    p = pkg.jac_params(h=1e-6, tau_a=0.0, tau_b=0.1, n=n)
    vA, vB, Mt, d = ctx.jacobian_dense_t(t, p)  # Mt[col, row]
    torch.cuda.synchronize()
    assert torch.equal(Mt, Mt.t())
    rs = Mt.sum(dim=0)
    assert (rs.abs() <= 1e-12 * Mt.abs().sum(dim=0)).all().item()
    assert torch.equal(vA, v)
    rng = np.random.default_rng(13)
    rows = rng.choice(n, 6, replace=False)
    Mh_rows = Mt[torch.as_tensor(rows, device=dev)].cpu().numpy()  # columns of M = rows by symmetry
    vh = v.cpu().numpy()
    b = bsign(n)
    for k, r in enumerate(rows.tolist()):
        others = np.delete(np.arange(n), r)
        lo = np.minimum(others, r)
        hi = np.maximum(others, r)
        dx = wrap(xy[lo, 0] - xy[hi, 0])
        dy = wrap(xy[lo, 1] - xy[hi, 1])
        e, c = orc.inputs(dx, dy)
        a, s0 = pair_allowance(orc, dx, dy, e, c)
        ad, d0 = pair_allowance(orc, dx, dy, e, c, deriv=True)
        sgn = np.where(r == lo, 1.0, -1.0)
        f = s0 * b[r] * b[others]
        vr = (sgn * f).sum()
        assert abs(vh[r] - vr) <= a.sum() + 1e-13 * np.abs(f).sum()
        got = Mh_rows[k][others]
        want = 1e-6 * b[r] * b[others] * d0
        assert np.all(np.abs(got - want) <= 1e-6 * ad)
