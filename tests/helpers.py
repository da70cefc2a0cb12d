"""Shared test helpers: synthetic configurations and the per-pair condition
scales used by the parity bounds.

Parity convention for floating-point results (the reference's field cancels
between image walls, and inside the central wall for near pairs, so a pure
relative bound is not attainable with any other evaluation order):
  (A) field values from the same transcendental inputs E, c:
      |got - ref| <= 1e-12 * max(|ref|, S), S = 2 pi^2 sum_k |w_k f_k| (sigma) or the
      matching sum for d sigma -- the magnitude the sum cancels from;
  (B) with the device's own exp2/cos: (A) plus 4x the reference's own change
      under a one-ulp change of E and of c (its conditioning at that pair).
"""
import numpy as np

TOL = 1e-12
PI2LN2INV = 9.0647202836543876192553658914333336203437229354476
CH = {1: 267.746761483748222245931879901, 2: 143375.656570070321051450310133,
      3: 7.67764676977233502175191858009e7, 4: 4.11131577927974976374895337541e10,
      5: 2.20157529303160145057002722946e13}
SH = {1: 267.744894041016514257117449688, 2: 143375.656566582978695241314640}


def bsign(n):
    return np.where(np.arange(n) < n // 2, 1.0, -1.0)


def sort_reference(xy):
    """SimulationData's order (simulation_data.cpp:116): y + b descending, b = +1 for the first half drawn."""
    n = xy.shape[0]
    b = np.where(np.arange(n) % 2 == 1, 1.0, -1.0)  # alternating, as init_config_gen writes them
    key = xy[:, 1] + b
    order = np.argsort(-key, kind="stable")
    return np.ascontiguousarray(xy[order])


def random_config(n, seed):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-0.5, 0.5, (n, 2))
    return sort_reference(xy)


def c5_config(n=2048, seed=2048):
    """C5's start: x, then y, uniform in the unit cell from numpy's generator with this seed, b alternating
    (-1 first), in SimulationData's sorted order."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, n)
    y = rng.uniform(-0.5, 0.5, n)
    return sort_reference(np.stack([x, y], axis=1))


def _walls(dx, E, c):
    """(u, C, S_quirk, weight) of the ten walls (Field.cpp:295-331 layout)."""
    R = 1.0 / E
    P, M = E + R, E - R
    C = {0: P / 2, 1: (P * CH[1] + M * SH[1]) / 2, -1: (P * CH[1] - M * SH[1]) / 2,
         2: (P * CH[2] + M * SH[2]) / 2, -2: (P * CH[2] - M * SH[2]) / 2}
    S = {0: M / 2, 1: (M * CH[1] + P * SH[1]) / 2, -1: (M * CH[1] - P * SH[1]) / 2,
         2: (M * CH[2] + P * SH[2]) / 2, -2: (M * CH[2] - P * SH[2]) / 2}
    for k in (3, 4, 5):
        C[k] = CH[k] * E
        C[-k] = CH[k] * R
        S[k], S[-k] = C[k], C[-k]
    pos = dx >= 0
    sg = np.where(pos, 1.0, -1.0)
    ad = np.abs(dx)
    out = [(dx + k, C[k], S[k], 1.0) for k in range(-3, 4)]
    out.append((dx - 4 * sg, np.where(pos, C[-4], C[4]), np.where(pos, S[-4], S[4]), 1.0))
    out.append((dx + 4 * sg, np.where(pos, C[4], C[-4]), np.where(pos, S[4], S[-4]), 1 - ad))
    out.append((dx - 5 * sg, np.where(pos, C[-5], C[5]), np.where(pos, S[-5], S[5]), ad))
    return out


def scale_sigma(dx, dy, E, c):
    """S = 2 pi^2 sum |w f| over the walls."""
    tot = 0.0
    for u, C, _, w in _walls(dx, E, c):
        tot = tot + np.abs(w * u * (C * c - 1) / (C - c) ** 2)
    return 2 * np.pi ** 2 * tot


def scale_dsigma(dx, dy, E, c):
    """The matching magnitude for d sigma: every additive piece of every wall's derivative."""
    tot = 0.0
    for k, (u, C, S, w) in enumerate(_walls(dx, E, c)):
        D = C - c
        N = C * c - 1
        a = 2 * np.pi * u
        tot = tot + np.abs(w) * (np.abs(N / D ** 2) + np.abs(a * S * c / D ** 2) + np.abs(2 * a * N * S / D ** 3))
        if k >= 8:  # blend-weight derivative terms f(dx+4s), f(dx-5s)
            tot = tot + np.abs(u * N / D ** 2)
    return 2 * np.pi ** 2 * tot


def wrap(v):
    """normalize() of sim_utils.cpp:6-13 in closed form: v - rint(v), with +0.5 -> -0.5."""
    v = np.asarray(v, dtype=np.float64)
    r = v - np.rint(v)
    return np.where(r >= 0.5, r - 1.0, r)


def all_pairs(xy):
    """(i, j, dx, dy) for every unordered pair i < j, oriented as simulation.cpp:446-450."""
    n = xy.shape[0]
    i, j = np.triu_indices(n, 1)
    dx = wrap(xy[i, 0] - xy[j, 0])
    dy = wrap(xy[i, 1] - xy[j, 1])
    return i, j, dx, dy


def pair_allowance(orc, dx, dy, E, c, deriv=False):
    """Bound (B) per pair: 1e-12 * max(|ref|, S) + 4 x the oracle's one-ulp input sensitivity."""
    s0, d0 = orc.field(dx, dy, E, c)
    sE, dE = orc.field(dx, dy, np.nextafter(E, np.inf), c)
    sc, dc = orc.field(dx, dy, E, np.nextafter(c, np.inf))
    if deriv:
        base = np.maximum(np.abs(d0), scale_dsigma(dx, dy, E, c))
        return TOL * base + 4 * (np.abs(dE - d0) + np.abs(dc - d0)), d0
    base = np.maximum(np.abs(s0), scale_sigma(dx, dy, E, c))
    return TOL * base + 4 * (np.abs(sE - s0) + np.abs(sc - s0)), s0




def _row_sum(n, i, j, vi, vj):
    out = np.zeros(n)
    np.add.at(out, i, vi)
    np.add.at(out, j, vj)
    return out


def speed_allowance(orc, xy):
    """Per-row bound for a speed pass: the sum of the row's per-pair bounds (B) plus 1e-13 of its
    absolute pair sum (a different, fixed summation order)."""
    n = xy.shape[0]
    i, j, dx, dy = all_pairs(xy)
    e, c = orc.inputs(dx, dy)
    a, s = pair_allowance(orc, dx, dy, e, c)
    return _row_sum(n, i, j, a, a) + 1e-13 * _row_sum(n, i, j, np.abs(s), np.abs(s))


def jacobian_allowance(orc, xy, h, R):
    """Per-entry bound for a CSC Jacobian pass R (the oracle's) from the pair bounds on d sigma, the
    weights' and the diagonal's sums over the column (the bound test_gpu_passes applies)."""
    n = xy.shape[0]
    i, j, dx, dy = all_pairs(xy)
    e, c = orc.inputs(dx, dy)
    ad, _ = pair_allowance(orc, dx, dy, e, c, deriv=True)
    aV = np.zeros((n, n))
    aV[i, j] = aV[j, i] = h * ad
    col_allow = aV.sum(axis=0)
    d = R["dvec"]
    rows = R["Ai"]
    cols = np.repeat(np.arange(n), np.diff(R["Ap"]))
    Vabs = np.abs(R["Ax"])
    bound = (1 + d[rows]) / 2 * aV[rows, cols] + Vabs * 0.5 * col_allow[rows] + 1e-13 * Vabs
    diag = rows == cols
    bound[diag] = col_allow[cols[diag]] + np.abs(R["Ax"][diag]) * (1e-13 + 0.5 * col_allow[cols[diag]])
    return bound, 0.5 * col_allow + 1e-13
