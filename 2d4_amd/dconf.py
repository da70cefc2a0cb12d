"""The .dconf text format, the configuration generator and the comparator.

  read_dconf / write_dconf   x y b per line (simulation_data.cpp:76-113, 177-191;
                             init_config_gen.cpp:312-320 writes 17 significant digits)
  generate                   uniform configurations as init_config_gen.cpp:239-263 draws
                             them (x then y per dislocation, b alternating from -1)
  MT19937, MT19937_64        the Mersenne Twister engines (Matsumoto & Nishimura 1998,
                             Nishimura 2000) with the seeding the C++ standard library uses
  sort_order                 the reference's sort: y + b descending (simulation_data.cpp:116)
  conf_compare               conf_compare/conf_compare.cpp:180-279 (max / average
                             x difference after normalize, y check)
"""
import numpy as np


def read_dconf(path):
    """Returns an (n, 3) array of x, y, b in file order."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if len(parts) < 3:
                raise ValueError(f"{path}: a line needs x, y and b: {line!r}")
            rows.append([float(parts[0]), float(parts[1]), float(parts[2])])
    return np.array(rows, dtype=np.float64).reshape(-1, 3)


def write_dconf(path, xyb, digits=17):
    """One dislocation per line, tab separated, `digits` significant digits
    (init_config_gen.cpp:318 uses max_digits10 = 17; writeDislocationDataToFile
    uses 14, simulation_data.cpp:181)."""
    with open(path, "w") as f:
        for x, y, b in np.asarray(xyb, dtype=np.float64).reshape(-1, 3):
            f.write(f"{x:.{digits}g}\t{y:.{digits}g}\t{int(b)}\n")


class _MT:
    """Mersenne Twister state machine for word size w (32 or 64), written from
    the published recurrences; seeding x_0 = seed, x_i = f (x_{i-1} xor
    (x_{i-1} >> (w-2))) + i as std::mersenne_twister_engine seeds."""

    def __init__(self, w, n, m, r, a, u, d, s, b, t, c, l, f, seed):
        self.w, self.n, self.m = w, n, m
        self.mask = (1 << w) - 1
        self.lower = (1 << r) - 1
        self.upper = self.mask & ~self.lower
        self.a, self.u, self.d, self.s, self.b, self.t, self.c, self.l = a, u, d, s, b, t, c, l
        mt = [seed & self.mask]
        for i in range(1, n):
            mt.append((f * (mt[-1] ^ (mt[-1] >> (w - 2))) + i) & self.mask)
        self.mt, self.i = mt, n

    def _twist(self):
        mt, n, m = self.mt, self.n, self.m
        for k in range(n):
            y = (mt[k] & self.upper) | (mt[(k + 1) % n] & self.lower)
            mt[k] = mt[(k + m) % n] ^ (y >> 1) ^ (self.a if y & 1 else 0)
        self.i = 0

    def __call__(self):
        if self.i >= self.n:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> self.u) & self.d
        y ^= (y << self.s) & self.b
        y ^= (y << self.t) & self.c
        y ^= y >> self.l
        return y & self.mask


def MT19937(seed=5489):
    return _MT(32, 624, 397, 31, 0x9908B0DF, 11, 0xFFFFFFFF, 7, 0x9D2C5680, 15, 0xEFC60000, 18, 1812433253, seed)


def MT19937_64(seed=5489):
    return _MT(64, 312, 156, 31, 0xB5026F5AA96619E9, 29, 0x5555555555555555, 17, 0x71D67FFFEDA60000, 37,
               0xFFF7EEE000000000, 43, 6364136223846793005, seed)


def _uniform(engine, bits):
    """A draw of uniform_real_distribution(-0.5, 0.5) on the engine: u / 2^bits + (-0.5),
    redrawn while it rounds to 0.5 (the result must stay below the upper bound)."""
    while True:
        v = float(engine()) / float(1 << bits) * 1.0 + (-0.5)
        if v < 0.5:
            return v


def generate(seed, n, engine="mt19937_64"):
    """(n, 3) x, y, b in the order init_config_gen.cpp:239-263 draws them: per
    dislocation x then y, b = (k % 2) * 2 - 1 (starting at -1).  The reference
    tool uses mt19937_64 now; the fixture 1000_64.dconf was made with the 32-bit
    engine, which engine="mt19937" reproduces."""
    if n % 2:
        raise ValueError("n must be even (half b = +1, half b = -1)")
    eng, bits = (MT19937(seed), 32) if engine == "mt19937" else (MT19937_64(seed), 64)
    out = np.empty((n, 3))
    for k in range(n):
        out[k, 0] = _uniform(eng, bits)
        out[k, 1] = _uniform(eng, bits)
        out[k, 2] = (k % 2) * 2 - 1
    return out


def sort_order(y, b):
    """Stable order by y + b descending (simulation_data.cpp:116, std::sort
    there; ties among distinct dislocations have probability zero)."""
    key = np.asarray(y, dtype=np.float64) + np.asarray(b, dtype=np.float64)
    return np.argsort(-key, kind="stable")


def normalize(v):
    """sim_utils.cpp:6-13 / conf_compare.cpp:161-168 for an array: v - k with
    k integral so that the result lies in [-0.5, 0.5)."""
    v = np.asarray(v, dtype=np.float64)
    r = v - np.rint(v)
    return np.where(r >= 0.5, r - 1.0, r)


def conf_compare(a, b, ind_tol=1e-8, sort=False):
    """conf_compare.cpp:180-279 on two (n, 3) x, y, b arrays (same order unless sort).

    As the reference does, the absolute x difference is normalised after
    taking fabs, so a difference in [0.5, 1) enters as |dx| - 1 (negative;
    it does not raise the maximum).  Returns a dict with max_diff, max_id,
    max_id_y, avg_fabs, avg_fabs_sq; raises ValueError where the reference
    reports an error (sizes differ, a y differs by more than ind_tol)."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 3)
    if a.shape != b.shape:
        raise ValueError("the number of dislocations are not equal")
    if sort:
        a = a[sort_order(a[:, 1], a[:, 2])]
        b = b[sort_order(b[:, 1], b[:, 2])]
    n = a.shape[0]
    ydiff = normalize(np.abs(a[:, 1] - b[:, 1]))
    bad = np.nonzero(ydiff > ind_tol)[0]
    if bad.size:
        raise ValueError(f"y values are different for dislocation with id = {bad[0]}")
    d = normalize(np.abs(a[:, 0] - b[:, 0]))
    max_diff, max_id = 0.0, 0
    for i in range(n):  # strict '>' from 0: the first maximum wins, negatives never count
        if d[i] > max_diff:
            max_diff, max_id = float(d[i]), i
    return {"max_diff": max_diff, "max_id": max_id, "max_id_y": float(a[max_id, 1]) if n else 0.0,
            "avg_fabs": float(d.sum() / n) if n else float("nan"),
            "avg_fabs_sq": float((d * d).sum() / n) if n else float("nan")}


def max_periodic_dx(xy_a, xy_b):
    """max_i |normalize(x_a - x_b)|: the periodic x deviation between two
    configurations in the same order (the quantity conf_compare's max_diff
    reports when no difference wraps past half a cell)."""
    d = normalize(np.asarray(xy_a)[:, 0] - np.asarray(xy_b)[:, 0])
    return float(np.max(np.abs(d))) if d.size else 0.0
