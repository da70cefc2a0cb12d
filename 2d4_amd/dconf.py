"""The .dconf text format and the configuration comparator.

  read_dconf / write_dconf   x y b per line (simulation_data.cpp:76-113, 177-191;
                             init_config_gen.cpp:312-320 writes 17 significant digits)
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
