"""Writes the golden fixtures of tests/golden/ from the reference's own compiled
sources (oracle/_ref, built by `make -C oracle ref` from /root/reference).

Run in a container that has /root/reference:  python tests/golden/make_golden.py
Outputs (data only: inputs and the reference's outputs):
  field_golden.npz   per-pair Field::xy / xy_diff_x (2- and 4-arg) at -O2 and at
                     the reference flags, with the libm inputs E, c used
  utils_golden.npz   normalize / weight / weightInv / PrecisionHandler /
                     stress protocols / step-size heuristic
  dconf_golden.npz   SimulationData's sort, disl_order, b and cutoff on
                     1000_64.dconf (copied here from init_config_gen/README_files)
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

REF_DCONF = "/root/reference/init_config_gen/README_files/1000_64.dconf"


def field_points(seed=20251203):
    rng = np.random.default_rng(seed)
    dx = [rng.uniform(-0.5, 0.5, 2500)]
    dy = [rng.uniform(-0.5, 0.5, 2500)]
    # near field: series branch (r^2 <= 1e-6) and just outside it
    r = 10 ** rng.uniform(-5, -2.5, 600)
    th = rng.uniform(0, 2 * np.pi, 600)
    dx.append(r * np.cos(th))
    dy.append(r * np.sin(th))
    # edges: wrap boundary, dx = 0 (sign switch), dy = 0 and +-0.5
    e = np.nextafter(0.5, 0)
    ex = [0.0, -0.0, -0.5, e, -e, 1e-300, -1e-300, 0.25, -0.25, 0.1, 0.45, -0.3, 0.25, 1e-3, -1e-3, 0.3]
    ey = [0.3, -0.3, 0.1, 0.2, -0.2, 0.4, 0.0, -0.5, e, 0.2, -0.4, 0.05, 0.25, 0.0, 0.0, 0.1]
    dx.append(np.array(ex))
    dy.append(np.array(ey))
    return np.concatenate(dx), np.concatenate(dy)


def main():
    if not os.path.exists(REF_DCONF):
        sys.exit("the reference tree is not present; fixtures are generated only where it is")
    subprocess_build()
    r2, rf = oracle.Ref("o2"), oracle.Ref("fast")
    dx, dy = field_points()
    a = r2.field_batch(dx, dy)
    b = rf.field_batch(dx, dy)
    np.savez_compressed(os.path.join(HERE, "field_golden.npz"), dx=dx, dy=dy, e=a["e"], c=a["c"],
                        xy2=a["xy2"], dxy2=a["dxy2"], xy4=a["xy4"], dxy4=a["dxy4"],
                        e_fast=b["e"], c_fast=b["c"], xy2_fast=b["xy2"], dxy2_fast=b["dxy2"],
                        xy4_fast=b["xy4"], dxy4_fast=b["dxy4"])

    L = r2.L
    norm_in = np.array([0.0, -0.0, 0.5, -0.5, np.nextafter(0.5, 0), np.nextafter(-0.5, 0), np.nextafter(-0.5, -1),
                        1.5, -1.5, 2.49, -10.3, 7.25, 0.9999999999999999, -0.9999999999999999, 123.456,
                        -0.49999999999999994, 0.49999999999999994, 3.0, -3.0, 2.5, -2.5, 0.5000000000000001,
                        -0.5000000000000001, 1.4999999999999998, -1.5000000000000002, 1024.5, -7.499999999999999])
    norm_out = np.array([L.ref_normalize(v) for v in norm_in])
    s_in = np.concatenate([[-1.0, 0.0, 1e-300, 1e-12, 1e-3, 0.5, 1.0, 2.0, 10.0, 1e6], 10 ** np.linspace(-6, 4, 41)])
    w = {k: np.array([L.ref_weight(v, k.encode()) for v in s_in]) for k in "cpm"}
    w_in = np.concatenate([[-0.5, 0.0], np.linspace(1e-4, 0.99, 40)])
    wi = {k: np.array([L.ref_weight_inv(v, k.encode()) for v in w_in]) for k in "cpm"}
    # PrecisionHandler replay: 16 particles, 200 random updates
    rng = np.random.default_rng(5)
    ids = rng.integers(0, 16, 200).astype(np.uint32)
    d2 = 10 ** rng.uniform(-10, -3, 200)
    ratios = {}
    for prec, dip in ((1e-5, 0.05), (1e-5, 0.0), (1e-3, 0.5)):
        out = np.empty(16)
        L.ref_tolerance_replay(prec, dip, 16, 200, d2.ctypes.data, ids.ctypes.data, out.ctypes.data)
        ratios[f"{prec}_{dip}"] = out
    steps = np.array([[L.ref_new_step_size(1e-5, 0.05, 4, err, 1e-3) for err in (0.0, 1e-9, 1e-7, 1e-6, 1e-5)]])
    t_in = np.linspace(0, 3.7, 23)
    stress = {f"k{k}": np.array([L.ref_ext_stress(k, 0.1, 0.3, 1.3, 0.7, t) for t in t_in]) for k in range(4)}
    np.savez_compressed(os.path.join(HERE, "utils_golden.npz"), norm_in=norm_in, norm_out=norm_out, s_in=s_in,
                        w_c=w["c"], w_p=w["p"], w_m=w["m"], w_in=w_in, wi_c=wi["c"], wi_p=wi["p"], wi_m=wi["m"],
                        ph_ids=ids, ph_d2=d2, **{f"ph_{k}": v for k, v in ratios.items()}, steps=steps, t_in=t_in,
                        **{f"stress_{k}": v for k, v in stress.items()})

    shutil.copyfile(REF_DCONF, os.path.join(HERE, "1000_64.dconf"))
    out = {}
    for mult, hv in ((1e20, 0), (0.05 * 8, 1), (0.05 * 8, 0), (2.0, 0)):
        xy, order, bb, cut = r2.read_dconf(REF_DCONF, mult, bool(hv))
        out[f"cut_{mult}_{hv}"] = cut
    out.update(xy=xy, order=order, b=bb)
    np.savez_compressed(os.path.join(HERE, "dconf_golden.npz"), **out)
    print("fixtures written to", HERE)


def subprocess_build():
    import subprocess
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)


if __name__ == "__main__":
    main()
