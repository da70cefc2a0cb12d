"""The LU kernels' DPP instructions against the hazard the compiler does not track through inline asm.

lu_panel_kernel updates its blocks with `v_fmac_f64_dpp ... row_newbcast` written as inline asm
(2d4_amd/csrc/d4_gpulu.hip).  On CDNA a DPP instruction must not read, as its DPP source, a VGPR that a
VALU instruction wrote in the two wait states before it, nor follow a write of EXEC by fewer than five;
the hardware does not interlock and the compiler's hazard recognizer does not look inside inline asm.
The kernels keep every DPP source a value loaded from LDS a step earlier.  This test compiles the file
for gfx950 and checks every DPP instruction of the assembly for both hazards, with the checker itself
pinned on planted cases."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "2d4_amd", "csrc", "d4_gpulu.hip")


def _vgprs(tok):
    tok = tok.lstrip("-")
    m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
    if m:
        return set(range(int(m[1]), int(m[2]) + 1))
    m = re.fullmatch(r"v(\d+)", tok)
    return {int(m[1])} if m else set()


def dpp_hazards(lines):
    """(number of DPP instructions, the ones whose source a VALU wrote within 2 wait states or that follow
    an EXEC write within 5), scanning straight-line code; a label resets nothing (conservative)."""
    hist, count, bad = [], 0, []
    for ln in lines:
        t = ln.strip()
        if not t or t.startswith((";", ".")) or t.endswith(":"):
            continue
        parts = t.replace(",", " ").split()
        op = parts[0]
        if op == "s_nop":
            hist.append((int(parts[1], 0) + 1, set(), False))
            continue
        if "_dpp" in op:
            count += 1
            src0 = _vgprs(parts[2])
            ws = 0
            for w, dst, exw in reversed(hist):
                if ws < 2 and dst & src0:
                    bad.append(t)
                    break
                if ws < 5 and exw:
                    bad.append(t + "  [exec]")
                    break
                ws += w
                if ws >= 5:
                    break
        valu_dst = (op.startswith("v_") and len(parts) > 1
                    and not op.startswith(("v_cmp", "v_readlane", "v_readfirstlane")))
        writes_exec = (len(parts) > 1 and parts[1] == "exec") or op.startswith("v_cmpx")
        hist.append((1, _vgprs(parts[1]) if valu_dst else set(), writes_exec))
    return count, bad


_FMAC = "\tv_fmac_f64_dpp v[0:1], -v[2:3], v[6:7] row_newbcast:1 row_mask:0xf bank_mask:0xf"


@pytest.mark.parametrize("prefix,hazard", [
    (["\tv_mov_b64_e32 v[2:3], v[4:5]"], True),                              # written just before
    (["\tv_add_f64 v[2:3], v[4:5], v[4:5]", "\ts_nop 1"], False),            # two wait states
    (["\tv_add_f64 v[2:3], v[4:5], v[4:5]", "\tv_add_f64 v[8:9], v[4:5], v[4:5]"], True),  # one between
    (["\tv_add_f64 v[2:3], v[4:5], v[4:5]", "\ts_nop 0"], True),             # one wait state
    (["\tv_add_f64 v[3:4], v[4:5], v[4:5]", "\tv_mov_b32 v9, v4"], True),    # half of the pair
    (["\tv_add_f64 v[3:4], v[4:5], v[4:5]", "\tv_mov_b32 v9, v4", "\ts_nop 0"], False),
    (["\tds_read_b64 v[2:3], v1", "\ts_waitcnt lgkmcnt(0)"], False),         # a load is no VALU write
    (["\ts_mov_b64 exec, -1", "\ts_nop 2"], True),                           # EXEC within 5
    (["\ts_mov_b64 exec, -1", "\ts_nop 4"], False),
])
def test_checker_on_planted_cases(prefix, hazard):
    n, bad = dpp_hazards(prefix + [_FMAC])
    assert n == 1 and bool(bad) == hazard


def test_lu_kernels_have_no_dpp_hazard(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path / "gpulu.s"
    cmd = [hipcc, "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-mcode-object-version=5",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
           "-I", os.path.join(ROOT, "include"), SRC, "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    lines = out.read_text().splitlines()
    n, bad = dpp_hazards(lines)
    # the panel's two blocks: 2 x 2016 updates, and the 64 pivot reciprocals broadcast by the compiler
    assert n >= 2 * 2016 + 64
    assert not bad, bad[:5]
