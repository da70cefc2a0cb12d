"""The C ABI library loads and exports every symbol include/*.h declares (no GPU needed)."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(d4\w*)\s*\(", src))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "d4_speed_pass" in names and "d4_jacobian_pass" in names and len(names) >= 15


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol(pkg):
    from importlib import import_module
    extra = import_module("2d4_amd._bind_extra").SIGNATURES
    bound = set(pkg.SIGNATURES) | set(extra)
    assert set(declared_symbols()) <= bound, set(declared_symbols()) - bound


def test_context_without_gpu_fails_loudly(pkg):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(pkg.D4Error):
        pkg.Context(0)
