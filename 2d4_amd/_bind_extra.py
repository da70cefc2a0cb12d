"""ctypes signatures of the entry points declared outside include/d4.h
(include/d4_batch.h, include/d4_sim.h); bound when those symbols exist."""
import ctypes as C

SIGNATURES = {}


def bind(L):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name, None)
        if f is not None:
            f.restype = res
            f.argtypes = args
