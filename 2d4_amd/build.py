"""Builds lib2d4.so (gfx950 kernels + C ABI + host Simulation) in-tree with hipcc.

The library goes to 2d4_amd/lib/lib2d4.so; nothing is installed elsewhere.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "lib2d4.so")

ARCH = "gfx950"
# -packed-fp32-ops: keep hipcc from forming packed FP32 instructions
# code object v5: loadable by the HIP runtime torch bundles as well as /opt/rocm's
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-mcode-object-version=5",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument"]

SOURCES = ["d4_kernels.hip", "d4_capi.hip", "d4_batch.hip", "d4_gpulu.hip", "d4_dstep.hip", "d4_adjoint.hip",
           "d4_sim.cpp", "d4_solver.cpp"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    return r


def build_lib(verbose=False):
    """Compiles every source under csrc/ into lib2d4.so (always recompiles)."""
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    # rocSOLVER / rocBLAS (dense LU on the GPU): with torch loaded first the
    # dynamic linker binds these sonames to the copies torch already holds
    libs = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    cmd = [_hipcc()] + HIPFLAGS + ["-shared", "-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs + libs
    _run(cmd, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_test_helpers(verbose=False):
    """Test-only helpers under tests/csrc: the kernel math compiled for the host,
    and the host sparse LU behind a C shim."""
    tdir = os.path.join(ROOT, "tests", "csrc")
    src = os.path.join(tdir, "hostmath.hip")
    if not os.path.exists(src):
        return None
    out = os.path.join(tdir, "libhostmath.so")
    cmd = [_hipcc(), "-O2", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
           "-Wno-unused-command-line-argument", "-o", out, src]
    _run(cmd, verbose)
    shim = os.path.join(tdir, "solver_shim.cpp")
    if os.path.exists(shim):
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", os.path.join(tdir, "libsolver_shim.so"),
              shim, os.path.join(CSRC, "d4_solver.cpp")], verbose)
    sim_shim = os.path.join(tdir, "sim_cpu_shim.cpp")
    if os.path.exists(sim_shim) and os.path.exists(os.path.join(CSRC, "d4_sim.cpp")):
        orc_src = os.path.join(ROOT, "oracle", "d4_oracle.c")
        orc_obj = os.path.join(tdir, "d4_oracle_sim.o")
        _run(["gcc", "-O2", "-fPIC", "-std=c11", "-c", "-o", orc_obj, orc_src], verbose)
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", os.path.join(tdir, "libsim_cpu.so"),
              sim_shim, os.path.join(CSRC, "d4_sim.cpp"), os.path.join(CSRC, "d4_solver.cpp"), orc_obj, "-lm"],
             verbose)
        # the same loop with the reference's optimisation flags (CMakeLists.txt:6; a portable -march in
        # place of native): the CPU side of bench.py's implicit-step measurement
        fast = ["-Ofast", "-ffast-math", "-march=x86-64-v3", "-fPIC"]
        orc_fast = os.path.join(tdir, "d4_oracle_sim_fast.o")
        _run(["gcc"] + fast + ["-std=c11", "-c", "-o", orc_fast, orc_src], verbose)
        _run(["g++"] + fast + ["-std=c++17", "-shared", "-o", os.path.join(tdir, "libsim_cpu_fast.so"), sim_shim,
              os.path.join(CSRC, "d4_sim.cpp"), os.path.join(CSRC, "d4_solver.cpp"), orc_fast, "-lm"], verbose)
    # a compiled caller that sees only the C ABI headers, linked against the built library
    drv = os.path.join(tdir, "capi_driver.cpp")
    if os.path.exists(drv) and os.path.exists(LIB):
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o",
              os.path.join(tdir, "capi_driver"), drv, "-L", LIBDIR, "-l2d4",
              "-Wl,-rpath,$ORIGIN/../../2d4_amd/lib", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    return out


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_test_helpers(verbose=True))
