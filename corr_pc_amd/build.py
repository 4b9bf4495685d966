"""Build recipe of the engine's shared library (corr_pc_amd/libcpc.so) for gfx950 and of the corr_pc
program over it (corr_pc_amd/corr_pc).

Run as ``python -m corr_pc_amd.build`` or through ``__graft_entry__.build()``.  Compiles with hipcc:
contraction off (the FP64 candidate test must round like the reference's x86-64 build) and packed
FP32 disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcpc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["cpc_kernels.hip", "cpc_sky.hip", "cpc_api.cpp", "cpc_out.cpp"]
# the corr_pc program: host code over the library's C ABI, built with the host compiler
PROGRAM = os.path.join(HERE, "corr_pc")
PROGRAM_SOURCE = "cpc_main.cpp"
CXX = os.environ.get("CXX", "g++")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Wall", "-Wno-unused-result",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    proc = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc


def build(verbose=False):
    """Compile every source and link libcpc.so; returns the library path."""
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + COMMON + ["-x", "hip"] + DEVICE + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_program(verbose)
    return LIB


def build_program(verbose=False):
    """Compile and link corr_pc_amd/corr_pc against libcpc.so (found beside it at run time)."""
    tmp = PROGRAM + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-o", tmp, PROGRAM_SOURCE, "-L" + HERE, "-lcpc",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, PROGRAM)
    return PROGRAM


def build_program_sanitized(verbose=False):
    """The corr_pc program built with AddressSanitizer and UndefinedBehaviorSanitizer (host code
    only; the library it calls is the ordinary build) as corr_pc_amd/_asan/corr_pc, for the CPU tests
    of its reader and planner.  Not part of build()."""
    out_dir = os.path.join(HERE, "_asan")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "corr_pc")
    tmp = out + ".tmp"
    cmd = [CXX, "-O1", "-g", "-std=c++17", "-Wall", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-o", tmp, PROGRAM_SOURCE, "-L" + HERE, "-lcpc",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
