"""Builds libsfsim_atmo.so (HIP kernels for gfx950 + the C ABI) in place with hipcc.

Usage: python -m sfsim_amd.build  (or sfsim_amd.build.build() from Python).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsfsim_atmo.so")
SOURCES = ["kernels.hip", "surrogate.hip", "atmo.cpp"]

# -packed-fp32-ops: no v_pk_*_f32 generated for the device code (hipcc passes the flag to both
# compilation passes; the host pass warns that it does not know the feature).
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; libsfsim_atmo needs ROCm to build")
    return path


def build(verbose=False):
    """Compile every source to an object and link the shared library; returns its path."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        _run(cmd, verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    # --no-undefined: a symbol missing from the objects fails the link instead of the first dlopen
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    noise = [ln for ln in (res.stdout + res.stderr).splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln]
    if verbose and noise:
        print("\n".join(noise))


if __name__ == "__main__":
    print(build(verbose=True))
