"""Build recipe for the gfx950 backend library (aurora_amd/libaph.so).

Compiles every HIP/C++ source under aurora_amd/csrc with hipcc for gfx950 and
links one shared library exporting the C-ABI of include/aph.h.  Packed FP32
instructions are disabled for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaph.so")
ARCH = os.environ.get("APH_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.cpp", "server.cpp", "assets.cpp", "gltf.cpp", "guides.cpp", "denoiser.cpp", "alias.cpp", "bvh.hip", "sah.hip", "kernels.hip",
           "wavefront.hip", "temporal.hip", "denoise.hip"]

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Xarch_host", "-ffp-contract=off"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return proc.stdout


def build(verbose=False, jobs=4):
    """Compiles all sources (always; no staleness check) and links libaph.so."""
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        cmd = ["hipcc"] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = ["hipcc", "-x", "hip"] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs, verbose)
    _drain(procs, verbose)
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def _drain(procs, verbose):
    failed = None
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = cmd
        elif verbose and out:
            sys.stderr.write(out)
    procs.clear()
    if failed:
        raise RuntimeError("build failed: " + " ".join(failed))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
