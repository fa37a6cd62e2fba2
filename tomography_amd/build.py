"""Build the gfx950 library (tomography_amd/libtomo.so) and the oracle.

    python -m tomography_amd.build          # library + oracle (+ oracle/_ref when the reference tree exists)

Every HIP source is compiled with hipcc for gfx950 only, with packed-FP32
instruction selection disabled (-Xclang -target-feature -Xclang -packed-fp32-ops)
and multiply-add contraction limited to single expressions (-ffp-contract=on).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtomo.so")
ARCH = os.environ.get("TOMO_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=on: multiply-adds are fused within one expression only, so a
# stage computes bit-identical values whichever specialised variant of the
# surrounding code (full / partial time block, with / without the source
# cell) it is compiled into.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=on",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
SOURCES = ["kernels.hip", "setup.cpp", "runtime.cpp", "search.cpp", "fdtd_gpu_shim.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature: drop that notice
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-x", "hip"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        if verbose and out:
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, LIB)
    return LIB


def build_dropin_caller():
    """tests/_bin/dropin_caller: a C++ caller of FDTD_GPU/FDTD_CPU shaped like the
    reference's search driver, linked against libtomo.so (host code only)."""
    out_dir = os.path.join(ROOT, "tests", "_bin")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "dropin_caller")
    _run(["g++", "-std=c++17", "-O2", "-Wall", os.path.join(ROOT, "tests", "dropin_caller.cpp"), "-L" + HERE,
          "-l:libtomo.so", "-Wl,-rpath,$ORIGIN/../../tomography_amd", "-o", exe + ".tmp"])
    os.replace(exe + ".tmp", exe)
    return exe


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main():
    build_oracle()
    print(build_library(verbose=True))
    print(build_dropin_caller())


if __name__ == "__main__":
    sys.exit(main())
