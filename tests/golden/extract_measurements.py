"""Extract the reference's hard-coded "measurement" arrays into data files.

The reference compiles its target far field into the source
(FDTD_GPU.cu:996, 8000 floats = D[4][10][100] then the phase P in the same
layout; FDTD_CPU.cxx:789, 4000 floats = D).  The drop-in shim reads them from
measurement_gpu.f32 / measurement_cpu.f32 (raw little-endian float32) so that
the target is data, not code.  It writes into tomography_amd/data/, the
package data the shim loads.  Run once in a tree that has the reference:

    python tests/golden/extract_measurements.py /root/reference
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(os.path.dirname(os.path.dirname(HERE)), "tomography_amd", "data")


def extract(path, n):
    text = open(path).read()
    m = re.search(r"float\s+measurement\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}", text)
    if not m:
        raise ValueError(f"no measurement array in {path}")
    vals = np.array([float(t) for t in m.group(1).replace("\n", " ").split(",") if t.strip()], np.float32)
    if vals.size != n:
        raise ValueError(f"{path}: expected {n} values, found {vals.size}")
    return vals


def main(ref):
    extract(os.path.join(ref, "FDTD_GPU.cu"), 8000).tofile(os.path.join(DATA, "measurement_gpu.f32"))
    extract(os.path.join(ref, "FDTD_CPU.cxx"), 4000).tofile(os.path.join(DATA, "measurement_cpu.f32"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
