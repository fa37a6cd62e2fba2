"""Write tests/golden/ref_fdtd_cpu_fitness.json: the value the reference's
whole FDTD_CPU() (FDTD_CPU.cxx:669-892, compiled unmodified into
oracle/_ref/fdtd_cpu_main by oracle/Makefile) returns for one 162-vector at its
native configuration (401^2, 4 excitations x 2000 steps; about four minutes
on one core), run in a fresh process.

Its far-field loop starts from an uninitialised index (FDTD_CPU.cxx:816).  In
a fresh process the projection does not run and D keeps the zero contents of
fresh heap memory, so the returned value is the fitness arithmetic of
FDTD_CPU.cxx:841-845 at D = 0 against the hard-coded measurement
(FDTD_CPU.cxx:789):
    fit (float) -= pow(D - m, 2) / (pow(m, 2) * 4 * 10)   over 4000 entries.
(Called inside a long-lived process instead, the same binary code returned
-inf: the undefined loop index and the recycled heap let garbage into D.)
The record keeps the vector, the returned value, the printed fitness line and
the last centre-Ez value the run prints (FDTD_CPU.cxx:808).

    python tests/golden/make_fdtd_cpu_fitness.py
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "fdtd_cpu_main")


def main():
    rng = np.random.default_rng(3)
    p = np.concatenate([rng.uniform(1.0, 40.0, 81), rng.uniform(0.0, 1.1, 81)])
    r = subprocess.run([EXE], input=" ".join(repr(float(x)) for x in p), capture_output=True, text=True,
                       check=True)
    lines = r.stdout.splitlines()
    fit = float(lines[-1].split()[-1])
    printed = [l for l in lines if "fitness" in l]
    trace = [l for l in lines[:-1] if "fitness" not in l and l.strip()]
    rec = dict(params=p.tolist(), fitness=fit, printed=printed[-1] if printed else None, trace_len=len(trace),
               last_centre_ez=float(trace[-1]) if trace else None)
    with open(os.path.join(HERE, "ref_fdtd_cpu_fitness.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print(rec["fitness"], rec["printed"], rec["trace_len"])


if __name__ == "__main__":
    main()
