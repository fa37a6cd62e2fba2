"""Write the golden fixtures of the reference CPU path (tests/golden/*.npz).

Runs the reference's own functions (FDTD_CPU.cxx, FDTD_common.cxx compiled
unmodified into oracle/_ref by oracle/Makefile) on seeded inputs and stores
small input/output vectors, so that tests on machines without the reference
tree (a GPU machine with only this repository) can still check against the
reference:

  ref_cpu_cpml.npz    CPML tables kex, kmx, aex, bex, amx, bmx (FDTD_CPU.cxx:765-776)
  ref_cpu_run.npz     162-vector; material patch map; per-material coefficient
                      table; contour (x, y); centre-Ez trace (FDTD_CPU.cxx:808);
                      Ez every 4th node; J and M accumulators; far field D of
                      excitation 1 after 900 steps (freq_index loop started at 0)
  ref_cpu_dispersionless.npz  same for excitation 0 after 700 steps, with a
                      vector whose sigma is zero everywhere

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.binding import Reference, unpack_ref_cj  # noqa: E402


def params162(seed, sigma=True):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.0, 40.0, 81)
    sig = rng.uniform(0.0, 1.1, 81) if sigma else np.zeros(81)
    return np.concatenate([eps, sig])


def material_table(ref, p):
    eps, sig, Ceze, Cezhy, Cezeic, Cezeip = ref.materials(p)
    # patch id per node: 0 outside the disc, 1 + patch inside (identified by eps/sigma values)
    keys = {}
    pid = np.zeros(eps.shape, np.uint8)
    table = [(1.0, 0.0)]
    for k in range(81):
        table.append((np.float32(p[k]), np.float32(p[81 + k])))
    inside = ~((eps == 1.0) & (sig == 0.0))
    for k in range(81):
        m = inside & (eps == np.float32(p[k])) & (sig == np.float32(p[81 + k]))
        pid[m] = 1 + k
    coef = np.zeros((82, 4), np.float32)
    for m in range(82):
        sel = pid == m
        if sel.any():
            i = np.argwhere(sel)[0]
            coef[m] = [Ceze[tuple(i)], Cezhy[tuple(i)], Cezeic[tuple(i)], Cezeip[tuple(i)]]
            assert np.all(Ceze[sel] == coef[m, 0]) and np.all(Cezhy[sel] == coef[m, 1])
    return pid, coef


def run_case(ref, p, exc, nsteps, path):
    out = ref.run(p, exc, nsteps)
    J, M = unpack_ref_cj(out["cj"])
    pid, coef = material_table(ref, p)
    xs, ys = ref.contour()
    np.savez_compressed(path, params=p, exc=exc, nsteps=nsteps, patch_map=pid, coef=coef, contour_x=xs,
                        contour_y=ys, trace=out["trace"], ez_sub=out["ez"][::4, ::4],
                        ez_sum=np.float64(out["ez"].astype(np.float64).sum()),
                        ez_l2=np.float64(np.sqrt((out["ez"].astype(np.float64) ** 2).sum())),
                        hx_l2=np.float64(np.sqrt((out["hx"].astype(np.float64) ** 2).sum())),
                        hy_l2=np.float64(np.sqrt((out["hy"].astype(np.float64) ** 2).sum())),
                        J=J.astype(np.complex64), M=M.astype(np.complex64), D=out["D"])


def main():
    ref = Reference()
    np.savez_compressed(os.path.join(HERE, "ref_cpu_cpml.npz"), tables=ref.cpml_tables())
    run_case(ref, params162(20261202), 1, 900, os.path.join(HERE, "ref_cpu_run.npz"))
    run_case(ref, params162(7, sigma=False), 0, 700, os.path.join(HERE, "ref_cpu_dispersionless.npz"))


if __name__ == "__main__":
    main()
