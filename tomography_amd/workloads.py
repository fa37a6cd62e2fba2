"""The benchmark workloads and the launch schedules measured on them, shared by
bench.py and the tests so that both run exactly the same configurations.

  cfg2 plain     1024^2, dx 0.25 mm, 8192 steps, one antenna (PLAIN_SCHEDULE)
  cfg3 batched   64 vectors x 16 antennas of 512^2 x 4096 steps (BATCH_SCHEDULE)
  cfg4 uneven    grids log-uniform in [256, 2048], steps clamp(8N, 2048, 16384)
  cfg5 serving   grids 256^2..2048^2, 1-16 antennas, 8-32-dim vectors
  REF_GPU        the FDTD_GPU() drop-in configuration (REF_SCHEDULE)
"""
import numpy as np

from ._abi import PRESET_CFG4, PRESET_CFG5
from .api import preset


def seeded_vectors(n, dim, seed):
    """Parameter vectors inside the physical bounds (tumour centre, radius, Debye values)."""
    rng = np.random.default_rng(seed)
    v = np.zeros((n, dim))
    for b in range(0, dim, 8):
        k = min(8, dim - b)
        blk = np.stack([rng.uniform(-15, 15, n), rng.uniform(-15, 15, n), rng.uniform(2, 6, n),
                        rng.uniform(10, 30, n), rng.uniform(10, 50, n), rng.uniform(8, 20, n),
                        rng.uniform(0, 1, n), rng.uniform(0, 4, n)], axis=1)
        v[:, b:b + k] = blk[:, :k]
    return v


def cell_steps(cfg):
    return float(cfg.nx) * cfg.ny * cfg.steps * cfg.n_exc


SCHED_KEYS = ("tsteps", "seg_rows", "edge_waves", "inner_waves", "chunk", "grid_mult", "form", "edge_rows",
              "inner_rows")
# Default schedules (swept on MI355X; fixed so that the profiled command and
# the timed one run the same launches).  The plain one: fused T = 4 blocks,
# one wave per workgroup, 1-row corner / row-edge tiles, 4-row column-edge
# tiles and 10-row interior tiles (on 1024^2 the most tiles that still run in
# one round of three waves per SIMD).  The batched one: four streams, T = 6
# blocks, 52-row column-edge and interior tiles (the 64-row cap) of four
# waves, 24-row row-edge tiles, 32 persistent workgroups per CU.  Both fold the running DFT over 60-step chunks: the chunk
# sets the FP32 partial sums, so equal chunks make batched and plain results
# equal bit for bit.  The drop-in one: resident T = 8 tiles
# of 12 rows over 64-step chunks.
PLAIN_SCHEDULE = dict(tsteps=4, seg_rows=4, edge_waves=1, inner_waves=1, chunk=60, grid_mult=16, form=2, edge_rows=1,
                      inner_rows=10)
BATCH_SCHEDULE = dict(tsteps=6, seg_rows=52, edge_waves=1, inner_waves=4, chunk=60, grid_mult=32, form=1,
                      edge_rows=24, inner_rows=0)
REF_SCHEDULE = dict(tsteps=8, seg_rows=12, edge_waves=1, inner_waves=1, chunk=64, grid_mult=16, form=3, edge_rows=1,
                    inner_rows=0)


def batch_vectors():
    return seeded_vectors(64, 8, 20261203)


SERVE_GRIDS = (256, 362, 512, 724, 1024, 1448, 2048)  # log-spaced 256^2..2048^2
SERVE_ANTENNAS = (1, 2, 4, 8, 16)


def serve_requests(n_req, seed=20261205):
    """Serving mix: grids log-spaced 256^2..2048^2, 1-16 antennas, 8-32-dim
    vectors (1-4 tumour blocks), drawn per request."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n_req):
        g = int(rng.choice(SERVE_GRIDS))
        cfg = preset(PRESET_CFG5, g).copy(n_exc=int(rng.choice(SERVE_ANTENNAS)))
        dim = int(rng.choice([8, 16, 24, 32]))
        reqs.append((cfg, seeded_vectors(1, dim, seed + 1 + i)[0]))
    return reqs


def uneven_requests(n, seed=20261204):
    """The uneven batch: grids log-uniform in [256, 2048] (steps clamp(8N, 2048,
    16384), one antenna each) and 0-3 tumours per phantom (0: a zero-radius
    tumour block)."""
    rng = np.random.default_rng(seed)
    grids = np.exp(rng.uniform(np.log(256), np.log(2048), n)).astype(int)
    ntum = rng.integers(0, 4, n)
    reqs = []
    for i in range(n):
        cfg = preset(PRESET_CFG4, int(grids[i]))
        v = seeded_vectors(1, 8 * max(1, int(ntum[i])), seed + 1 + i)[0]
        if ntum[i] == 0:
            v[2] = 0.0
        reqs.append((cfg, v))
    return reqs
