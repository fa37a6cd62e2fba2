"""Small configurations shared by the parity tests (sizes the oracle runs in seconds)."""
import numpy as np

import tomography_amd as t


def ref_small(steps=300, n=101, flags=0):
    """REF_GPU physics on an n x n grid (disc r = 31.5 mm, dx = 1 mm)."""
    c = t.preset(t.PRESET_REF_GPU)
    return c.copy(nx=n, ny=n, ntff_offset=12, steps=steps, flags=flags)


def ref_cpu_small(steps=300, n=101):
    c = t.preset(t.PRESET_REF_CPU)
    return c.copy(nx=n, ny=n, ntff_offset=12, steps=steps)


def phys_small(steps=512, n=128, n_exc=2, flags=0):
    """Physical mode on a 64 mm domain (dx = 0.5 mm) with a small phantom."""
    c = t.preset(t.PRESET_CFG1)
    return c.copy(nx=n, ny=n, steps=steps, n_exc=n_exc, n_freq=4, n_angles=36, breast_radius=0.018,
                  skin_thickness=0.0015, src_radius=0.022, flags=flags)


def ref_params(seed=1, npch=3):
    rng = np.random.default_rng(seed)
    n2 = npch * npch
    return np.concatenate([rng.uniform(1.0, 7.0, n2), rng.uniform(1.5, 41.0, n2), rng.uniform(0.0, 0.3, n2)])


def ref_cpu_params(seed=1):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(1.0, 40.0, 81), rng.uniform(0.0, 1.1, 81)])


PHYS_PARAMS = np.array([3.0, -2.0, 4.0, 20.0, 34.0, 13.0, 0.2, 0.0])


def rel_l2(a, b):
    """||a - b|| / ||b|| in double precision (complex-aware)."""
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    n = np.sqrt((np.abs(b) ** 2).sum())
    return float(np.sqrt((np.abs(a - b) ** 2).sum()) / (n if n > 0 else 1.0))
