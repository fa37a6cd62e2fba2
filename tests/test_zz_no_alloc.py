"""tomo_eval_noalloc, the allocation-free variant of tomo_eval, called raw
through the C ABI: the same fitness, far field and phase as tomo_eval, bit
for bit (same kernels, same schedule, same reduction), on two calls in a row
of which the second is smaller (its slot, tables and records sit at other
offsets of the same scratch array, which still holds the first call's
contents); and the size and argument limits return their status codes."""
import ctypes as C

import numpy as np
import pytest

import tomography_amd as t
from cases import PHYS_PARAMS, phys_small, ref_params, ref_small

pytestmark = pytest.mark.gpu


def noalloc(ctx, cfg, p, target):
    L = cfg.farfield_len()
    p = np.ascontiguousarray(p, np.float64)
    tg = np.ascontiguousarray(target, np.float32)
    D = np.zeros(L, np.float32)
    P = np.zeros(L, np.float32)
    f = C.c_double()
    rc = t.lib().tomo_eval_noalloc(ctx.h, C.byref(cfg), p.ctypes.data_as(C.POINTER(C.c_double)), p.size,
                                   tg.ctypes.data_as(C.POINTER(C.c_float)), tg.size, C.byref(f),
                                   D.ctypes.data_as(C.POINTER(C.c_float)), P.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, f.value, D, P


def test_noalloc_equals_eval_on_two_calls_second_smaller():
    big = phys_small(steps=600, n=256, n_exc=2)
    small = ref_small(steps=300, n=101).copy(n_exc=2)
    with t.Context() as ctx:
        ctx.set_schedule(tsteps=4, seg_rows=4, edge_waves=1, inner_waves=1, chunk=64, grid_mult=16,
                         form=t.FORM_FUSED, edge_rows=1, inner_rows=14)
        tgt_big = ctx.target_farfield(big)
        rng = np.random.default_rng(3)
        tgt_small = rng.uniform(0.0, 0.05, small.farfield_len()).astype(np.float32)
        for cfg, p, tgt in ((big, PHYS_PARAMS, tgt_big), (small, ref_params(5), tgt_small)):
            rc, f, D, P = noalloc(ctx, cfg, p, tgt)
            assert rc == t.TOMO_OK, t.lib().tomo_last_error().decode()
            fe, De, Pe = ctx.eval(cfg, p, target=tgt, want_farfield=True)
            assert f == fe and np.array_equal(D, De) and np.array_equal(P, Pe), (cfg.nx, f, fe)
            assert np.isfinite(f) and f < 0


def test_noalloc_limits():
    with t.Context() as ctx:
        cfg = t.preset(t.PRESET_CFG4, 2048)
        tgt = np.zeros(cfg.farfield_len(), np.float32)
        rc, *_ = noalloc(ctx, cfg, PHYS_PARAMS, tgt)
        assert rc == t.TOMO_E_NOMEM and "scratch" in t.lib().tomo_last_error().decode()
        small = phys_small(steps=64, n=96, n_exc=1)
        rc, *_ = noalloc(ctx, small, PHYS_PARAMS, np.zeros(3, np.float32))
        assert rc == t.TOMO_E_INVALID
