"""Python host wrapper over the C ABI of libtomo.so (include/tomo.h).

The library is the product: every call here goes through the HIP kernels.
There is no CPU fallback; loading fails loudly when libtomo.so is missing.
"""
import ctypes as _c
import os

import numpy as np

from ._abi import (MAX_MATERIALS, SimConfig, TOMO_OK, TOMO_E_PENDING)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtomo.so")

_dp = _c.POINTER(_c.c_double)
_fp = _c.POINTER(_c.c_float)
_ip = _c.POINTER(_c.c_int)
_lib = None

# Every symbol declared in include/tomo.h.
EXPORTS = [
    "tomo_preset", "tomo_farfield_len", "tomo_contour_len", "tomo_ctx_create", "tomo_ctx_destroy",
    "tomo_last_error", "tomo_set_schedule", "tomo_get_schedule", "tomo_autotune", "tomo_eval", "tomo_eval_batch",
    "tomo_target_farfield", "tomo_submit", "tomo_poll", "tomo_wait", "tomo_cancel", "tomo_progress",
    "tomo_request_bytes", "tomo_sim_create", "tomo_sim_destroy", "tomo_sim_step", "tomo_sim_fields",
    "tomo_sim_materials", "tomo_sim_dft", "tomo_sim_farfield", "tomo_profile", "tomo_use_graphs", "tomo_get_stats",
    "tomo_search_run", "tomo_fitness", "tomo_eval_noalloc",
]


class Schedule(_c.Structure):
    """tomo_schedule (include/tomo.h)."""

    _fields_ = [("tsteps", _c.c_int32), ("seg_rows", _c.c_int32), ("edge_waves", _c.c_int32),
                ("inner_waves", _c.c_int32), ("chunk", _c.c_int32), ("grid_mult", _c.c_int32),
                ("form", _c.c_int32), ("edge_rows", _c.c_int32), ("inner_rows", _c.c_int32)]

    def as_tuple(self):
        return tuple(getattr(self, k) for k, _ in self._fields_)


KERNEL_NAMES = ("k_tblock", "k_tedge", "k_trow", "k_tcol", "k_tinner", "k_fold", "k_project_mfma", "k_fitness",
                "k_tfused", "k_tres", "k_twave")


class KernelStats(_c.Structure):
    """tomo_kernel_stats (include/tomo.h)."""

    _fields_ = [("launches", _c.c_int64), ("ms", _c.c_double), ("cell_steps", _c.c_double),
                ("alg_bytes", _c.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(_c.Structure):
    """tomo_stats (include/tomo.h): one record per kernel (TOMO_KERNEL_*) and
    one for the whole update phase."""

    _fields_ = [("kernel", KernelStats * len(KERNEL_NAMES)), ("update", KernelStats), ("requests", _c.c_int64),
                ("resident_refused", _c.c_int64)]

    def as_dict(self):
        d = {name: self.kernel[i].as_dict() for i, name in enumerate(KERNEL_NAMES)}
        d["update"] = self.update.as_dict()
        d["requests"] = self.requests
        d["resident_refused"] = self.resident_refused
        return d


class SearchConfig(_c.Structure):
    """tomo_search_config (include/tomo.h)."""

    _fields_ = [("method", _c.c_int32), ("pop", _c.c_int32), ("dim", _c.c_int32), ("max_evals", _c.c_int32),
                ("inflight", _c.c_int32), ("pad", _c.c_int32), ("F", _c.c_double), ("CR", _c.c_double),
                ("seed", _c.c_uint64)]


class SearchResult(_c.Structure):
    """tomo_search_result (include/tomo.h)."""

    _fields_ = [("best_fitness", _c.c_double), ("evals", _c.c_int64), ("failed", _c.c_int64),
                ("wall_s", _c.c_double), ("evals_per_s", _c.c_double), ("p50_latency_s", _c.c_double),
                ("p99_latency_s", _c.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TomoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tomo error {code}: {msg}")
        self.code = code


def lib():
    """Load libtomo.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tomography_amd.build`")
    L = _c.CDLL(LIB_PATH)
    vp = _c.c_void_p
    L.tomo_last_error.restype = _c.c_char_p
    L.tomo_preset.argtypes = [_c.c_int, _c.c_int, vp]
    L.tomo_farfield_len.argtypes = [vp]
    L.tomo_farfield_len.restype = _c.c_int64
    L.tomo_contour_len.argtypes = [vp]
    L.tomo_contour_len.restype = _c.c_int32
    L.tomo_ctx_create.argtypes = [_c.c_int, _c.c_uint64, _c.POINTER(vp)]
    L.tomo_ctx_destroy.argtypes = [vp]
    L.tomo_set_schedule.argtypes = [vp, vp]
    L.tomo_get_schedule.argtypes = [vp, vp]
    L.tomo_autotune.argtypes = [vp, vp, _c.c_int, vp, _dp]
    L.tomo_eval.argtypes = [vp, vp, _dp, _c.c_size_t, _fp, _c.c_size_t, _dp, _fp, _fp]
    L.tomo_eval_batch.argtypes = [vp, vp, _dp, _c.c_size_t, _c.c_size_t, _fp, _c.c_size_t, _dp, _fp]
    L.tomo_eval_noalloc.argtypes = [vp, vp, _dp, _c.c_size_t, _fp, _c.c_size_t, _dp, _fp, _fp]
    L.tomo_target_farfield.argtypes = [vp, vp, _fp]
    L.tomo_submit.argtypes = [vp, vp, _dp, _c.c_size_t, _c.POINTER(_c.c_uint64)]
    L.tomo_poll.argtypes = [vp, _c.c_uint64, _dp, _dp]
    L.tomo_wait.argtypes = [vp, _c.c_uint64, _dp, _dp]
    L.tomo_cancel.argtypes = [vp, _c.c_uint64]
    L.tomo_progress.argtypes = [vp, _ip]
    L.tomo_request_bytes.argtypes = [vp]
    L.tomo_request_bytes.restype = _c.c_int64
    L.tomo_sim_create.argtypes = [vp, vp, _dp, _c.c_size_t, _c.c_int, _c.POINTER(vp)]
    L.tomo_sim_destroy.argtypes = [vp]
    L.tomo_sim_step.argtypes = [vp, _c.c_int]
    L.tomo_sim_fields.argtypes = [vp, _fp, _fp, _fp, _fp]
    L.tomo_sim_materials.argtypes = [vp, _c.POINTER(_c.c_uint8), _fp, _ip]
    L.tomo_sim_dft.argtypes = [vp, _fp, _fp]
    L.tomo_sim_farfield.argtypes = [vp, _fp, _fp]
    L.tomo_profile.argtypes = [vp, _c.c_int]
    L.tomo_use_graphs.argtypes = [vp, _c.c_int]
    L.tomo_get_stats.argtypes = [vp, _c.c_void_p]
    L.tomo_search_run.argtypes = [vp, vp, _dp, _dp, vp, _dp, vp]
    L.tomo_fitness.argtypes = [vp, _c.c_int, _c.c_int32, _dp, _c.c_size_t, _fp, _c.c_size_t, _dp]
    _lib = L
    return L


def _check(rc):
    if rc != TOMO_OK:
        raise TomoError(rc, lib().tomo_last_error().decode())


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def preset(preset_id, grid=0):
    c = SimConfig()
    _check(lib().tomo_preset(preset_id, grid, _c.byref(c)))
    return c


class Context:
    """A device context: the owner of resident simulations and the request loop."""

    def __init__(self, device=0, hbm_budget=0):
        h = _c.c_void_p()
        _check(lib().tomo_ctx_create(device, hbm_budget, _c.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().tomo_ctx_destroy(self.h)  # frees the slots of live Sim handles too
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_schedule(self, tsteps=0, seg_rows=0, edge_waves=0, chunk=0, inner_waves=0, grid_mult=0, form=0,
                     edge_rows=0, inner_rows=0):
        s = Schedule(tsteps, seg_rows, edge_waves, inner_waves, chunk, grid_mult, form, edge_rows, inner_rows)
        _check(lib().tomo_set_schedule(self.h, _c.byref(s)))

    def autotune(self, cfg, n_sims):
        """Pick and install the fastest schedule for n_sims resident simulations of cfg."""
        s, r = Schedule(), _c.c_double()
        _check(lib().tomo_autotune(self.h, _c.byref(cfg), int(n_sims), _c.byref(s), _c.byref(r)))
        return s.as_tuple(), r.value

    def profile(self, enable=True):
        """Start (or stop) per-launch HIP-event timing; resets the counters."""
        _check(lib().tomo_profile(self.h, int(enable)))

    def use_graphs(self, enable=True):
        """Replay full epochs from a captured hipGraph (default) or launch each kernel."""
        _check(lib().tomo_use_graphs(self.h, int(enable)))

    def stats(self):
        st = Stats()
        _check(lib().tomo_get_stats(self.h, _c.byref(st)))
        return st.as_dict()

    def fitness(self, preset_id, params, flags=0, measurement=None):
        """The objective entry: preset + vector -> fitness (packaged measurement when None)."""
        p = np.ascontiguousarray(params, np.float64)
        m = None if measurement is None else np.ascontiguousarray(measurement, np.float32)
        f = _c.c_double()
        _check(lib().tomo_fitness(self.h, preset_id, flags, _p(p, _dp), p.size, _p(m, _fp),
                                  0 if m is None else m.size, _c.byref(f)))
        return f.value

    def search(self, cfg, lo, hi, method, pop, max_evals, inflight=0, F=0.5, CR=0.9, seed=1):
        """Asynchronous DE / PSO over the request loop: returns (best_x, result dict)."""
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        sc = SearchConfig(method, pop, lo.size, max_evals, inflight, 0, F, CR, seed)
        best = np.zeros(lo.size)
        res = SearchResult()
        _check(lib().tomo_search_run(self.h, _c.byref(cfg), _p(lo, _dp), _p(hi, _dp), _c.byref(sc), _p(best, _dp),
                                     _c.byref(res)))
        return best, res.as_dict()

    def schedule(self):
        """(tsteps, seg_rows, edge_waves, inner_waves, chunk, grid_mult, form, edge_rows, inner_rows)"""
        s = Schedule()
        _check(lib().tomo_get_schedule(self.h, _c.byref(s)))
        return s.as_tuple()

    def eval(self, cfg, params, target=None, want_farfield=False):
        """One fitness evaluation (the FDTD_GPU() twin): returns fitness[, D, P]."""
        p = np.ascontiguousarray(params, np.float64)
        t = None if target is None else np.ascontiguousarray(target, np.float32)
        L = cfg.farfield_len()
        D = np.zeros(L, np.float32) if want_farfield else None
        P = np.zeros(L, np.float32) if want_farfield else None
        fit = _c.c_double()
        _check(lib().tomo_eval(self.h, _c.byref(cfg), _p(p, _dp), p.size, _p(t, _fp), 0 if t is None else t.size,
                               _c.byref(fit), _p(D, _fp), _p(P, _fp)))
        return (fit.value, D, P) if want_farfield else fit.value

    def eval_noalloc(self, cfg, params, target, want_farfield=False):
        """tomo_eval without device allocation (the target is required)."""
        p = np.ascontiguousarray(params, np.float64)
        t = np.ascontiguousarray(target, np.float32)
        L = cfg.farfield_len()
        D = np.zeros(L, np.float32) if want_farfield else None
        P = np.zeros(L, np.float32) if want_farfield else None
        fit = _c.c_double()
        _check(lib().tomo_eval_noalloc(self.h, _c.byref(cfg), _p(p, _dp), p.size, _p(t, _fp), t.size, _c.byref(fit),
                                       _p(D, _fp), _p(P, _fp)))
        return (fit.value, D, P) if want_farfield else fit.value

    def eval_batch(self, cfg, params, target=None, want_farfield=False):
        p = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        n, dim = p.shape
        t = None if target is None else np.ascontiguousarray(target, np.float32)
        fit = np.zeros(n)
        D = np.zeros(n * cfg.farfield_len(), np.float32) if want_farfield else None
        _check(lib().tomo_eval_batch(self.h, _c.byref(cfg), _p(p, _dp), n, dim, _p(t, _fp),
                                     0 if t is None else t.size, _p(fit, _dp), _p(D, _fp)))
        return (fit, D.reshape(n, -1)) if want_farfield else fit

    def target_farfield(self, cfg):
        D = np.zeros(cfg.farfield_len(), np.float32)
        _check(lib().tomo_target_farfield(self.h, _c.byref(cfg), _p(D, _fp)))
        return D

    # ---- request loop
    def submit(self, cfg, params):
        p = np.ascontiguousarray(params, np.float64)
        t = _c.c_uint64()
        _check(lib().tomo_submit(self.h, _c.byref(cfg), _p(p, _dp), p.size, _c.byref(t)))
        return t.value

    def poll(self, ticket):
        """(done, fitness, latency_s); raises on a failed request."""
        f, lat = _c.c_double(), _c.c_double()
        rc = lib().tomo_poll(self.h, ticket, _c.byref(f), _c.byref(lat))
        if rc == TOMO_E_PENDING:
            return False, None, None
        _check(rc)
        return True, f.value, lat.value

    def wait(self, ticket):
        f, lat = _c.c_double(), _c.c_double()
        _check(lib().tomo_wait(self.h, ticket, _c.byref(f), _c.byref(lat)))
        return f.value, lat.value

    def cancel(self, ticket):
        _check(lib().tomo_cancel(self.h, ticket))

    def progress(self):
        n = _c.c_int()
        _check(lib().tomo_progress(self.h, _c.byref(n)))
        return n.value

    def sim(self, cfg, params, exc=0):
        return Sim(self, cfg, params, exc)


def request_bytes(cfg):
    return int(lib().tomo_request_bytes(_c.byref(cfg)))


class Sim:
    """One resident simulation (parity-test access to fields and accumulators)."""

    def __init__(self, ctx, cfg, params, exc=0):
        self.ctx, self.cfg = ctx, cfg
        if not ctx.h:
            raise TomoError(-8, "context is closed")
        p = np.ascontiguousarray(params, np.float64)
        h = _c.c_void_p()
        _check(lib().tomo_sim_create(ctx.h, _c.byref(cfg), _p(p, _dp), p.size, exc, _c.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            # safe in either order: after tomo_ctx_destroy the call frees the handle alone
            lib().tomo_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n):
        _check(lib().tomo_sim_step(self.h, int(n)))

    def fields(self):
        c = self.cfg
        ez = np.zeros((c.ny + 1, c.nx + 1), np.float32)
        jp = np.zeros_like(ez)
        hx = np.zeros((c.ny, c.nx), np.float32)
        hy = np.zeros_like(hx)
        _check(lib().tomo_sim_fields(self.h, _p(ez, _fp), _p(hx, _fp), _p(hy, _fp), _p(jp, _fp)))
        return ez, hx, hy, jp

    def materials(self):
        c = self.cfg
        mat = np.zeros((c.ny + 1, c.nx + 1), np.uint8)
        coef = np.zeros(MAX_MATERIALS * 8, np.float32)
        n = _c.c_int()
        _check(lib().tomo_sim_materials(self.h, mat.ctypes.data_as(_c.POINTER(_c.c_uint8)), _p(coef, _fp),
                                        _c.byref(n)))
        return mat, coef[:n.value * 8].reshape(n.value, 8)

    def dft(self):
        c = self.cfg
        P = c.contour_len()
        J = np.zeros(2 * c.n_freq * P, np.float32)
        M = np.zeros_like(J)
        _check(lib().tomo_sim_dft(self.h, _p(J, _fp), _p(M, _fp)))
        return ((J[0::2] + 1j * J[1::2]).reshape(c.n_freq, P), (M[0::2] + 1j * M[1::2]).reshape(c.n_freq, P))

    def farfield(self):
        c = self.cfg
        D = np.zeros(c.n_freq * c.n_angles, np.float32)
        P = np.zeros_like(D)
        _check(lib().tomo_sim_farfield(self.h, _p(D, _fp), _p(P, _fp)))
        return D.reshape(c.n_freq, c.n_angles), P.reshape(c.n_freq, c.n_angles)


def measurement(kind="gpu"):
    """The reference's target far field (FDTD_GPU.cu:996 / FDTD_CPU.cxx:789) as shipped data."""
    path = os.path.join(_HERE, "data", f"measurement_{kind}.f32")
    return np.fromfile(path, np.float32)
