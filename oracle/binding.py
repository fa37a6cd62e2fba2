"""ctypes binding of oracle/liboracle.so and oracle/_ref/libfdtd_ref.so.

TEST INFRASTRUCTURE ONLY: what the HIP product's results are checked against,
never the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libfdtd_ref.so")

_c = ctypes
_dp = _c.POINTER(_c.c_double)
_fp = _c.POINTER(_c.c_float)
_ip = _c.POINTER(_c.c_int)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Oracle:
    """The CPU restatement (fdtd_oracle.cpp)."""

    def __init__(self, path=_ORACLE):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = _c.CDLL(path)
        L.orc_last_error.restype = _c.c_char_p
        L.orc_sim_create.argtypes = [_c.c_void_p, _dp, _c.c_size_t, _c.c_int, _c.c_int, _c.POINTER(_c.c_void_p)]
        L.orc_sim_destroy.argtypes = [_c.c_void_p]
        L.orc_sim_step.argtypes = [_c.c_void_p, _c.c_int]
        L.orc_sim_fields.argtypes = [_c.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_sim_probe.argtypes = [_c.c_void_p, _c.c_int, _c.c_int]
        L.orc_sim_probe.restype = _c.c_double
        L.orc_sim_materials.argtypes = [_c.c_void_p, _c.POINTER(_c.c_uint8), _dp]
        L.orc_sim_cpml.argtypes = [_c.c_void_p, _dp]
        L.orc_sim_dft.argtypes = [_c.c_void_p, _dp, _dp]
        L.orc_sim_farfield.argtypes = [_c.c_void_p, _dp, _dp]
        L.orc_eval.argtypes = [_c.c_void_p, _dp, _c.c_size_t, _fp, _c.c_size_t, _c.c_int, _dp, _dp, _dp]
        L.orc_contour.argtypes = [_c.c_void_p, _ip, _ip, _ip]
        L.orc_target_vector.argtypes = [_c.c_void_p, _dp]
        L.orc_source_cell.argtypes = [_c.c_void_p, _c.c_int, _ip, _ip]
        L.orc_incident_power.argtypes = [_c.c_void_p, _c.c_double]
        L.orc_incident_power.restype = _c.c_double
        L.orc_farfield.argtypes = [_c.c_void_p, _dp, _c.c_size_t, _c.c_int, _c.c_int, _dp, _dp]
        L.orc_fitness.argtypes = [_c.c_void_p, _dp, _dp, _fp, _fp, _c.c_size_t, _dp]

    def err(self):
        return self.lib.orc_last_error().decode()

    def sim(self, cfg, params, exc=0, fp64=False):
        return OracleSim(self, cfg, params, exc, fp64)

    def contour(self, cfg):
        P = 2 * cfg.nx + 2 * cfg.ny - 8 * cfg.ntff_offset - 4
        xs, ys, lg = (np.zeros(P, np.int32) for _ in range(3))
        self.lib.orc_contour(_c.byref(cfg), _ptr(xs, _ip), _ptr(ys, _ip), _ptr(lg, _ip))
        return xs, ys, lg

    def eval(self, cfg, params, target=None, fp64=False):
        p = np.ascontiguousarray(params, np.float64)
        L = cfg.n_exc * cfg.n_freq * cfg.n_angles
        D = np.zeros(L)
        P = np.zeros(L)
        fit = _c.c_double()
        t = None if target is None else np.ascontiguousarray(target, np.float32)
        rc = self.lib.orc_eval(_c.byref(cfg), _ptr(p, _dp), p.size, _ptr(t, _fp), 0 if t is None else t.size,
                               int(fp64), _c.byref(fit), _ptr(D, _dp), _ptr(P, _dp))
        if rc != 0:
            raise ValueError(self.err())
        return fit.value, D, P

    def farfield(self, cfg, params, exc, fp64=True):
        """Far field (D, P), [freq][angle], of one excitation run to cfg.steps.
        Releases the GIL for the whole run, so host threads run several at once."""
        p = np.ascontiguousarray(params, np.float64)
        D = np.zeros(cfg.n_freq * cfg.n_angles)
        P = np.zeros_like(D)
        if self.lib.orc_farfield(_c.byref(cfg), _ptr(p, _dp), p.size, int(exc), int(fp64), _ptr(D, _dp),
                                 _ptr(P, _dp)) != 0:
            raise ValueError(self.err())
        return D.reshape(cfg.n_freq, cfg.n_angles), P.reshape(cfg.n_freq, cfg.n_angles)

    def fitness(self, cfg, D, target, P=None, target_phase=None):
        """The oracle's fitness arithmetic on given far-field records (cfg.flags select the form)."""
        D = np.ascontiguousarray(D, np.float64).ravel()
        t = np.ascontiguousarray(target, np.float32).ravel()
        Pp = None if P is None else np.ascontiguousarray(P, np.float64).ravel()
        tp = None if target_phase is None else np.ascontiguousarray(target_phase, np.float32).ravel()
        assert t.size >= D.size and (Pp is None or Pp.size == D.size) and (tp is None or tp.size >= D.size)
        f = _c.c_double()
        self.lib.orc_fitness(_c.byref(cfg), _ptr(D, _dp), _ptr(Pp, _dp), _ptr(t, _fp), _ptr(tp, _fp), D.size,
                             _c.byref(f))
        return f.value

    def target_vector(self, cfg):
        v = np.zeros(8)
        self.lib.orc_target_vector(_c.byref(cfg), _ptr(v, _dp))
        return v

    def source_cell(self, cfg, exc):
        x, y = _c.c_int(), _c.c_int()
        self.lib.orc_source_cell(_c.byref(cfg), exc, _c.byref(x), _c.byref(y))
        return x.value, y.value


class OracleSim:
    def __init__(self, orc, cfg, params, exc, fp64):
        self.o, self.cfg = orc, cfg
        p = np.ascontiguousarray(params, np.float64)
        h = _c.c_void_p()
        if orc.lib.orc_sim_create(_c.byref(cfg), _ptr(p, _dp), p.size, exc, int(fp64), _c.byref(h)) != 0:
            raise ValueError(orc.err())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.orc_sim_destroy(self.h)
            self.h = None

    def step(self, n):
        self.o.lib.orc_sim_step(self.h, int(n))

    def probe(self, x, y):
        return self.o.lib.orc_sim_probe(self.h, int(x), int(y))

    def fields(self):
        c = self.cfg
        ez = np.zeros((c.ny + 1) * (c.nx + 1))
        jp = np.zeros_like(ez)
        hx = np.zeros(c.nx * c.ny)
        hy = np.zeros_like(hx)
        self.o.lib.orc_sim_fields(self.h, _ptr(ez, _dp), _ptr(hx, _dp), _ptr(hy, _dp), _ptr(jp, _dp))
        return (ez.reshape(c.ny + 1, c.nx + 1), hx.reshape(c.ny, c.nx), hy.reshape(c.ny, c.nx),
                jp.reshape(c.ny + 1, c.nx + 1))

    def materials(self):
        c = self.cfg
        n1 = (c.nx + 1) * (c.ny + 1)
        mat = np.zeros(n1, np.uint8)
        coef = np.zeros(9 * n1)
        self.o.lib.orc_sim_materials(self.h, mat.ctypes.data_as(_c.POINTER(_c.c_uint8)), _ptr(coef, _dp))
        return mat.reshape(c.ny + 1, c.nx + 1), coef.reshape(9, c.ny + 1, c.nx + 1)

    def cpml(self):
        out = np.zeros(6 * self.cfg.ncpml)
        self.o.lib.orc_sim_cpml(self.h, _ptr(out, _dp))
        return out.reshape(6, self.cfg.ncpml)

    def dft(self):
        c = self.cfg
        P = 2 * c.nx + 2 * c.ny - 8 * c.ntff_offset - 4
        J = np.zeros(2 * c.n_freq * P)
        M = np.zeros_like(J)
        self.o.lib.orc_sim_dft(self.h, _ptr(J, _dp), _ptr(M, _dp))
        J = (J[0::2] + 1j * J[1::2]).reshape(c.n_freq, P)
        M = (M[0::2] + 1j * M[1::2]).reshape(c.n_freq, P)
        return J, M

    def farfield(self):
        c = self.cfg
        D = np.zeros(c.n_freq * c.n_angles)
        P = np.zeros_like(D)
        self.o.lib.orc_sim_farfield(self.h, _ptr(D, _dp), _ptr(P, _dp))
        return D.reshape(c.n_freq, c.n_angles), P.reshape(c.n_freq, c.n_angles)


class Reference:
    """The reference's own FDTD_CPU functions (oracle/_ref), when built."""

    NX = 401

    def __init__(self, path=_REF):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = _c.CDLL(path)
        L.ref_cpml_tables.argtypes = [_fp]
        L.ref_materials.argtypes = [_dp, _fp, _fp, _fp, _fp, _fp, _fp]
        L.ref_run.argtypes = [_dp, _c.c_int, _c.c_int, _fp, _fp, _fp, _fp, _fp, _fp]
        L.ref_contour.argtypes = [_ip, _ip]
        L.ref_incident_power.argtypes = [_c.c_float]
        L.ref_incident_power.restype = _c.c_float
        L.ref_fitness.argtypes = [_fp, _c.c_int, _fp]
        L.ref_fitness.restype = _c.c_float

    @staticmethod
    def available(path=_REF):
        return os.path.exists(path)

    def cpml_tables(self):
        out = np.zeros(60, np.float32)
        self.lib.ref_cpml_tables(_ptr(out, _fp))
        return out.reshape(6, 10)

    def materials(self, params162):
        p = np.ascontiguousarray(params162, np.float64)
        n1 = 402 * 402
        arrs = [np.zeros(n1, np.float32) for _ in range(6)]
        self.lib.ref_materials(_ptr(p, _dp), *[_ptr(a, _fp) for a in arrs])
        return [a.reshape(402, 402) for a in arrs]  # eps, sigma, Ceze, Cezhy, Cezeic, Cezeip

    def run(self, params162, phi_index, nsteps, want_d=True):
        p = np.ascontiguousarray(params162, np.float64)
        ez = np.zeros(402 * 402, np.float32)
        hx = np.zeros(401 * 401, np.float32)
        hy = np.zeros_like(hx)
        tr = np.zeros(max(nsteps, 1), np.float32)
        cj = np.zeros(2 * 10 * (2 * 199 + 2 * 201) * 2, np.float32)
        D = np.zeros(1000, np.float32) if want_d else None
        self.lib.ref_run(_ptr(p, _dp), phi_index, nsteps, _ptr(ez, _fp), _ptr(hx, _fp), _ptr(hy, _fp),
                         _ptr(tr, _fp), _ptr(cj, _fp), _ptr(D, _fp))
        return dict(ez=ez.reshape(402, 402), hx=hx.reshape(401, 401), hy=hy.reshape(401, 401),
                    trace=tr[:nsteps], cj=cj, D=None if D is None else D.reshape(10, 100))

    def fitness(self, D, measurement):
        """fitness() of FDTD_common.cxx:53-61, compiled from the reference."""
        D = np.ascontiguousarray(D, np.float32).ravel()
        m = np.ascontiguousarray(measurement, np.float32).ravel()
        assert m.size >= D.size
        return float(self.lib.ref_fitness(_ptr(D, _fp), D.size, _ptr(m, _fp)))

    def contour(self):
        xs = np.zeros(800, np.int32)
        ys = np.zeros(800, np.int32)
        self.lib.ref_contour(_ptr(xs, _ip), _ptr(ys, _ip))
        return xs, ys


def unpack_ref_cj(cj, nfreq=10, ny_side=199, nx_side=201):
    """Split the reference's 8 accumulator arrays into J, M ordered like the
    contour index (legs yn, xp, yp, xn; yp and xn already in contour order)."""
    out = []
    o = 0
    sizes = [ny_side, nx_side, ny_side, nx_side, ny_side, nx_side, ny_side, nx_side]
    arrs = []
    for s in sizes:
        a = cj[o:o + 2 * nfreq * s]
        o += 2 * nfreq * s
        arrs.append((a[0::2] + 1j * a[1::2]).reshape(nfreq, s))
    cjzyn, cjzxp, cjzyp, cjzxn, cmxyn, cmyxp, cmxyp, cmyxn = arrs
    J = np.concatenate([cjzyn, cjzxp, cjzyp, cjzxn], axis=1)
    M = np.concatenate([cmxyn, cmyxp, cmxyp, cmyxn], axis=1)
    return J, M
