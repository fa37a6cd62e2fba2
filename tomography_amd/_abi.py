"""ctypes mirror of the C ABI declared in include/tomo.h."""
import ctypes as _c

TOMO_OK = 0
TOMO_E_INVALID = -1
TOMO_E_SHORT = -2
TOMO_E_HIP = -3
TOMO_E_NOMEM = -4
TOMO_E_NONFINITE = -5
TOMO_E_CANCELLED = -6
TOMO_E_PENDING = -7
TOMO_E_STATE = -8

MODE_REF_COMPAT = 0
MODE_PHYSICAL = 1

FLAG_FWF_ABS_ORIGIN = 0x1
FLAG_CPU_COEFS = 0x2
FLAG_NO_NAN_CLAMP = 0x4
FLAG_PARAMS_EPS_SIGMA = 0x8
FLAG_PHASE_FITNESS = 0x10
FLAG_DFT_FP64 = 0x20
FLAG_REL_FITNESS = 0x40
FLAG_FITNESS_PER_EXC = 0x80
FLAG_CPU_REF_VARIANT = (FLAG_FWF_ABS_ORIGIN | FLAG_CPU_COEFS | FLAG_NO_NAN_CLAMP | FLAG_PARAMS_EPS_SIGMA
                        | FLAG_REL_FITNESS)

PRESET_REF_GPU = 0
PRESET_REF_CPU = 1
PRESET_CFG1 = 2
PRESET_CFG2 = 3
PRESET_CFG3 = 4
PRESET_CFG4 = 5
PRESET_CFG5 = 6

MAX_MATERIALS = 128
NOALLOC_SCRATCH_BYTES = 64 << 20

FORM_STREAMS = 1
FORM_FUSED = 2
FORM_RESIDENT = 3
FORM_TILE = 4

SEARCH_DE = 1
SEARCH_PSO = 2


class SimConfig(_c.Structure):
    """tomo_sim_config (include/tomo.h)."""

    _fields_ = [
        ("mode", _c.c_int32), ("flags", _c.c_int32), ("nx", _c.c_int32), ("ny", _c.c_int32),
        ("steps", _c.c_int32), ("n_exc", _c.c_int32), ("ncpml", _c.c_int32), ("ntff_offset", _c.c_int32),
        ("n_freq", _c.c_int32), ("n_angles", _c.c_int32), ("numpatches", _c.c_int32), ("n_params", _c.c_int32),
        ("dx", _c.c_double), ("dy", _c.c_double), ("dt", _c.c_double),
        ("f_start", _c.c_double), ("f_stop", _c.c_double),
        ("cpml_order", _c.c_double), ("kappa_max", _c.c_double), ("alpha_max", _c.c_double),
        ("sigma_max", _c.c_double), ("pulse_tau", _c.c_double), ("pulse_t0", _c.c_double),
        ("debye_tau", _c.c_double), ("breast_radius", _c.c_double), ("skin_thickness", _c.c_double),
        ("src_radius", _c.c_double), ("ff_norm", _c.c_double), ("seed", _c.c_uint64),
    ]

    def copy(self, **over):
        c = SimConfig()
        _c.pointer(c)[0] = self
        for k, v in over.items():
            setattr(c, k, v)
        return c

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    def farfield_len(self):
        return self.n_exc * self.n_freq * self.n_angles

    def contour_len(self):
        return 2 * self.nx + 2 * self.ny - 8 * self.ntff_offset - 4
