"""MI355X-native 2D TMz FDTD fitness evaluation (microwave breast tomography).

The product is the HIP library libtomo.so (include/tomo.h); this package is
its Python host wrapper.
"""
from ._abi import *  # noqa: F401,F403
from .api import (Context, Sim, TomoError, lib, measurement, preset, request_bytes)  # noqa: F401
