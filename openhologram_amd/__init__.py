"""openhologram_amd -- MI355X-native hologram generation (OpenHolo hot path).

The compute path is liboph_hip.so (HIP kernels for gfx950 behind the C ABI of
include/oph_hip.h); this package is a thin ctypes host layer over it, used by
the tests and the benchmark.  C++ callers use include/oph_hip.h directly or
the reference-shaped API of include/openholo_hip.hpp.
"""
from . import scenes
from .lib import (ACCUMULATE, ASYNC, DEVICE_PTR, OUT_C64, OUT_C128, OUT_FRINGE, Context,
                  DeviceArray, OphError, Spec, load, load_keti_points, make_spec,
                  spec_from_keti_config)

__all__ = ["Context", "DeviceArray", "OphError", "Spec", "load", "make_spec",
           "spec_from_keti_config", "load_keti_points", "OUT_C128", "OUT_C64", "OUT_FRINGE",
           "DEVICE_PTR", "ACCUMULATE", "ASYNC"]
