"""MI355X-native immersed-boundary projection time step (IB_Parallel's advance path).

The product is the C-ABI library ``libibgpu.so`` (include/ib_gpu.h) built from
``csrc/`` with hipcc for gfx950.  This package loads it with ctypes and offers
``GpuSim``, a host-side handle that mirrors the reference's advance/operator
interface.  There is no CPU fallback: if the library or a GPU is missing,
``load()`` raises.
"""
import ctypes as C
import os

from . import abi
from .sim import IBError, Simulation

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libibgpu.so")

_lib = None
_fns = None


def load():
    """Load libibgpu.so (built in-tree by ib_parallel_amd.build)."""
    global _lib, _fns
    if _fns is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ib_parallel_amd.build`")
        _lib = C.CDLL(LIB_PATH)
        _fns = abi.bind(_lib, "ib_", abi.GPU_ONLY)
    return _fns


def library():
    load()
    return _lib


class GpuSim(Simulation):
    """Simulation whose state lives in HBM of the current HIP device."""

    def __init__(self, params, bodies):
        super().__init__(load(), params, bodies, prefix="ib_")

    def schur_info(self):
        """(form, face-set size, G computations): form 1 matrix-free, 2 assembled, 3 face-space."""
        mode, nfaces, builds = C.c_int32(), C.c_int32(), C.c_int64()
        self._check(self._f["schur_info"](self.h, C.byref(mode), C.byref(nfaces), C.byref(builds)))
        return mode.value, nfaces.value, builds.value

    def kernel_timing(self, kclass, enable=True):
        self._check(self._f["kernel_timing"](self.h, kclass, 1 if enable else 0))

    def kernel_timing_read(self, kclass):
        """(launches, total_ms, total algorithmic bytes) of the timed launches."""
        n = C.c_int64()
        ms = C.c_double()
        by = C.c_double()
        self._check(self._f["kernel_timing_read"](self.h, kclass, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value


__all__ = ["GpuSim", "IBError", "Simulation", "abi", "load", "library"]
