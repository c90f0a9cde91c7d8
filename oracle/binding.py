"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ib_parallel_amd import abi
from ib_parallel_amd.sim import Simulation

HERE = os.path.dirname(os.path.abspath(__file__))
# IB_ORACLE_LIB selects another build of the same sources (the sanitizer
# build, liboracle_san.so, in a process that preloads its runtimes)
LIB = os.environ.get("IB_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_lib = None
_fns = None

EXTRA = {
    "op_rot": (C.c_int, [C.c_void_p, abi._D, abi._D, abi._D]),
    "op_curl": (C.c_int, [C.c_void_p, abi._D, abi._D, abi._D, abi._D]),
    "op_coarsify": (C.c_int, [C.c_void_p, abi._D, abi._D]),
    "op_get_bc": (C.c_int, [C.c_void_p, abi._D, C.c_double, abi._D]),
    "op_apply_bc": (C.c_int, [C.c_void_p, abi._D, abi._D, C.c_double]),
    "op_weights": (C.c_int, [C.c_void_p, abi._D, abi._I, abi._I]),
    "op_ddti": (C.c_int, [C.c_void_p, abi._D, abi._D]),
    "op_coarsify_pressure": (C.c_int, [C.c_void_p, abi._D, abi._D]),
    "deltafnc": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int32]),
    "dst1d": (C.c_int, [C.c_int32, C.c_int32, abi._D, abi._D]),
    "op_btimes": (C.c_int, [C.c_void_p, abi._D, abi._D]),
    "set_theta": (C.c_int, [C.c_void_p, abi._D, abi._D, abi._D]),
    "op_body_update": (C.c_int, [C.c_void_p, abi._D, abi._D]),
    "set_threads": (C.c_int, [C.c_int32]),
    "get_threads": (C.c_int, []),
    "op_subdomain_interp": (C.c_int, [C.c_void_p, abi._I, abi._D, abi._D, abi._D]),
    "op_predictor": (C.c_int, [C.c_void_p, C.c_int32, abi._D, abi._D, abi._D, abi._D, abi._D]),
    "op_redistribute": (C.c_int, [C.c_void_p, abi._D, abi._D]),
    "op_beam_matrices": (C.c_int, [C.c_void_p, C.c_int32, abi._D, abi._D, abi._D]),
    "op_beam_newmark": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, abi._D, abi._D, abi._D, abi._D]),
}


def build():
    """Compile liboracle.so in place (make; g++)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib, _fns
    if _fns is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _fns = abi.bind(_lib, "orc_", EXTRA)
    return _fns


class OracleSim(Simulation):
    """Simulation on the CPU restatement, plus the oracle-only operators."""

    def __init__(self, params, bodies, subdomain=None):
        super().__init__(load(), params, bodies, prefix="orc_", subdomain=subdomain)

    def subdomain_interp(self, npts_sd, np_=4):
        """(indices (nf, np), weights (nf, np), sub-domain x, y) of P."""
        import numpy as np
        idx = np.empty(self.nf * np_, dtype=np.int32); w = np.empty(self.nf * np_)
        sx = np.empty(npts_sd); sy = np.empty(npts_sd)
        self._check(self._f["op_subdomain_interp"](self.h, idx.ctypes.data_as(abi._I), abi.dptr(w), abi.dptr(sx),
                                                   abi.dptr(sy)))
        return idx.reshape(self.nf, np_), w.reshape(self.nf, np_), sx, sy

    def predictor(self, k, omega, rhs, rhs_old, rhsbc=None):
        """fluid_predictor of level k: the new vorticity from omega and the
        nonlinear terms rhs (now) and rhs_old (previous step)."""
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        rhsbc = np.zeros(self.P) if rhsbc is None else rhsbc
        omega, rhs, rhs_old, rhsbc = map(c, (omega, rhs, rhs_old, rhsbc))
        out = np.empty(self.P)
        d = abi.dptr
        self._check(self._f["op_predictor"](self.h, k, d(omega), d(rhs), d(rhs_old), d(rhsbc), d(out)))
        return out

    def beam_matrices(self, k=0):
        """(M, K, H) of beam k: mass and stiffness (nd x nd) and the shape
        functions at the points (np x nd)."""
        nd = self.beam(k)["u"].size
        npts = nd + 1
        M, K, H = np.empty((nd, nd)), np.empty((nd, nd)), np.empty((npts, nd))
        self._check(self._f["op_beam_matrices"](self.h, k, abi.dptr(M), abi.dptr(K), abi.dptr(H)))
        return M, K, H

    def beam_newmark(self, k, nsteps, fb, u, ud, udd):
        """nsteps Newmark steps of beam k alone under fixed point forces fb
        (x forces, then y forces); returns (u, ud, udd)."""
        c = lambda a: np.array(a, dtype=np.float64, copy=True)
        fb, u, ud, udd = map(c, (fb, u, ud, udd))
        self._check(self._f["op_beam_newmark"](self.h, k, int(nsteps), abi.dptr(fb), abi.dptr(u), abi.dptr(ud),
                                               abi.dptr(udd)))
        return u, ud, udd

    def redistribute(self, f):
        """f_rdst = E M^-1 E^T f at the current weights, M = diag(E^T 1)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.empty(self.nf)
        self._check(self._f["op_redistribute"](self.h, abi.dptr(f), abi.dptr(out)))
        return out

    def rot(self, hx, hy):
        import numpy as np
        v = np.empty(self.P)
        self._check(self._f["op_rot"](self.h, abi.dptr(np.ascontiguousarray(hx, float)),
                                      abi.dptr(np.ascontiguousarray(hy, float)), abi.dptr(v)))
        return v

    def curl(self, s, sbc):
        import numpy as np
        qx = np.empty(self.Fx); qy = np.empty(self.Fy)
        self._check(self._f["op_curl"](self.h, abi.dptr(np.ascontiguousarray(s, float)),
                                       abi.dptr(np.ascontiguousarray(sbc, float)), abi.dptr(qx), abi.dptr(qy)))
        return qx, qy

    def coarsify(self, fine, coarse):
        import numpy as np
        out = np.array(coarse, dtype=np.float64, copy=True)
        self._check(self._f["op_coarsify"](self.h, abi.dptr(np.ascontiguousarray(fine, float)), abi.dptr(out)))
        return out

    def get_bc(self, coarse, fac):
        import numpy as np
        bc = np.full(self.nbc, np.nan)
        self._check(self._f["op_get_bc"](self.h, abi.dptr(np.ascontiguousarray(coarse, float)), fac, abi.dptr(bc)))
        return bc

    def apply_bc(self, vort, bc, fac):
        import numpy as np
        out = np.array(vort, dtype=np.float64, copy=True)
        self._check(self._f["op_apply_bc"](self.h, abi.dptr(out), abi.dptr(np.ascontiguousarray(bc, float)), fac))
        return out

    def ddti(self, rhs):
        import numpy as np
        out = np.empty(self.m * self.n)
        self._check(self._f["op_ddti"](self.h, abi.dptr(np.ascontiguousarray(rhs, float).ravel()), abi.dptr(out)))
        return out.reshape(self.n, self.m)

    def coarsify_pressure(self, fine, coarse=None):
        import numpy as np
        out = np.zeros(self.m * self.n) if coarse is None else np.array(coarse, dtype=np.float64).ravel()
        self._check(self._f["op_coarsify_pressure"](self.h, abi.dptr(np.ascontiguousarray(fine, float).ravel()),
                                                    abi.dptr(out)))
        return out.reshape(self.n, self.m)

    def weights(self):
        import numpy as np
        w = np.empty(49 * self.nf)
        # x entries and y entries each get 49 taps
        b = self.body()
        nfx = int(round(self.nf / 2))
        ix = np.empty(49 * self.nf, dtype=np.int32); iy = np.empty(49 * self.nf, dtype=np.int32)
        self._check(self._f["op_weights"](self.h, abi.dptr(w), ix.ctypes.data_as(abi._I), iy.ctypes.data_as(abi._I)))
        return w.reshape(self.nf, 49), ix, iy, b


def _arr(a):
    import numpy as np
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def btimes(sim, x):
    import numpy as np
    y = np.empty(sim.nf)
    sim._check(sim._f["op_btimes"](sim.h, abi.dptr(_arr(x)), abi.dptr(y)))
    return y


def set_theta(sim, theta=None, thetad=None, thetadd=None):
    sim._check(sim._f["set_theta"](sim.h, abi.dptr(_arr(theta)), abi.dptr(_arr(thetad)), abi.dptr(_arr(thetadd))))


def body_update(sim, f):
    import numpy as np
    d = np.empty(max(1, sim.ntors))
    sim._check(sim._f["op_body_update"](sim.h, abi.dptr(_arr(f)), abi.dptr(d)))
    return d[: sim.ntors]


def deltafnc(x1, x2, dr, f32=True):
    return load()["deltafnc"](x1, x2, dr, 1 if f32 else 0)


def dst1d(x):
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    N = x.shape[-1]
    cnt = x.size // N
    out = np.empty_like(x)
    load()["dst1d"](N, cnt, abi.dptr(x), abi.dptr(out))
    return out


def set_threads(n):
    load()["set_threads"](int(n))
