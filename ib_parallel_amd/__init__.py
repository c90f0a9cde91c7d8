"""MI355X-native immersed-boundary projection time step (IB_Parallel's advance path).

The product is the C-ABI library ``libibgpu.so`` (include/ib_gpu.h) built from
``csrc/`` with hipcc for gfx950.  This package loads it with ctypes and offers
``GpuSim``, a host-side handle that mirrors the reference's advance/operator
interface.  There is no CPU fallback: if the library or a GPU is missing,
``load()`` raises.
"""
import ctypes as C
import os

import numpy as np

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

    def __init__(self, params, bodies, subdomain=None):
        super().__init__(load(), params, bodies, prefix="ib_", subdomain=subdomain)

    def schur_info(self):
        """(form, face-set size, G computations): form 1 matrix-free, 2 assembled, 3 face-space."""
        mode, nfaces, builds = C.c_int32(), C.c_int32(), C.c_int64()
        self._check(self._f["schur_info"](self.h, C.byref(mode), C.byref(nfaces), C.byref(builds)))
        return mode.value, nfaces.value, builds.value

    def gmres_noalloc(self, b_dev, x_dev):
        """ib_op_gmres_noalloc on device buffers (addresses of nf doubles; x
        holds the initial guess and receives the solution): (status, its).
        The status is returned, not raised, so a caller can see IB_E_INVAL
        for a workspace above the static maximum."""
        its = C.c_int32()
        rc = self._f["op_gmres_noalloc"](self.h, C.c_void_p(b_dev), C.c_void_p(x_dev), C.byref(its))
        return rc, its.value

    def gmres_info(self):
        """(form, workgroups, exchange) of the dense GMRES solve: form 0 per-stage
        kernels, 1 one workgroup, 2 cooperative (nf <= 256), 3 cooperative for
        larger systems; exchange of the last form-3 solve 1 one XCD's L2, 2
        device scope, 0 none yet."""
        f, p, x = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self._f["gmres_info"](self.h, C.byref(f), C.byref(p), C.byref(x)))
        return f.value, p.value, x.value

    def dst_tiles(self):
        """(rp, cp, source): row pairs per workgroup of the DST row pass, column
        pairs of the column pass, 0 fixed rule / 1 timed / 2 forced (IB_DST_TILES)."""
        rp, cp, src = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self._f["dst_tiles"](self.h, C.byref(rp), C.byref(cp), C.byref(src)))
        return rp.value, cp.value, src.value

    def dst_form(self):
        """(forms, us): forms of the DST passes for (row, column) x (time step,
        batched assembly), 0 Stockham FFT, k > 0 the FP64 matrix-core form in
        configuration k-1; us the timed microseconds per field,
        us[4*batched + 2*column + matrix_core] (negative: not timed)."""
        f = (C.c_int32 * 4)()
        us = np.zeros(8)
        self._check(self._f["dst_form"](self.h, f, us.ctypes.data_as(C.POINTER(C.c_double))))
        return tuple(f), us

    def adjoint(self, nsteps, cx, cy, ckpt_every=None, omega_grad=True, position_grad=False, body_grad=False):
        """Advance nsteps steps and return (J, dJ/dRe, dJ/domega0) for the
        time-mean objective J = mean over the steps of sum_b cx[b] fx_b +
        cy[b] fy_b; dJ/domega0 has the layout of vort2flux's omega (None
        unless omega_grad).  position_grad appends dJ/dxb and dJ/dyb per body
        point (deck order; rigid bodies), body_grad appends dJ/dK and
        dJ/dgamma0 per body (torsional stiffness and pitch angle in radians;
        0 for rigid bodies).  ckpt_every None: the whole run is taped in HBM
        when it fits, else checkpoints every ceil(sqrt(nsteps)) steps.  Rigid
        bodies with num_stat = 1 or GMRES on the assembled Schur matrix (no
        position gradients), or the face-space form (moving bodies)."""
        cx = np.ascontiguousarray(np.broadcast_to(np.asarray(cx, dtype=np.float64), (self.nbodies,)))
        cy = np.ascontiguousarray(np.broadcast_to(np.asarray(cy, dtype=np.float64), (self.nbodies,)))
        K = int(ckpt_every) if ckpt_every else 0
        J, dRe = C.c_double(), C.c_double()
        g = np.empty(self.L * self.P) if omega_grad else None
        npts = self.nf // 2
        gx = np.empty(npts) if position_grad else None
        gy = np.empty(npts) if position_grad else None
        dk = np.empty(self.nbodies) if body_grad else None
        dg = np.empty(self.nbodies) if body_grad else None
        self._check(self._f["adjoint_run"](self.h, int(nsteps), K, abi.dptr(cx), abi.dptr(cy), C.byref(J),
                                           C.byref(dRe), abi.dptr(g), abi.dptr(gx), abi.dptr(gy), abi.dptr(dk),
                                           abi.dptr(dg)))
        self.itime.value += int(nsteps)
        out = [J.value, dRe.value, g]
        if position_grad:
            out += [gx, gy]
        if body_grad:
            out += [dk, dg]
        return tuple(out)

    def hinge_grad(self):
        """(dJ/dxh, dJ/dyh, dJ/dds) per body: the gradient of the last
        adjoint's J with respect to each torsional body's hinge and point
        spacing (0 for the others); needs the last adjoint run with
        position_grad in the face-space form."""
        gx, gy, gs = np.empty(self.nbodies), np.empty(self.nbodies), np.empty(self.nbodies)
        self._check(self._f["adjoint_hinge_grad"](self.h, abi.dptr(gx), abi.dptr(gy), abi.dptr(gs)))
        return gx, gy, gs

    def kernel_timing(self, kclass, enable=True):
        self._check(self._f["kernel_timing"](self.h, kclass, 1 if enable else 0))

    def kernel_timing_read(self, kclass):
        """(launches, total_ms, total algorithmic bytes) of the timed launches."""
        n = C.c_int64()
        ms = C.c_double()
        by = C.c_double()
        self._check(self._f["kernel_timing_read"](self.h, kclass, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value


class GpuBatch:
    """Independent simulations advanced one step together (ib_batch_*).

    ``sims[i]`` is simulation i as a ``GpuSim`` view for state access; it is
    owned by the batch and valid while the batch is open."""

    def __init__(self, cases):
        """cases: list of (params, bodies) pairs."""
        fns = load()
        self._b = abi.bind(_lib, "ib_", abi.BATCH)
        n = len(cases)
        self._params = (abi.IBParams * n)(*[p for p, _ in cases])
        self._nb = (C.c_int32 * n)(*[len(b) for _, b in cases])
        flat = [b for _, bs in cases for b in bs]
        self._bodies_keep = abi.make_bodies(flat)
        h = C.c_void_p()
        rc = self._b["batch_create"](n, self._params, self._nb, self._bodies_keep[0], C.byref(h))
        if rc != abi.IB_OK:
            raise IBError(rc, self._err(None))
        self.h = h
        self.sims = []
        for i in range(n):
            hi = C.c_void_p()
            self._check(self._b["batch_handle"](h, i, C.byref(hi)))
            s = GpuSim.__new__(GpuSim)
            Simulation.__init__(s, fns, cases[i][0], cases[i][1], prefix="ib_", handle=hi)
            self.sims.append(s)
        self.itimes = (C.c_int32 * n)()
        self.infos = (abi.IBStepInfo * n)()

    def _err(self, h):
        buf = C.create_string_buffer(1024)
        self._b["batch_last_error"](h, buf, 1024)
        return buf.value.decode(errors="replace")

    def _check(self, rc, allow=()):
        if rc != abi.IB_OK and rc not in allow:
            raise IBError(rc, self._err(self.h))
        return rc

    @property
    def mode(self):
        """1: fused (one graph launch per step), 2: concurrent (a host thread per simulation)."""
        n, m = C.c_int32(), C.c_int32()
        self._check(self._b["batch_info"](self.h, C.byref(n), C.byref(m)))
        return m.value

    def dst_tiles(self):
        """(row rule, column rule) of the fused form's DST launches: 0 each
        simulation's own tiles, -1 the fixed rule, k > 0 k pairs per workgroup."""
        x, y = C.c_int32(), C.c_int32()
        self._check(self._b["batch_dst_tiles"](self.h, C.byref(x), C.byref(y)))
        return x.value, y.value

    def advance(self, allow_noconv=False):
        rc = self._b["batch_advance"](self.h, self.itimes, self.infos)
        allow = (abi.IB_E_NOCONV_GMRES, abi.IB_E_NOCONV_FSI) if allow_noconv else ()
        self._check(rc, allow)
        for s, t in zip(self.sims, self.itimes):
            s.itime.value = t
        # copies: the records of the next step reuse the same buffer
        return [abi.IBStepInfo.from_buffer_copy(i) for i in self.infos]

    def adjoint(self, nsteps, cx, cy, ckpt_every=None):
        """ib_adjoint_run of every simulation, concurrently: cx, cy are per
        body over all simulations (as the bodies of the cases); returns
        (J per simulation, dJ/dRe per simulation, dJ/dK and dJ/dgamma0 per
        body)."""
        nb = sum(s.nbodies for s in self.sims)
        cx = np.ascontiguousarray(np.broadcast_to(np.asarray(cx, dtype=np.float64), (nb,)))
        cy = np.ascontiguousarray(np.broadcast_to(np.asarray(cy, dtype=np.float64), (nb,)))
        n = len(self.sims)
        J, dre, dk, dg = np.empty(n), np.empty(n), np.empty(nb), np.empty(nb)
        K = int(ckpt_every) if ckpt_every else 0
        self._check(self._b["batch_adjoint"](self.h, int(nsteps), K, abi.dptr(cx), abi.dptr(cy), abi.dptr(J),
                                             abi.dptr(dre), abi.dptr(dk), abi.dptr(dg)))
        for s in self.sims:
            s.itime.value += int(nsteps)
        return J, dre, dk, dg

    def close(self):
        if getattr(self, "h", None):
            for s in self.sims:
                s.h = None
            self._b["batch_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["GpuBatch", "GpuSim", "IBError", "Simulation", "abi", "load", "library"]
