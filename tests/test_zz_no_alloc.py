"""ib_op_gmres_noalloc (the dense GMRES solve with its workspace in a static
device array) against ib_op_gmres on the same handle: the same kernel on
the same matrix and inputs, so the same bits, for each dense kernel (the
cooperative solve for larger systems, the cooperative solve for nf <= 256,
the one-workgroup solve), on calls in a row whose workspaces shrink; and the
calls it refuses."""
import ctypes as C

import numpy as np
import pytest

from ib_parallel_amd import abi, decks

pytestmark = pytest.mark.gpu


class _Dev:
    """A device copy of a float64 array through the HIP runtime that
    libibgpu.so is linked against (its symbols looked up through the
    library's handle: a runtime that another module, such as torch, brought
    into the process would not see the device the library uses)."""
    _hip = None

    def __init__(self, a):
        if _Dev._hip is None:
            from ib_parallel_amd import library
            h = library()
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            _Dev._hip = h
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.n = a.size
        self.p = C.c_void_p()
        assert _Dev._hip.hipMalloc(C.byref(self.p), max(8, 8 * self.n)) == 0
        assert _Dev._hip.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, 1) == 0  # host to device

    def get(self):
        out = np.empty(self.n)
        assert _Dev._hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        _Dev._hip.hipFree(self.p)


def _pair(sim, seed):
    """One ib_op_gmres and one ib_op_gmres_noalloc solve of the same system."""
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(sim.nf)
    x0 = 0.1 * rng.standard_normal(sim.nf)
    x_ref = x0.copy()
    its_ref = C.c_int32()
    rc_ref = sim._f["op_gmres"](sim.h, abi.dptr(b), abi.dptr(x_ref), C.byref(its_ref))
    bd, xd = _Dev(b), _Dev(x0)
    rc, its = sim.gmres_noalloc(bd.p.value, xd.p.value)
    x = xd.get()
    bd.free()
    xd.free()
    print(f"nf {sim.nf} form {sim.gmres_info()}: status {rc_ref} / {rc}, its {its_ref.value} / {its}", flush=True)
    assert rc_ref in (abi.IB_OK, abi.IB_E_NOCONV_GMRES)
    assert (rc, its) == (rc_ref, its_ref.value)
    assert np.array_equal(x, x_ref)
    return x


def test_noalloc_gmres_equals_op_gmres():
    from ib_parallel_amd import GpuSim
    cases = [
        ("airfoil + flap, 1024x512x5, nf 800", decks.airfoil_flap_case(), 3),
        ("two bodies, 40x32, cooperative solve for larger systems",
         decks.tiny_case(m=40, n=32, gmres_form=4, gmres_restart=80), 3),
        ("two bodies, 32x24, cooperative solve", decks.tiny_case(gmres_form=2), 2),
        ("two bodies, 32x24, one workgroup", decks.tiny_case(gmres_form=1), 1),
    ]
    for name, (p, b), form in cases:
        print(name, flush=True)
        s = GpuSim(p, b)
        assert s.advance().status == abi.IB_OK
        assert s.gmres_info()[0] == form, s.gmres_info()
        x1 = _pair(s, 1)
        # the array keeps the last call's contents: a second call starts clean
        x2 = _pair(s, 1)
        assert np.array_equal(x1, x2)
        _pair(s, 2)
        s.close()


def test_noalloc_gmres_refuses():
    from ib_parallel_amd import GpuSim

    def status(sim):
        bd, xd = _Dev(np.zeros(sim.nf)), _Dev(np.zeros(sim.nf))
        rc = sim.gmres_noalloc(bd.p.value, xd.p.value)[0]
        bd.free()
        xd.free()
        return rc

    # the cooperative solve for nf <= 256 at restart 1000 keeps a Hessenberg
    # matrix per workgroup: 4 x 1001 x 1000 doubles, 32 MB of workspace
    s = GpuSim(*decks.tiny_case(gmres_form=2, gmres_restart=1000))
    assert s.gmres_info()[:2] == (2, 4), s.gmres_info()
    assert status(s) == abi.IB_E_INVAL
    assert "static maximum" in s._err(s.h), s._err(s.h)
    s.close()
    s = GpuSim(*decks.tiny_case(gmres_form=3))
    assert status(s) == abi.IB_E_UNSUPPORTED
    s.close()
    s = GpuSim(*decks.tiny_case(schur_mode=1))
    assert status(s) == abi.IB_E_UNSUPPORTED
    s.close()
    # null buffers
    s = GpuSim(*decks.tiny_case())
    its = C.c_int32(7)
    assert s._f["op_gmres_noalloc"](s.h, None, None, C.byref(its)) == abi.IB_E_INVAL and its.value == 0
    s.close()
