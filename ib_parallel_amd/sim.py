"""Host-side handle around a library exporting the ib_* C ABI.

``Simulation`` plays the role of the reference's module-global state plus
``advance`` (src/timestepper.F90:18-556): create once from a parameter set and
body decks, then advance step by step and read fields/forces back.  The same
class drives the GPU library (prefix ``ib_``) and, in tests only, the CPU
oracle (prefix ``orc_``) so parity tests call both with identical arguments.
"""
import ctypes as C

import numpy as np

from . import abi


class IBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Simulation:
    def __init__(self, fns, params, bodies, prefix="ib_", handle=None, subdomain=None):
        """Create a simulation; with `handle`, wrap one that a batch owns; with
        `subdomain` (abi.make_subdomain), the sub-domain approach."""
        self._f = fns
        self.prefix = prefix
        self.params = params
        self.subdomain = subdomain
        self._owned = handle is None
        if handle is None:
            self._bodies_keep = abi.make_bodies(bodies)
            h = C.c_void_p()
            if subdomain is not None:
                rc = fns["create_subdomain"](C.byref(params), C.byref(subdomain), len(bodies), self._bodies_keep[0],
                                             C.byref(h))
            else:
                rc = fns["create"](C.byref(params), len(bodies), self._bodies_keep[0], C.byref(h))
            if rc != abi.IB_OK:
                raise IBError(rc, self._err(None))
        else:
            h = handle
        self.h = h
        sz = abi.IBSizes()
        fns["get_sizes"](h, C.byref(sz))
        self.sizes = sz
        self.m, self.n, self.L = sz.m, sz.n, sz.mgridlev
        self.P, self.Fx, self.Fy, self.nbc = sz.n_nodes, sz.n_xfaces, sz.n_yfaces, sz.nbc
        self.nf, self.nbodies, self.ntors = sz.nf, sz.nbodies, sz.ntors
        self.itime = C.c_int32(params.istart)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                self._f["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, h):
        buf = C.create_string_buffer(1024)
        self._f["last_error"](h, buf, 1024)
        return buf.value.decode(errors="replace")

    def _check(self, rc, allow=()):
        if rc != abi.IB_OK and rc not in allow:
            raise IBError(rc, self._err(self.h))
        return rc

    # -- stepping -----------------------------------------------------------
    def advance(self, allow_noconv=False):
        info = abi.IBStepInfo()
        rc = self._f["advance"](self.h, C.byref(self.itime), C.byref(info))
        allow = (abi.IB_E_NOCONV_GMRES, abi.IB_E_NOCONV_FSI) if allow_noconv else ()
        self._check(rc, allow)
        return info

    # -- state --------------------------------------------------------------
    def level(self, k):
        w = np.empty(self.P); s = np.empty(self.P); qx = np.empty(self.Fx); qy = np.empty(self.Fy)
        ro = np.empty(self.P)
        self._check(self._f["get_level_fields"](self.h, k, abi.dptr(w), abi.dptr(s), abi.dptr(qx),
                                                abi.dptr(qy), abi.dptr(ro)))
        return {"omega": w, "s": s, "qx": qx, "qy": qy, "rhs_old": ro}

    def set_level(self, k, omega=None, s=None, qx=None, qy=None, rhs_old=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        omega, s, qx, qy, rhs_old = map(c, (omega, s, qx, qy, rhs_old))
        self._check(self._f["set_level_fields"](self.h, k, abi.dptr(omega), abi.dptr(s), abi.dptr(qx),
                                                abi.dptr(qy), abi.dptr(rhs_old)))

    def set_body(self, f=None, xb=None, yb=None, theta=None, thetad=None, thetadd=None):
        """Restore body state (restart); the weights are rebuilt from xb, yb."""
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        args = [c(a) for a in (f, xb, yb, theta, thetad, thetadd)]
        self._check(self._f["set_body"](self.h, *[abi.dptr(a) for a in args]))

    def body(self):
        nf, nt = self.nf, max(1, self.ntors)
        out = {k: np.empty(nf) for k in ("f", "f_rdst", "xb", "yb")}
        th = {k: np.zeros(nt) for k in ("theta", "thetad", "thetadd")}
        self._check(self._f["get_body"](self.h, abi.dptr(out["f"]), abi.dptr(out["f_rdst"]), abi.dptr(out["xb"]),
                                        abi.dptr(out["yb"]), abi.dptr(th["theta"]), abi.dptr(th["thetad"]),
                                        abi.dptr(th["thetadd"])))
        for k in th:
            out[k] = th[k][: self.ntors]
        return out

    def forces(self, which=0):
        fx = np.empty(max(1, self.nbodies)); fy = np.empty(max(1, self.nbodies))
        self._check(self._f["get_forces"](self.h, which, abi.dptr(fx), abi.dptr(fy)))
        return fx[: self.nbodies], fy[: self.nbodies]

    # -- single operators ---------------------------------------------------
    def dst2d(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        batch = a.size // self.P
        out = np.empty_like(a)
        self._check(self._f["op_dst2d"](self.h, batch, abi.dptr(a), abi.dptr(out)))
        return out

    def ctci(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64); s = np.empty(self.P)
        self._check(self._f["op_ctci"](self.h, abi.dptr(w), abi.dptr(s)))
        return s

    def ainv(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64); o = np.empty(self.P)
        self._check(self._f["op_ainv"](self.h, abi.dptr(w), abi.dptr(o)))
        return o

    def vort2flux(self, omega):
        w = np.ascontiguousarray(omega, dtype=np.float64).copy()
        s = np.empty(self.L * self.P); qx = np.empty(self.L * self.Fx); qy = np.empty(self.L * self.Fy)
        self._check(self._f["op_vort2flux"](self.h, abi.dptr(w), abi.dptr(s), abi.dptr(qx), abi.dptr(qy)))
        return w, s, qx, qy

    def reg(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64); hx = np.empty(self.Fx); hy = np.empty(self.Fy)
        self._check(self._f["op_reg"](self.h, abi.dptr(f), abi.dptr(hx), abi.dptr(hy)))
        return hx, hy

    def regT(self, hx, hy, f_in=None):
        hx = np.ascontiguousarray(hx, dtype=np.float64); hy = np.ascontiguousarray(hy, dtype=np.float64)
        f_in = np.zeros(self.nf) if f_in is None else np.ascontiguousarray(f_in, dtype=np.float64)
        out = np.empty(self.nf)
        self._check(self._f["op_regT"](self.h, abi.dptr(hx), abi.dptr(hy), abi.dptr(f_in), abi.dptr(out)))
        return out

    def schur_matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64); y = np.empty(self.nf)
        self._check(self._f["op_schur_matvec"](self.h, abi.dptr(x), abi.dptr(y)))
        return y

    def pressure(self):
        """Pressure coefficient on the cells of every level, shape (mgridlev, n, m)."""
        p = np.empty(self.L * self.m * self.n)
        self._check(self._f["pressure"](self.h, abi.dptr(p)))
        return p.reshape(self.L, self.n, self.m)

    def schur_matrix(self):
        """The Schur matrix by columns, S[:, j] = ab_times(e_j) (nf x nf)."""
        S = np.empty((self.nf, self.nf))  # row j of this buffer receives column j
        self._check(self._f["op_schur_matrix"](self.h, abi.dptr(S)))
        return S.T.copy()

    def subdomain_info(self):
        """(sub-domain points, B entries evaluated, B entries kept)."""
        nb, ne, nk = C.c_int32(), C.c_int64(), C.c_int64()
        self._check(self._f["subdomain_info"](self.h, C.byref(nb), C.byref(ne), C.byref(nk)))
        return nb.value, ne.value, nk.value

    def subdomain_matrix(self):
        """lumat = P B^T P^T + body_mat at the current positions (nf x nf)."""
        A = np.empty((self.nf, self.nf))  # row j of this buffer receives column j
        self._check(self._f["op_subdomain_matrix"](self.h, abi.dptr(A)))
        return A.T.copy()

    def beam(self, k=0):
        """State of family-3 body k: u, ud, udd (normal deflection and slope
        per free node)."""
        nd = C.c_int32()
        self._check(self._f["get_beam"](self.h, k, None, None, None, C.byref(nd)))
        u, ud, udd = np.empty(nd.value), np.empty(nd.value), np.empty(nd.value)
        self._check(self._f["get_beam"](self.h, k, abi.dptr(u), abi.dptr(ud), abi.dptr(udd), C.byref(nd)))
        return {"u": u, "ud": ud, "udd": udd}

    def gmres(self, b, x0=None, allow_noconv=False):
        """Solve the Schur system (ab_times) by GMRES from x0; returns (x, its)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.nf) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        its = C.c_int32()
        rc = self._f["op_gmres"](self.h, abi.dptr(b), abi.dptr(x), C.byref(its))
        self._check(rc, (abi.IB_E_NOCONV_GMRES,) if allow_noconv else ())
        return x, its.value

    def nonlinear_rot(self, k, omega, qx, qy, bc):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        omega, qx, qy, bc = map(c, (omega, qx, qy, bc))
        rhs = np.empty(self.P)
        self._check(self._f["op_nonlinear_rot"](self.h, k, abi.dptr(omega), abi.dptr(qx), abi.dptr(qy),
                                                abi.dptr(bc), abi.dptr(rhs)))
        return rhs
