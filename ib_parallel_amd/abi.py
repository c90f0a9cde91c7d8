"""ctypes mirror of include/ib_gpu.h and a thin binder for libraries that export it.

The structs follow the reference's input model: ``ib_params`` mirrors namelist
READ_PARAMETERS (src/parameters.F90:55-59) and ``ib_body`` one body deck
input/body.NNN.inp (src/grid.F90:252-288).
"""
import ctypes as C

import numpy as np

IB_OK = 0
IB_E_INVAL = 1
IB_E_NOCONV_GMRES = 2
IB_E_NOCONV_FSI = 3
IB_E_HIP = 4
IB_E_UNSUPPORTED = 5

STATUS_NAMES = {
    IB_OK: "IB_OK",
    IB_E_INVAL: "IB_E_INVAL",
    IB_E_NOCONV_GMRES: "IB_E_NOCONV_GMRES",
    IB_E_NOCONV_FSI: "IB_E_NOCONV_FSI",
    IB_E_HIP: "IB_E_HIP",
    IB_E_UNSUPPORTED: "IB_E_UNSUPPORTED",
}


class IBParams(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("mgridlev", C.c_int32), ("istart", C.c_int32),
        ("dt", C.c_double), ("re", C.c_double), ("len", C.c_double),
        ("offsetx", C.c_double), ("offsety", C.c_double),
        ("fsi_tol", C.c_double), ("atol", C.c_double), ("rtol", C.c_double),
        ("gmres_restart", C.c_int32), ("gmres_max_it", C.c_int32), ("fsi_max_count", C.c_int32),
        ("deltafnc_f32_literals", C.c_int32), ("num_stat", C.c_int32), ("sub_domain", C.c_int32),
        ("launch_mode", C.c_int32), ("schur_mode", C.c_int32),
        ("face_margin", C.c_int32), ("gmres_form", C.c_int32), ("reserved", C.c_int32 * 2),
    ]


class IBBody(C.Structure):
    _fields_ = [
        ("npts", C.c_int32), ("fam", C.c_int32),
        ("gamma0", C.c_double), ("alpha0", C.c_double),
        ("itheta", C.c_double), ("ktheta", C.c_double), ("ctheta", C.c_double),
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
    ]


class IBSubdomain(C.Structure):
    _fields_ = [
        ("sdxi", C.c_double), ("sdxe", C.c_double), ("sdyi", C.c_double), ("sdye", C.c_double),
        ("delta_sd", C.c_double), ("full_rectangular", C.c_int32), ("precomputed", C.c_int32),
        ("droptol", C.c_double), ("interp_points", C.c_int32), ("support_nz", C.c_int32),
    ]


def make_subdomain(sdxi, sdxe, sdyi, sdye, delta_sd, full_rectangular=0, precomputed=0, droptol=0.0,
                   interp_points=0, support_nz=0):
    """Sub-domain parameters of ib.inp (0 in the last three: the reference's values)."""
    sd = IBSubdomain()
    sd.sdxi, sd.sdxe, sd.sdyi, sd.sdye, sd.delta_sd = sdxi, sdxe, sdyi, sdye, delta_sd
    sd.full_rectangular, sd.precomputed = int(full_rectangular), int(precomputed)
    sd.droptol, sd.interp_points, sd.support_nz = droptol, interp_points, support_nz
    return sd


class IBStepInfo(C.Structure):
    _fields_ = [
        ("itime", C.c_int32), ("ksp_iter", C.c_int32), ("fsi_iter", C.c_int32),
        ("status", C.c_int32), ("err_fsi", C.c_double),
    ]


class IBSizes(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("mgridlev", C.c_int32), ("nbc", C.c_int32),
        ("n_nodes", C.c_int32), ("n_xfaces", C.c_int32), ("n_yfaces", C.c_int32),
        ("nb", C.c_int32), ("nf", C.c_int32), ("nbodies", C.c_int32), ("ntors", C.c_int32),
    ]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)

# name -> (restype, argtypes); the prefix ("ib_" or "orc_") is added by bind()
SIGNATURES = {
    "create": (C.c_int, [C.POINTER(IBParams), C.c_int32, C.POINTER(IBBody), C.POINTER(_P)]),
    "destroy": (C.c_int, [_P]),
    "get_sizes": (C.c_int, [_P, C.POINTER(IBSizes)]),
    "last_error": (C.c_int, [_P, C.c_char_p, C.c_int32]),
    "advance": (C.c_int, [_P, _I, C.POINTER(IBStepInfo)]),
    "get_level_fields": (C.c_int, [_P, C.c_int32, _D, _D, _D, _D, _D]),
    "set_level_fields": (C.c_int, [_P, C.c_int32, _D, _D, _D, _D, _D]),
    "get_body": (C.c_int, [_P, _D, _D, _D, _D, _D, _D, _D]),
    "set_body": (C.c_int, [_P, _D, _D, _D, _D, _D, _D]),
    "get_forces": (C.c_int, [_P, C.c_int32, _D, _D]),
    "op_dst2d": (C.c_int, [_P, C.c_int32, _D, _D]),
    "op_ctci": (C.c_int, [_P, _D, _D]),
    "op_ainv": (C.c_int, [_P, _D, _D]),
    "op_vort2flux": (C.c_int, [_P, _D, _D, _D, _D]),
    "op_reg": (C.c_int, [_P, _D, _D, _D]),
    "op_regT": (C.c_int, [_P, _D, _D, _D, _D]),
    "op_schur_matvec": (C.c_int, [_P, _D, _D]),
    "op_schur_matrix": (C.c_int, [_P, _D]),
    "pressure": (C.c_int, [_P, _D]),
    "op_nonlinear_rot": (C.c_int, [_P, C.c_int32, _D, _D, _D, _D, _D]),
    "op_gmres": (C.c_int, [_P, _D, _D, _I]),
    "get_beam": (C.c_int, [_P, C.c_int32, _D, _D, _D, _I]),
    "create_subdomain": (C.c_int, [C.POINTER(IBParams), C.POINTER(IBSubdomain), C.c_int32, C.POINTER(IBBody),
                                   C.POINTER(_P)]),
    "subdomain_info": (C.c_int, [_P, _I, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "op_subdomain_matrix": (C.c_int, [_P, _D]),
}

GPU_ONLY = {
    "schur_info": (C.c_int, [_P, _I, _I, C.POINTER(C.c_int64)]),
    "dst_tiles": (C.c_int, [_P, _I, _I, _I]),
    "dst_form": (C.c_int, [_P, _I, _D]),
    "adjoint_hinge_grad": (C.c_int, [_P, _D, _D, _D]),
    "gmres_info": (C.c_int, [_P, _I, _I, _I]),
    # b and x are device addresses
    "op_gmres_noalloc": (C.c_int, [_P, _P, _P, _I]),
    "kernel_timing": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "kernel_timing_read": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "adjoint_run": (C.c_int, [_P, C.c_int32, C.c_int32, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
}


# the batched form: GPU library only
BATCH = {
    "batch_create": (C.c_int, [C.c_int32, C.POINTER(IBParams), _I, C.POINTER(IBBody), C.POINTER(_P)]),
    "batch_destroy": (C.c_int, [_P]),
    "batch_advance": (C.c_int, [_P, _I, C.POINTER(IBStepInfo)]),
    "batch_handle": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "batch_info": (C.c_int, [_P, _I, _I]),
    "batch_dst_tiles": (C.c_int, [_P, _I, _I]),
    "batch_last_error": (C.c_int, [_P, C.c_char_p, C.c_int32]),
    "batch_adjoint": (C.c_int, [_P, C.c_int32, C.c_int32, _D, _D, _D, _D, _D, _D]),
}


def bind(lib, prefix, extra=None):
    """Attach restype/argtypes for every SIGNATURES entry present in `lib`."""
    table = dict(SIGNATURES)
    if extra:
        table.update(extra)
    fns = {}
    for name, (res, args) in table.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    return fns


def dptr(a):
    """double* of a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_D)


def make_params(m, n, mgridlev, dt, re, len_, offsetx, offsety, fsi_tol=1e-7, atol=1e-7, rtol=1e-7,
                gmres_restart=30, gmres_max_it=10000, fsi_max_count=1000, deltafnc_f32_literals=1,
                istart=0, num_stat=0, sub_domain=0, launch_mode=0, schur_mode=0, face_margin=0, gmres_form=0):
    p = IBParams()
    p.m, p.n, p.mgridlev, p.istart = m, n, mgridlev, istart
    p.dt, p.re, p.len, p.offsetx, p.offsety = dt, re, len_, offsetx, offsety
    p.fsi_tol, p.atol, p.rtol = fsi_tol, atol, rtol
    p.gmres_restart, p.gmres_max_it, p.fsi_max_count = gmres_restart, gmres_max_it, fsi_max_count
    p.deltafnc_f32_literals, p.num_stat, p.sub_domain = deltafnc_f32_literals, num_stat, sub_domain
    p.launch_mode = launch_mode
    p.schur_mode = schur_mode
    p.face_margin = face_margin
    p.gmres_form = gmres_form
    return p


class BodySpec:
    """One body deck: coordinates plus family-2 parameters (angles in radians)."""

    def __init__(self, x, y, fam=1, gamma0=0.0, alpha0=0.0, itheta=0.0, ktheta=0.0, ctheta=0.0):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        assert self.x.shape == self.y.shape and self.x.ndim == 1
        self.fam, self.gamma0, self.alpha0 = int(fam), float(gamma0), float(alpha0)
        self.itheta, self.ktheta, self.ctheta = float(itheta), float(ktheta), float(ctheta)

    @property
    def npts(self):
        return self.x.size


def make_bodies(specs):
    """ctypes array of IBBody; keep the returned tuple alive while in use."""
    arr = (IBBody * max(1, len(specs)))()
    for i, b in enumerate(specs):
        arr[i].npts, arr[i].fam = b.npts, b.fam
        arr[i].gamma0, arr[i].alpha0 = b.gamma0, b.alpha0
        arr[i].itheta, arr[i].ktheta, arr[i].ctheta = b.itheta, b.ktheta, b.ctheta
        arr[i].x = dptr(b.x)
        arr[i].y = dptr(b.y)
    return arr, specs
