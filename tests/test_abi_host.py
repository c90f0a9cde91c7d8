"""C ABI and host logic that run without a GPU."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from ib_parallel_amd import abi, decks
from oracle import binding as ob


def declared(header):
    txt = open(header).read()
    return sorted(set(re.findall(r"\bint\s+(ib_\w+)\s*\(", txt)))


def test_gpu_library_exports_every_declared_symbol():
    from ib_parallel_amd import LIB_PATH
    assert os.path.exists(LIB_PATH), "build libibgpu.so first (python -m ib_parallel_amd.build)"
    lib = C.CDLL(LIB_PATH)
    names = declared(os.path.join(ROOT, "include", "ib_gpu.h"))
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_mirror_of_abi():
    lib = C.CDLL(ob.LIB)
    for n in declared(os.path.join(ROOT, "include", "ib_gpu.h")):
        gpu_only = ("ib_schur_info", "ib_adjoint_run", "ib_adjoint_hinge_grad", "ib_dst_tiles", "ib_dst_form",
                    "ib_gmres_info", "ib_op_gmres_noalloc")
        if n.startswith("ib_kernel_timing") or n in gpu_only or n.startswith("ib_batch_"):
            continue
        assert hasattr(lib, "orc_" + n[3:]), n


def test_struct_layout_matches_header():
    # offsets the Fortran bind(C) module and the C header agree on
    assert C.sizeof(abi.IBParams) == 4 * 4 + 8 * 8 + 4 * 7 + 4 * 5
    assert abi.IBParams.dt.offset == 16 and abi.IBParams.gmres_restart.offset == 80
    assert C.sizeof(abi.IBBody) == 8 + 8 * 5 + 16
    assert C.sizeof(abi.IBStepInfo) == 24


def test_reference_deck_is_read():
    nml = decks.read_ib_inp(os.path.join(GOLDEN, "decks", "airfoil_flap", "ib.inp"))
    assert nml["m"] == 860 and nml["n"] == 860 and nml["mgridlev"] == 5
    assert nml["dt"] == pytest.approx(0.0004375) and nml["re"] == 1000.0
    assert nml["rtol"] == pytest.approx(1e-7) and nml["sub_domain"] is True
    p, bodies, _ = decks.read_case(os.path.join(GOLDEN, "decks", "airfoil_flap"))
    assert [b.fam for b in bodies] == [1, 2, 2, 2]
    assert bodies[0].npts == 285 and bodies[1].npts == 29
    assert bodies[1].gamma0 == pytest.approx(-13.5619092638379701 * math.pi / 180)
    assert bodies[1].itheta == pytest.approx(0.001) and bodies[1].ktheta == pytest.approx(0.001)
    assert p.sub_domain == 1


def test_deck_roundtrip(tmp_path):
    p, bodies = decks.plate_case()
    decks.write_ib_inp(tmp_path / "ib.inp", p, istop=5)
    for i, b in enumerate(bodies):
        decks.write_body(tmp_path / f"body.{i + 1:03d}.inp", b)
    p2, b2, nml = decks.read_case(str(tmp_path))
    assert (p2.m, p2.n, p2.mgridlev) == (p.m, p.n, p.mgridlev) and nml["istop"] == 5
    assert p2.dt == p.dt and p2.len == p.len
    assert np.allclose(b2[0].x, bodies[0].x, atol=1e-15) and b2[0].gamma0 == pytest.approx(bodies[0].gamma0)


def test_create_errors_on_the_oracle():
    with pytest.raises(Exception) as e:
        ob.OracleSim(abi.make_params(30, 32, 2, 0.01, 100, 1.0, 0.5, 0.5), [])
    assert e.value.code == abi.IB_E_INVAL                      # grid not divisible by 4
    p, bodies, _ = decks.read_case(os.path.join(GOLDEN, "decks", "airfoil_flap"))
    with pytest.raises(Exception) as e:
        ob.OracleSim(p, bodies)
    assert e.value.code == abi.IB_E_UNSUPPORTED                # sub_domain path not built
    pt, bt = decks.tiny_case()
    with pytest.raises(Exception) as e:
        ob.OracleSim(pt, bt[::-1])                             # rigid body after torsional one
    assert e.value.code == abi.IB_E_INVAL
    far = abi.BodySpec(np.array([9.0, 9.1]), np.array([0.0, 0.0]))
    with pytest.raises(Exception) as e:
        ob.OracleSim(pt, [far])                                # stencil leaves the grid
    assert e.value.code == abi.IB_E_INVAL


def test_oracle_status_codes_for_nonconvergence():
    p, b = decks.tiny_case(tol=1e-12, gmres_max_it=3)
    s = ob.OracleSim(p, b)
    info = s.advance(allow_noconv=True)
    assert info.status == abi.IB_E_NOCONV_GMRES
    p, b = decks.tiny_case(tol=1e-30, fsi_max_count=1, gmres_restart=60)
    s = ob.OracleSim(p, b)
    info = s.advance(allow_noconv=True)
    assert info.status == abi.IB_E_NOCONV_FSI and info.fsi_iter == 2


def test_oracle_cylinder_drag_positive_lift_symmetric():
    p, b = decks.cylinder_case(m=64, n=64, mgridlev=2, dt=0.01, npts=50, tol=1e-10, gmres_restart=120)
    s = ob.OracleSim(p, b)
    for _ in range(3):
        s.advance()
    fx, fy = s.forces()
    assert fx[0] > 0 and np.isfinite(fx[0])
    assert abs(fy[0]) < 1e-6 * abs(fx[0])
