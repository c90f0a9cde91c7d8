"""The PETSc-free Fortran driver (ib_parallel_amd/bin/ib) on the GPU: reads an
ib.inp namelist and body decks, advances through libibgpu.so and writes the
reference's output files; checked against the oracle on the same deck."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel
from ib_parallel_amd import decks, varfile

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "ib_parallel_amd", "bin", "ib")


def test_driver_matches_oracle(tmp_path):
    from oracle import binding as ob
    assert os.path.exists(DRIVER), "build the driver first (python -m ib_parallel_amd.build)"
    p, bodies = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12)
    (tmp_path / "input").mkdir()
    decks.write_ib_inp(tmp_path / "input" / "ib.inp", p, istop=3, isave=3)
    for i, b in enumerate(bodies):
        decks.write_body(tmp_path / "input" / f"body.{i + 1:03d}.inp", b)
    env = dict(os.environ, IB_GMRES_RESTART="80")
    r = subprocess.run([DRIVER], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the oracle on the deck as the driver read it back
    p2, b2, _ = decks.read_case(str(tmp_path / "input"))
    p2.gmres_restart = 80
    o = ob.OracleSim(p2, b2)
    forces, iters = [], []
    for _ in range(3):
        info = o.advance()
        fx, fy = o.forces()
        forces.append(np.concatenate([fx, fy]))
        iters.append((info.ksp_iter, info.fsi_iter))
    got = np.loadtxt(tmp_path / "output" / "force01_00.dat")
    assert got.shape == (3, 1 + 2 * len(bodies))
    assert np.array_equal(got[:, 0], [1, 2, 3])
    for k in range(3):
        assert rel(got[k, 1:], forces[k]) < 1e-10
    it = np.loadtxt(tmp_path / "output" / "iter01.dat")
    assert [tuple(int(v) for v in row[1:]) for row in it] == iters
    th = np.loadtxt(tmp_path / "output" / "theta01_02.dat")
    assert th.shape == (3, 2)
    assert th[-1, 1] == pytest.approx(o.body()["theta"][0], rel=1e-10)
    var = varfile.read_var(tmp_path / "output" / "data00_0000003.var")
    assert (var["m"], var["n"], var["mgridlev"]) == (40, 32, 2)
    for k in (1, 2):
        lv = o.level(k)
        assert rel(var["levels"][k - 1]["omega"], lv["omega"]) < 1e-10
        assert rel(var["levels"][k - 1]["qx"], lv["qx"]) < 1e-10
    assert rel(var["f"], o.body()["f"]) < 1e-10
