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


def test_driver_runs_a_beam_deck(tmp_path):
    """A family-3 body deck (the flexible plate as a beam: the five
    parameter lines of family 2, read as pitch, alpha0, mass per unit
    length, EI, damping) through the driver: forces against the oracle on
    the deck as the driver read it back."""
    from oracle import binding as ob
    p, bodies = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, beam=True)
    (tmp_path / "input").mkdir()
    decks.write_ib_inp(tmp_path / "input" / "ib.inp", p, istop=3, isave=3)
    for i, b in enumerate(bodies):
        decks.write_body(tmp_path / "input" / f"body.{i + 1:03d}.inp", b)
    env = dict(os.environ, IB_GMRES_RESTART="80")
    r = subprocess.run([DRIVER], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    p2, b2, _ = decks.read_case(str(tmp_path / "input"))
    assert b2[1].fam == 3 and b2[1].ktheta == pytest.approx(bodies[1].ktheta)
    p2.gmres_restart = 80
    o = ob.OracleSim(p2, b2)
    got = np.loadtxt(tmp_path / "output" / "force01_00.dat")
    for k in range(3):
        info = o.advance()
        assert info.status == 0
        fx, fy = o.forces()
        assert rel(got[k, 1:], np.concatenate([fx, fy])) < 1e-8, k


def test_driver_restart_matches_continuous_run(tmp_path):
    """istart > 0: the driver reads output/data00_<istart>.var (read_variables)
    and the resumed steps reproduce the uninterrupted run's."""
    import shutil
    p, bodies = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12)

    def case(d, istart):
        (d / "input").mkdir(parents=True)
        p.istart = istart
        decks.write_ib_inp(d / "input" / "ib.inp", p, istop=4, isave=2)
        for i, b in enumerate(bodies):
            decks.write_body(d / "input" / f"body.{i + 1:03d}.inp", b)

    env = dict(os.environ, IB_GMRES_RESTART="80")
    a, b = tmp_path / "a", tmp_path / "b"
    case(a, 0)
    r = subprocess.run([DRIVER], cwd=a, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    case(b, 2)
    (b / "output").mkdir()
    shutil.copy(a / "output" / "data00_0000002.var", b / "output")
    r = subprocess.run([DRIVER], cwd=b, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "reading variables" in r.stdout
    fa = np.loadtxt(a / "output" / "force01_00.dat")
    fb = np.loadtxt(b / "output" / "force01_00.dat")
    assert np.array_equal(fb[:, 0], [3, 4])
    assert rel(fb[:, 1:].ravel(), fa[2:, 1:].ravel()) < 1e-11
    va = varfile.read_var(a / "output" / "data00_0000004.var")
    vb = varfile.read_var(b / "output" / "data00_0000004.var")
    for k in range(2):
        for key in ("omega", "s", "qx", "qy"):
            assert rel(vb["levels"][k][key], va["levels"][k][key]) < 1e-11, (k, key)
    assert rel(vb["f"], va["f"]) < 1e-11
    ta = np.loadtxt(a / "output" / "theta01_02.dat")
    tb = np.loadtxt(b / "output" / "theta01_02.dat")
    assert rel(tb[:, 1], ta[2:, 1]) < 1e-11


def test_driver_writes_pressure(tmp_path):
    """compute_pressure = T: output/pressure00_<it>.var on save steps
    (write_pressure layout) equals the oracle's calculate_pressure."""
    from oracle import binding as ob
    p, bodies = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12)
    (tmp_path / "input").mkdir()
    decks.write_ib_inp(tmp_path / "input" / "ib.inp", p, istop=2, isave=2, compute_pressure=True)
    for i, b in enumerate(bodies):
        decks.write_body(tmp_path / "input" / f"body.{i + 1:03d}.inp", b)
    env = dict(os.environ, IB_GMRES_RESTART="80")
    r = subprocess.run([DRIVER], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    pv = varfile.read_pressure(tmp_path / "output" / "pressure00_0000002.var")
    p2, b2, _ = decks.read_case(str(tmp_path / "input"))
    p2.gmres_restart = 80
    o = ob.OracleSim(p2, b2)
    o.advance(); o.advance()
    po = o.pressure()
    assert (pv["m"], pv["n"], len(pv["levels"])) == (40, 32, 2)
    for k in range(2):
        assert rel(pv["levels"][k].ravel(), po[k].ravel()) < 1e-10


def test_driver_runs_a_subdomain_deck(tmp_path):
    # sub_domain = T in ib.inp (the switch of the reference's example deck):
    # the driver creates the handle with the deck's sub-domain parameters
    from ib_parallel_amd import abi
    from oracle import binding as ob
    p, bodies = decks.tiny_case(m=64, n=48, mgridlev=3, tol=1e-10)
    p.sub_domain = 1
    sd = abi.make_subdomain(-0.8, 1.0, -0.45, 0.4, 4.0 / 64, full_rectangular=1)
    (tmp_path / "input").mkdir()
    decks.write_ib_inp(tmp_path / "input" / "ib.inp", p, istop=3, isave=3, subdomain=sd)
    for i, b in enumerate(bodies):
        decks.write_body(tmp_path / "input" / f"body.{i + 1:03d}.inp", b)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    p2, b2, nml = decks.read_case(str(tmp_path / "input"))
    sd2 = decks.subdomain_of(nml)
    assert p2.sub_domain == 1 and sd2.full_rectangular == 1 and sd2.delta_sd == sd.delta_sd
    o = ob.OracleSim(p2, b2, subdomain=sd2)
    forces, iters = [], []
    for _ in range(3):
        info = o.advance()
        fx, fy = o.forces()
        forces.append(np.concatenate([fx, fy]))
        iters.append((info.ksp_iter, info.fsi_iter))
    got = np.loadtxt(tmp_path / "output" / "force01_00.dat")
    for k in range(3):
        assert rel(got[k, 1:], forces[k]) < 1e-8
    it = np.loadtxt(tmp_path / "output" / "iter01.dat")
    assert [tuple(int(v) for v in row[1:]) for row in it] == iters  # ksp_iter 0, as the reference writes it
    # the run wrote input/sd_precomputed.chd; a second run with
    # sub_domain_precomputed = T reads B from it and writes the same forces
    assert (tmp_path / "input" / "sd_precomputed.chd").stat().st_size > 0
    sd.precomputed = 1
    decks.write_ib_inp(tmp_path / "input" / "ib.inp", p, istop=3, isave=3, subdomain=sd)
    first = (tmp_path / "output" / "force01_00.dat").read_bytes()
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert decks.subdomain_of(decks.read_case(str(tmp_path / "input"))[2]).precomputed == 1
    assert (tmp_path / "output" / "force01_00.dat").read_bytes() == first
