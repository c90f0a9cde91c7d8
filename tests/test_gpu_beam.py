"""Family-3 bodies (a flexible plate as a linear Euler-Bernoulli beam) on
the GPU: steps against the oracle on a small grid, the flexible-plate
workload at its full size (400x400 cells, 4 levels, 101 points, 50
elements), and the configurations the beam is refused in."""
import numpy as np
import pytest

from conftest import rel
from ib_parallel_amd import abi, decks

pytestmark = pytest.mark.gpu


def test_beam_steps_match_oracle():
    from ib_parallel_amd import GpuSim
    from oracle import binding as ob
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, gmres_restart=80, beam=True)
    g, o = GpuSim(p, b), ob.OracleSim(p, b)
    assert g.schur_info()[0] == 3
    for step in range(3):
        ig, io = g.advance(), o.advance()
        assert ig.status == io.status == abi.IB_OK
        assert ig.fsi_iter == io.fsi_iter, (step, ig.fsi_iter, io.fsi_iter)
        for k in range(1, g.L + 1):
            assert rel(g.level(k)["omega"], o.level(k)["omega"]) < 1e-10, (step, k)
        assert rel(g.body()["f"], o.body()["f"]) < 1e-8, step
        bg, bo = g.beam(0), o.beam(0)
        for key in ("u", "ud", "udd"):
            assert rel(bg[key], bo[key]) < 1e-9, (step, key)
        assert rel(g.body()["yb"], o.body()["yb"]) < 1e-12, step
    g.close()
    o.close()


def test_flexible_plate_full_size():
    """The strongly coupled flexible plate: 101 points, 50 elements, mass
    ratio 0.2, K_B = 0.1, pitch 10 deg, Re 200, 400x400x4, dt 1e-3, tolerance
    1e-10.  After three steps from rest: converged FSI, field invariants on
    every level, and the plate's points moving with the beam (E(q + q0) =
    delta n (H ud) to the GMRES tolerance)."""
    from ib_parallel_amd import GpuSim
    from oracle import binding as ob
    p, b = decks.beam_plate_case()
    g = GpuSim(p, b)
    H = ob.OracleSim(p, b).beam_matrices()[2]
    d = p.len / p.m
    infos = [g.advance() for _ in range(3)]
    print("GMRES / FSI iterations:", [(i.ksp_iter, i.fsi_iter) for i in infos])
    assert all(i.status == abi.IB_OK for i in infos)
    m, n = g.m, g.n
    for k in range(1, g.L + 1):
        lv = g.level(k)
        QX = lv["qx"].reshape(n, m + 1)
        QY = lv["qy"].reshape(n + 1, m)
        div = QX[:, 1:] - QX[:, :-1] + QY[1:, :] - QY[:-1, :]
        assert np.abs(div).max() < 1e-12 * np.abs(lv["s"]).max(), k
        rot = QY[1:-1, 1:] - QY[1:-1, :-1] - QX[1:, 1:-1] + QX[:-1, 1:-1]
        assert rel(rot.ravel(), lv["omega"]) < 1e-10, k
    gam = b[0].gamma0
    nx, ny = -np.sin(gam), np.cos(gam)
    st = g.beam(0)
    vn = H @ st["ud"]
    lv = g.level(1)
    e = g.regT(lv["qx"] + d, lv["qy"])
    npts = b[0].npts
    scale = np.abs(d * vn).max()
    print(f"tip deflection {H[-1] @ st['u']:.3e}, max normal velocity {np.abs(vn).max():.3e}")
    assert scale > 0
    assert np.abs(e[:npts] - d * nx * vn).max() < 1e-6 * max(scale, d)
    assert np.abs(e[npts:] - d * ny * vn).max() < 1e-6 * max(scale, d)
    g.close()


def test_beam_refused_configurations():
    from ib_parallel_amd import GpuSim, IBError
    for kw in (dict(schur_mode=2), dict(schur_mode=1)):
        p, b = decks.tiny_case(m=40, n=32, mgridlev=2, beam=True, **kw)
        with pytest.raises(IBError) as e:
            GpuSim(p, b)
        assert e.value.code == abi.IB_E_UNSUPPORTED, kw
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, beam=True)
    g = GpuSim(p, b)
    with pytest.raises(IBError) as e:
        g.adjoint(2, [1.0, 1.0], [0.0, 0.0], omega_grad=False)
    assert e.value.code == abi.IB_E_UNSUPPORTED
    g.close()
