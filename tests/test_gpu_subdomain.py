"""Sub-domain approach on the GPU against the oracle (same sparsified B,
same P, same lumat, same steps)."""
import struct

import numpy as np
import pytest

from conftest import rel
from ib_parallel_amd import abi, decks

pytestmark = pytest.mark.gpu

TINY = 1e-300


def sd_params(p):
    q = abi.IBParams.from_buffer_copy(p)
    q.sub_domain = 1
    return q


CASES = {
    # torsional plate interpolated, cylinder points included exactly (no
    # entry dropped: cond(lumat) ~ 2e4, so GMRES runs to 1e-12 for 1e-10 parity)
    "rigid_included": (dict(m=64, n=48, mgridlev=3, tol=1e-12), (0.0, 1.0, -0.4, 0.3, 4.0 / 64, 0, TINY)),
    # every body interpolated from the rectangle, reference drop tolerance
    "full_rectangle": (dict(m=64, n=48, mgridlev=3, tol=1e-10), (-0.8, 1.0, -0.45, 0.4, 4.0 / 64, 1, 0.0)),
}


def make(case):
    from ib_parallel_amd import GpuSim
    from oracle import binding as ob
    kw, (x0, x1, y0, y1, d, full, dtol) = CASES[case]
    p, b = decks.tiny_case(**kw)
    sd = abi.make_subdomain(x0, x1, y0, y1, d, full_rectangular=full, droptol=dtol)
    return GpuSim(sd_params(p), b, subdomain=sd), ob.OracleSim(sd_params(p), b, subdomain=sd)


@pytest.mark.parametrize("case", sorted(CASES))
def test_setup_and_lumat_match_oracle(case):
    g, o = make(case)
    ng, no = g.subdomain_info(), o.subdomain_info()
    assert ng[:2] == no[:2]
    assert abs(ng[2] - no[2]) <= 2  # entries at the drop threshold may round either way
    assert rel(g.subdomain_matrix(), o.subdomain_matrix()) < 1e-12


@pytest.mark.parametrize("case", sorted(CASES))
def test_steps_match_oracle(case):
    g, o = make(case)
    for _ in range(3):
        ig, io = g.advance(), o.advance()
        assert (ig.ksp_iter, ig.fsi_iter) == (io.ksp_iter, io.fsi_iter)
    # per level against the largest vorticity of the run: a coarse level
    # the flow has not reached after three steps holds only rounding noise
    # (~1e-18), whose own relative difference means nothing.  The bound is
    # 1e-10 or ten times the case's tolerance: the full-rectangle case
    # (tolerance 1e-10, no drop tolerance) amplifies rounding differences,
    # measured at 5.6e-11 of the scale with the Stockham DST passes, 3.6e-11
    # with the matrix-core ones and 4.5e-10 with matrix-core rows and
    # Stockham columns (the iteration counts agree in each)
    wg = [g.level(k)["omega"] for k in range(1, g.L + 1)]
    wo = [o.level(k)["omega"] for k in range(1, g.L + 1)]
    scale = max(np.abs(w).max() for w in wo)
    bound = max(1e-10, 10 * CASES[case][0]["tol"])
    for k in range(g.L):
        assert np.abs(wg[k] - wo[k]).max() <= bound * scale, k
    bg, bo = g.body(), o.body()
    assert rel(bg["f"], bo["f"]) < 1e-8
    assert rel(bg["theta"], bo["theta"]) < 1e-10
    assert rel(bg["xb"], bo["xb"]) < 1e-13
    assert rel(g.subdomain_matrix(), o.subdomain_matrix()) < 1e-12  # P follows the moved plate


def test_one_level_rigid_included_equals_plain_step():
    from ib_parallel_amd import GpuSim
    # (the GPU sparsity windows must stay inside the grid)
    p, b = decks.tiny_case(m=64, n=48, mgridlev=1, torsional=False, tol=1e-12)
    sd = abi.make_subdomain(-0.8, 0.0, -0.4, 0.4, 4.0 / 64, full_rectangular=0, droptol=TINY)
    g, pl = GpuSim(sd_params(p), b, subdomain=sd), GpuSim(p, b)
    for _ in range(3):
        g.advance()
        pl.advance()
    assert rel(g.level(1)["omega"], pl.level(1)["omega"]) < 1e-8
    assert rel(g.body()["f"], pl.body()["f"]) < 1e-6


def test_window_at_the_grid_edge_is_rejected():
    from ib_parallel_amd import GpuSim, IBError
    p, b = decks.tiny_case(m=48, n=40, mgridlev=1, torsional=False, tol=1e-12)
    with pytest.raises(IBError) as e:
        GpuSim(sd_params(p), b, subdomain=abi.make_subdomain(-0.8, 0.0, -0.4, 0.4, 4.0 / 48, full_rectangular=0))
    assert e.value.code == abi.IB_E_UNSUPPORTED and "edge of the grid" in str(e.value)


def test_plain_create_rejects_subdomain_flag():
    from ib_parallel_amd import GpuSim, IBError
    p, b = decks.tiny_case(m=32, n=24, mgridlev=2)
    with pytest.raises(IBError) as e:
        GpuSim(sd_params(p), b)
    assert e.value.code == abi.IB_E_UNSUPPORTED


def test_reference_example_deck():
    """The reference's shipped deck (input/ib.inp with body.001/002: rigid
    airfoil + torsional flap, 860x860x5, sub_domain = T) at full size: the
    sub-domain has the airfoil's 287 points and the 136 x 95 rectangle; one
    step converges and leaves q = C s with div q = 0 and C^T q = omega."""
    import os
    from conftest import GOLDEN
    from ib_parallel_amd import GpuSim
    p, b, nml = decks.read_case(os.path.join(GOLDEN, "decks", "example"))
    g = GpuSim(p, b, subdomain=decks.subdomain_of(nml))
    nb_sd, ne, nk = g.subdomain_info()
    assert nb_sd == 287 + 136 * 95 and 0 < nk < ne
    info = g.advance()
    assert info.status == abi.IB_OK and info.err_fsi < p.fsi_tol and info.ksp_iter == 0
    m, n = g.m, g.n
    for k in range(1, g.L + 1):
        lv = g.level(k)
        QX = lv["qx"].reshape(n, m + 1); QY = lv["qy"].reshape(n + 1, m)
        div = QX[:, 1:] - QX[:, :-1] + QY[1:, :] - QY[:-1, :]
        assert np.abs(div).max() < 1e-12 * np.abs(lv["s"]).max()
        rot = QY[1:-1, 1:] - QY[1:-1, :-1] - QX[1:, 1:-1] + QX[:-1, 1:-1]
        assert rel(rot.ravel(), lv["omega"]) < 1e-10
    fx, fy = g.forces()
    assert np.all(np.isfinite(fx)) and np.all(np.isfinite(fy)) and fx[0] > 0


def read_chd(path):
    """input/sd_precomputed.chd read by its record layout (Fortran unformatted
    sequential, 4-byte markers; per column a count record, then values and
    1-based rows), independently of the library."""
    raw = open(path, "rb").read()
    cols, o = [], 0

    def i32(at):
        return struct.unpack_from("<i", raw, at)[0]
    while o < len(raw):
        m = i32(o)
        nz = int.from_bytes(raw[o + 4:o + 4 + m], "little", signed=True)
        assert i32(o + 4 + m) == m
        o += 8 + m
        m2 = i32(o)
        assert m2 == nz * (8 + m)
        vals = np.frombuffer(raw, dtype="<f8", count=nz, offset=o + 4)
        rows = np.frombuffer(raw, dtype="<i4" if m == 4 else "<i8", count=nz, offset=o + 4 + 8 * nz)
        assert i32(o + 4 + m2) == m2
        o += 8 + m2
        cols.append((vals.copy(), rows.astype(np.int64)))
    return cols


def write_chd(path, cols, isz=4):
    it = "<i4" if isz == 4 else "<i8"
    with open(path, "wb") as f:
        for vals, rows in cols:
            nz = len(vals)
            f.write(struct.pack("<i", isz) + np.array([nz], it).tobytes() + struct.pack("<i", isz))
            m2 = nz * (8 + isz)
            f.write(struct.pack("<i", m2) + np.asarray(vals, "<f8").tobytes() + np.asarray(rows, it).tobytes() +
                    struct.pack("<i", m2))


def test_precomputed_b_file_round_trip(tmp_path, monkeypatch):
    """sub_domain_precomputed: a handle that computes B writes
    input/sd_precomputed.chd in the reference's record layout (where the
    input folder exists); one created with the flag reads it back instead of
    computing B and steps bit for bit alike.  The values really come from
    the file: B scaled by 2 and 3 (rows reversed, 8-byte integers as a
    64-bit-index PETSc writes them) moves lumat linearly.  A damaged or
    missing file is an error."""
    from ib_parallel_amd import GpuSim, IBError
    monkeypatch.chdir(tmp_path)
    (tmp_path / "input").mkdir()
    kw, (x0, x1, y0, y1, d, full, dtol) = CASES["full_rectangle"]
    p, b = decks.tiny_case(**kw)

    def sim(pre):
        sd = abi.make_subdomain(x0, x1, y0, y1, d, full_rectangular=full, droptol=dtol, precomputed=pre)
        return GpuSim(sd_params(p), b, subdomain=sd)
    g0 = sim(0)
    nb_sd, nev, nkept = g0.subdomain_info()
    path = tmp_path / "input" / "sd_precomputed.chd"
    cols = read_chd(path)
    assert len(cols) == 2 * nb_sd and sum(len(v) for v, _ in cols) == nkept < nev
    for j, (v, r) in enumerate(cols):
        assert np.all(np.diff(r) > 0) and r[0] >= 1 and r[-1] <= 2 * nb_sd
        assert j + 1 in r  # the diagonal is always kept
    g1 = sim(1)
    assert g1.subdomain_info() == (nb_sd, 0, nkept)
    L0 = g0.subdomain_matrix()
    assert np.array_equal(g1.subdomain_matrix(), L0)
    for _ in range(2):
        i0, i1 = g0.advance(), g1.advance()
        assert (i0.ksp_iter, i0.fsi_iter, i0.status) == (i1.ksp_iter, i1.fsi_iter, i1.status)
    for k in range(1, g0.L + 1):
        assert np.array_equal(g0.level(k)["omega"], g1.level(k)["omega"]), k
    assert np.array_equal(g0.body()["f"], g1.body()["f"])
    g0.close()
    g1.close()
    scaled = []
    for c in (2.0, 3.0):
        write_chd(path, [(c * v[::-1], r[::-1]) for v, r in cols], isz=8)
        g = sim(1)
        scaled.append(g.subdomain_matrix())
        g.close()
    # lumat = body_mat + P B^T P^T: L3 - L1 = 2 (L2 - L1)
    assert rel((scaled[1] - L0).ravel(), (2 * (scaled[0] - L0)).ravel()) < 1e-12
    write_chd(path, cols)
    path.write_bytes(path.read_bytes()[:-10])
    with pytest.raises(IBError) as e:
        sim(1)
    assert e.value.code == abi.IB_E_INVAL and "sd_precomputed.chd" in str(e.value)
    path.unlink()
    with pytest.raises(IBError) as e:
        sim(1)
    assert e.value.code == abi.IB_E_INVAL
