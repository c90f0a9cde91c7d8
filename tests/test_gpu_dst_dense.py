"""The matrix-core DST form (each pass an FP64 MFMA product with the DST-I
matrix S[t][k] = 2 sin(pi (t+1)(k+1)/(N+1))), forced with IB_DST_FORM:
golden vectors and scipy at the large shapes in every tile configuration,
each fused mode through the spectral operators and whole steps against the
oracle, the batched Schur assembly against the Stockham form, the same bits
run to run, and the autotuner's report.  The two forms sum in different
orders, so they agree to rounding, with the tolerances of the Stockham
form's own tests."""
import numpy as np
import pytest

from conftest import rel
from ib_parallel_amd import abi, decks

pytestmark = pytest.mark.gpu

CFGS = (0, 1, 2, 3, 4)  # (T, KS) = (1,4), (1,8), (1,16), (2,2), (2,4): T x T blocks, KS groups


def gpu(p, b):
    from ib_parallel_amd import GpuSim
    return GpuSim(p, b)


@pytest.fixture
def form(monkeypatch):
    """Sets IB_DST_FORM for the handles created after the call."""
    monkeypatch.delenv("IB_DST_TILES", raising=False)
    monkeypatch.delenv("IB_DST_TUNE", raising=False)

    def set_form(v):
        monkeypatch.setenv("IB_DST_FORM", v)
    return set_form


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("shape", ((31, 23), (63, 47), (99, 67)))
def test_dense_dst2d_golden(form, golden, shape, cfg):
    form(f"dense:{cfg}")
    nx, ny = shape
    g = gpu(abi.make_params(nx + 1, ny + 1, 1, 0.01, 100.0, 1.0, 0.5, 0.5), [])
    assert g.dst_form()[0] == (cfg + 1,) * 4
    x = golden[f"d2_in_{nx}x{ny}"]
    assert rel(g.dst2d(x.ravel()), golden[f"d2_out_{nx}x{ny}"].ravel()) < 1e-14
    g.close()


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("m,n", ((200, 200), (400, 400), (1024, 512), (268, 300), (256, 512)))
def test_dense_dst2d_large_vs_scipy(form, m, n, cfg):
    import scipy.fft as sf
    form(f"dense:{cfg}")
    g = gpu(abi.make_params(m, n, 1, 0.01, 100.0, 1.0, 0.5, 0.5), [])
    x = np.random.default_rng(m + n).standard_normal((n - 1, m - 1))
    assert rel(g.dst2d(x.ravel()), sf.dstn(x, type=1).ravel()) < 5e-15 * np.log2(m * n)
    g.close()


@pytest.mark.parametrize("cfg", CFGS)
def test_dense_spectral_ops_match_oracle(form, cfg):
    """ctci (apply_bc on the coarser levels through vort2flux), ainv, and the
    Schur product (rot formed in the row pass's load) against the oracle."""
    from oracle import binding as ob
    form(f"dense:{cfg}")
    p, b = decks.tiny_case(m=40, n=32, mgridlev=3, tol=1e-12, gmres_restart=80, schur_mode=1)
    g, o = gpu(p, b), ob.OracleSim(p, b)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(g.P)
    assert rel(g.ctci(w), o.ctci(w)) < 1e-13
    assert rel(g.ainv(w), o.ainv(w)) < 1e-13
    W = rng.standard_normal(g.L * g.P)
    for a, c in zip(g.vort2flux(W), o.vort2flux(W)):
        assert rel(a, c) < 1e-12
    x = rng.standard_normal(g.nf)
    assert rel(g.schur_matvec(x), o.schur_matvec(x)) < 1e-12
    g.close()
    o.close()


@pytest.mark.parametrize("schur_mode", (1, 3))
def test_dense_steps_match_oracle(form, schur_mode):
    """Three FSI steps (predictor with its two-input column product, ainv of
    rot, the corrector's subtraction, the multi-level vort2flux) against the
    oracle, matrix-free and face-space."""
    from oracle import binding as ob
    form("dense")
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, gmres_restart=80, schur_mode=schur_mode)
    g, o = gpu(p, b), ob.OracleSim(p, b)
    assert all(f > 0 for f in g.dst_form()[0])
    for step in range(3):
        ig, io = g.advance(), o.advance()
        assert ig.status == io.status == abi.IB_OK
        assert abs(ig.fsi_iter - io.fsi_iter) <= 1, step
        for k in range(1, g.L + 1):
            assert rel(g.level(k)["omega"], o.level(k)["omega"]) < 1e-10, (step, k)
        assert rel(g.body()["f"], o.body()["f"]) < 1e-8, step
    g.close()
    o.close()


def test_dense_schur_assembly_matches_stockham(form):
    """The assembled Schur matrix (batched column passes, hundreds of
    columns per launch) from the two forms."""
    p, b = decks.tiny_case(m=64, n=48, mgridlev=3, tol=1e-12, gmres_restart=80, schur_mode=2)
    form("fft")
    g0 = gpu(p, b)
    assert g0.dst_form()[0] == (0, 0, 0, 0)
    S0 = g0.schur_matrix()
    g0.close()
    for cfg in CFGS:
        form(f"dense:{cfg}")
        g1 = gpu(p, b)
        S1 = g1.schur_matrix()
        g1.close()
        assert rel(S1, S0) < 1e-12, cfg


def test_dense_form_same_bits_run_to_run(form):
    """The split sums are added in a fixed order: two handles give the same
    bits, in every configuration."""
    p, b = decks.tiny_case(m=64, n=48, mgridlev=2, tol=1e-12, gmres_restart=80)
    for cfg in CFGS:
        form(f"dense:{cfg}")
        runs = []
        for _ in range(2):
            g = gpu(p, b)
            for _ in range(2):
                assert g.advance().status == abi.IB_OK
            runs.append((np.concatenate([g.level(k)["omega"] for k in range(1, g.L + 1)]), g.body()["f"].copy()))
            g.close()
        assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1]), cfg


def test_autotuner_reports_both_forms(form):
    """Timed at create: both forms are measured for the one-field class at a
    small grid (where the matrix-core form may win), and the reported choice
    follows the 10% rule from the reported times."""
    form("auto")
    p, b = decks.tiny_case(m=64, n=48, mgridlev=2)
    g = gpu(p, b)
    forms, us = g.dst_form()
    print("forms", forms, "us per field", np.round(us, 2))
    assert us[0] > 0 and us[2] > 0 and us[4] > 0 and us[6] > 0
    for q in range(4):
        fft_us, dense_us = us[2 * q], us[2 * q + 1]
        if dense_us < 0:
            assert forms[q] == 0
        else:
            assert (forms[q] > 0) == (dense_us < 0.9 * fft_us), q
    assert g.dst_tiles()[2] == 1
    g.close()
    form("fft")
    g = gpu(p, b)
    forms, us = g.dst_form()
    assert forms == (0, 0, 0, 0) and all(us[1::2] < 0)
    g.close()


def test_bad_form_is_refused(form):
    from ib_parallel_amd import IBError
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2)
    for v in ("dense:5", "mfma", "dense:x"):
        form(v)
        with pytest.raises(IBError):
            gpu(p, b)
