"""Batched form (ib_batch_*): every simulation of a batch gives exactly the
results of its own plain step, and the plain step is pinned to the oracle."""
import numpy as np
import pytest

from conftest import rel
from ib_parallel_amd import decks

pytestmark = pytest.mark.gpu

# small uneven instances: grids 32..72 cells, 2..4 levels, 20..60 points
SMALL = dict(m_range=(32, 72), levels=(2, 4), npts_range=(20, 60))


def state(sim):
    out = []
    for k in range(1, sim.L + 1):
        lv = sim.level(k)
        out += [lv["omega"], lv["s"], lv["qx"], lv["qy"], lv["rhs_old"]]
    b = sim.body()
    out += [b["f"], b["f_rdst"], b["xb"], b["yb"], b["theta"], b["thetad"], b["thetadd"]]
    return out


def same(a_list, b_list):
    # identical values (signed zeros compare equal)
    return all(a.shape == b.shape and np.array_equal(a, b) for a, b in zip(a_list, b_list))


def test_fused_batch_equals_plain_steps(monkeypatch):
    from ib_parallel_amd import GpuBatch, GpuSim
    # the fused launches run the Stockham DST passes: plain handles in that form round alike
    monkeypatch.setenv("IB_DST_FORM", "fft")
    cases = decks.ellipse_batch_cases(count=5, seed=7, **SMALL)
    assert len({(p.m, p.n, p.mgridlev) for p, _ in cases}) == 5     # uneven sizes
    batch = GpuBatch(cases)
    assert batch.mode == 1
    plain = [GpuSim(p, b) for p, b in cases]
    for step in range(1, 4):
        infos = batch.advance()
        for s in plain:
            s.advance()
        assert [i.itime for i in infos] == [step] * 5 and all(i.status == 0 for i in infos)
        for i in range(5):
            assert same(state(batch.sims[i]), state(plain[i])), (step, i)
    batch.close()


def test_fused_batch_tile_rules_leave_results_unchanged(monkeypatch):
    """The fused batch's DST tile rules (timed over its table launches, or
    the fixed rule, or forced) change only how transforms share a
    workgroup: three batched steps give the same bits under each."""
    from ib_parallel_amd import GpuBatch
    cases = decks.ellipse_batch_cases(count=4, seed=11, **SMALL)
    runs = []
    monkeypatch.setenv("IB_DST_FORM", "fft")  # the Schur assembly in one form for every run
    for env in ({"IB_DST_TUNE": "0"}, {}, {"IB_DST_TILES": "4,4"}):
        monkeypatch.delenv("IB_DST_TUNE", raising=False)
        monkeypatch.delenv("IB_DST_TILES", raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        batch = GpuBatch(cases)
        assert batch.mode == 1
        for _ in range(3):
            batch.advance()
        runs.append((batch.dst_tiles(), [state(s) for s in batch.sims]))
        batch.close()
    assert runs[0][0] == (0, 0) and runs[2][0] == (0, 0)  # the environment's choice stays with the handles
    for rules, st in runs[1:]:
        for i in range(len(cases)):
            assert same(st[i], runs[0][1][i]), (rules, i)


def test_fused_batch_matches_oracle():
    from ib_parallel_amd import GpuBatch
    from oracle import binding as ob
    cases = decks.ellipse_batch_cases(count=3, seed=11, **SMALL)
    batch = GpuBatch(cases)
    orcs = [ob.OracleSim(p, b) for p, b in cases]
    for _ in range(3):
        batch.advance()
        for o in orcs:
            o.advance()
    for g, o in zip(batch.sims, orcs):
        for k in range(1, g.L + 1):
            assert rel(g.level(k)["omega"], o.level(k)["omega"]) < 1e-10
        assert rel(g.body()["f"], o.body()["f"]) < 1e-8
    batch.close()


def test_fused_batch_restart_first_step_flag(monkeypatch):
    # a simulation resumed at istart > 0 takes the AB2 branch on its first
    # batched step, one at istart = 0 the Euler branch, in the same launch
    from ib_parallel_amd import GpuBatch, GpuSim
    monkeypatch.setenv("IB_DST_FORM", "fft")  # the fused launches' form, for the plain reference
    cases = decks.ellipse_batch_cases(count=2, seed=3, **SMALL)
    ref = GpuSim(*cases[1])
    ref.advance()
    snap = [ref.level(k) for k in range(1, ref.L + 1)]
    bref = ref.body()
    p1 = cases[1][0]
    p1.istart = 1
    batch = GpuBatch(cases)
    for k, lv in enumerate(snap, start=1):
        batch.sims[1].set_level(k, **lv)
    batch.sims[1].set_body(f=bref["f"])
    infos = batch.advance()
    assert infos[0].itime == 1 and infos[1].itime == 2
    ref.advance()
    assert same(state(batch.sims[1]), state(ref))
    batch.close()


def test_concurrent_batch_equals_plain_steps():
    # torsional bodies: data-dependent FSI iterations, one host thread each
    from ib_parallel_amd import GpuBatch, GpuSim
    cases = [decks.tiny_case(m=32 + 8 * i, n=24 + 4 * i, mgridlev=2 + i % 2, tol=1e-12) for i in range(3)]
    batch = GpuBatch(cases)
    assert batch.mode == 2
    plain = [GpuSim(p, b) for p, b in cases]
    for _ in range(2):
        infos = batch.advance()
        refs = [s.advance() for s in plain]
        for i in range(3):
            assert (infos[i].ksp_iter, infos[i].fsi_iter) == (refs[i].ksp_iter, refs[i].fsi_iter)
            assert same(state(batch.sims[i]), state(plain[i])), i
    batch.close()


def test_crowded_concurrent_batch_takes_turns_in_cooperative_solves():
    """33 concurrent torsional simulations, each with a cooperative GMRES
    solve of 8 workgroups; simulation i solves on XCD i mod 8.  XCD 0 holds
    five solves (40 workgroups, more than its 32 CUs), which take turns; the
    other XCDs hold four each (32 workgroups), which run at once.  Every
    simulation keeps the solve of its plain step, bit for bit, and its
    workgroups exchange through one XCD's L2."""
    from ib_parallel_amd import GpuBatch, GpuSim
    cases = [decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, gmres_form=4, gmres_restart=80, re=50.0 + i)
             for i in range(33)]
    batch = GpuBatch(cases)
    assert batch.mode == 2
    assert all(s.gmres_info()[:2] == (3, 8) for s in batch.sims)
    plain = [GpuSim(p, b) for p, b in cases]
    for _ in range(2):
        infos = batch.advance()
        refs = [s.advance() for s in plain]
        for i in range(33):
            assert (infos[i].ksp_iter, infos[i].fsi_iter) == (refs[i].ksp_iter, refs[i].fsi_iter), i
            assert same(state(batch.sims[i]), state(plain[i])), i
    assert all(s.gmres_info()[2] == 1 for s in batch.sims)
    batch.close()


def test_batch_create_errors():
    from ib_parallel_amd import GpuBatch, IBError, abi
    good = decks.ellipse_batch_cases(count=1, seed=1, **SMALL)[0]
    bad = (abi.make_params(30, 32, 2, 0.01, 100, 1.0, 0.5, 0.5), [])
    with pytest.raises(IBError) as e:
        GpuBatch([good, bad])
    assert e.value.code == abi.IB_E_INVAL and "simulation 1" in str(e.value)
