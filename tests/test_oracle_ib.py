"""Oracle immersed-boundary coupling and torsional-body algebra.

deltafnc: partition of unity, zero first moment, symmetry and continuity at
r = 1, 2, 3 (the defining properties of the 6-point kernel); the REAL(4)
literal variant satisfies them only to ~1e-6.  reg/regT: exact adjointness,
weights summing to one per force entry.  b_times2: rank one, symmetric, equal
to c BR BR^T / den assembled densely.  Newmark body update with zero fluid
force: the average-acceleration recurrence of a free torsional spring.
GMRES: the solution of the dense Schur matrix assembled column by column from
the matvec (the reference's num_stat preprocess) solved by LAPACK.  Schur
matrix: on one grid level C (Delta^-1 ainv) C^T is symmetric positive definite
in exact arithmetic (Delta^-1 and ainv are diagonal in the same DST basis), so
a transposed or dropped factor breaks symmetry; with more levels the boundary
coupling makes it slightly non-symmetric.  num_stat (precomputed matrix +
Cholesky): equal to tight GMRES on one level; on several levels it solves the
matrix built from the upper triangle.
"""
import numpy as np
import pytest

from conftest import rel
from ib_parallel_amd import decks
from oracle import binding as ob


def phi(r, f32):
    return np.array([ob.deltafnc(float(x), 0.0, 1.0, f32) for x in np.atleast_1d(r)])


@pytest.mark.parametrize("f32,tol", ((False, 1e-8), (True, 3e-6)))
def test_deltafnc_moments(f32, tol):
    for r in np.linspace(0.0, 1.0, 23, endpoint=False):
        k = np.arange(-3, 4)
        x = r + k
        w = phi(np.abs(x), f32)
        assert abs(w.sum() - 1.0) < tol
        assert abs((x * w).sum()) < tol


def test_deltafnc_continuity_and_support():
    # the kernel's coefficients carry 10 significant digits (1e-8 level);
    # the REAL(4) literals leave ~1e-6
    for f32, tol in ((False, 1e-8), (True, 3e-6)):
        for r0 in (1.0, 2.0):
            lo, hi = phi([r0 - 1e-12, r0 + 1e-12], f32)
            assert abs(lo - hi) < tol
        assert abs(phi([3.0 - 1e-12], f32)[0]) < tol
        assert phi([3.0 + 1e-9], f32)[0] == 0.0


def test_deltafnc_f32_literals_change_the_kernel():
    r = np.linspace(0, 2.9, 59)
    a, b = phi(r, False), phi(r, True)
    d = np.abs(a - b)[np.abs(a) > 1e-3] / np.abs(a[np.abs(a) > 1e-3])
    assert 1e-6 < d.max() < 1e-3


@pytest.fixture(scope="module")
def tiny():
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2)
    return ob.OracleSim(p, b)


def test_reg_regT_adjoint(tiny):
    rng = np.random.default_rng(7)
    f = rng.standard_normal(tiny.nf)
    ux, uy = rng.standard_normal(tiny.Fx), rng.standard_normal(tiny.Fy)
    hx, hy = tiny.reg(f)
    lhs = hx @ ux + hy @ uy
    rhs = f @ tiny.regT(ux, uy)
    assert abs(lhs - rhs) < 1e-13 * (np.abs(f) @ np.abs(tiny.regT(np.abs(ux), np.abs(uy))))


def test_reg_weights_sum_to_one(tiny):
    for k in range(tiny.nf):
        e = np.zeros(tiny.nf); e[k] = 1.0
        hx, hy = tiny.reg(e)
        assert abs(hx.sum() + hy.sum() - 1.0) < 3e-6   # f32-literal kernel


def test_regT_of_constant_field(tiny):
    out = tiny.regT(np.full(tiny.Fx, 2.0), np.full(tiny.Fy, 2.0))
    assert np.allclose(out, 2.0, rtol=3e-6)


def test_btimes_rank_one(tiny):
    b = tiny.body()
    n = tiny.nf
    B = np.stack([ob.btimes(tiny, np.eye(n)[i]) for i in range(n)], 1)
    assert np.linalg.matrix_rank(B, tol=1e-12 * np.abs(B).max()) == 1
    assert rel(B, B.T) < 1e-14
    # dense c BR BR^T / den on the torsional body's slice
    body = [bd for bd in decks.tiny_case(m=40, n=32, mgridlev=2)[1] if bd.fam == 2][0]
    npts = body.npts
    p = tiny.params
    d = p.len / p.m
    ds = np.hypot(body.x[1] - body.x[0], body.y[1] - body.y[0])
    j = np.arange(npts)
    BR = np.concatenate([j * ds * np.sin(body.gamma0), -j * ds * np.cos(body.gamma0)])
    den = 4 / p.dt ** 2 * body.itheta + 2 / p.dt * body.ctheta + body.ktheta
    c = (d * d / ds / p.dt) * (2 * ds / p.dt)
    off = n - 2 * npts
    assert rel(B[off:, off:], c * np.outer(BR, BR) / den) < 1e-13


def test_free_spring_newmark(tiny):
    p = tiny.params
    body = [bd for bd in decks.tiny_case(m=40, n=32, mgridlev=2)[1] if bd.fam == 2][0]
    M, K, dt = body.itheta, body.ktheta, p.dt
    th, thd = 0.01, -0.02
    thdd = -K / M * th
    ob.set_theta(tiny, [th], [thd], [thdd])
    for step in range(5):
        ob.body_update(tiny, np.zeros(tiny.nf))
        # average acceleration: th1 (1 + dt^2 K/(4M)) = th + dt thd + dt^2/4 thdd
        th1 = (th + dt * thd + dt * dt / 4 * thdd) / (1 + dt * dt * K / (4 * M))
        thdd1 = -K / M * th1
        thd1 = thd + dt / 2 * (thdd + thdd1)
        th, thd, thdd = th1, thd1, thdd1
        b = tiny.body()
        assert b["theta"][0] == pytest.approx(th, rel=1e-12)
        assert b["thetad"][0] == pytest.approx(thd, rel=1e-10)
        assert b["thetadd"][0] == pytest.approx(thdd, rel=1e-10)


def test_gmres_matches_dense_solve():
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-13, gmres_restart=200)
    sim = ob.OracleSim(p, b)
    n = sim.nf
    A = np.stack([sim.schur_matvec(np.eye(n)[i]) for i in range(n)], 1)
    rhs = A @ np.random.default_rng(8).standard_normal(n)
    x, its = sim.gmres(rhs)
    assert its <= n + 1
    assert rel(x, np.linalg.solve(A, rhs)) < 1e-8
    assert rel(A @ x, rhs) < 1e-12


@pytest.mark.parametrize("torsional", (False, True))
def test_schur_matrix_single_level_spd(torsional):
    p, b = decks.tiny_case(m=40, n=32, mgridlev=1, torsional=torsional)
    sim = ob.OracleSim(p, b)
    S = sim.schur_matrix()
    x = np.random.default_rng(9).standard_normal(sim.nf)
    assert rel(S @ x, sim.schur_matvec(x)) < 1e-14
    assert np.abs(S - S.T).max() < 1e-13 * np.abs(S).max()
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > 0


def test_schur_matrix_multilevel_columns():
    p, b = decks.tiny_case(m=40, n=32, mgridlev=3)
    sim = ob.OracleSim(p, b)
    S = sim.schur_matrix()
    for j in (0, sim.nf // 2, sim.nf - 1):
        assert np.array_equal(S[:, j], sim.schur_matvec(np.eye(sim.nf)[j]))
    ev = np.linalg.eigvals(S)
    assert np.abs(ev.imag).max() < 1e-12 and ev.real.min() > 0


def test_schur_matrix_face_space_factorisation():
    """S = E_F G E_F^T + B, with E_F the regT matrix restricted to the faces the
    weights touch and G the level-1 flux operator C (multi-level CtC)^-1 ainv
    C^T between those faces (unit face sources) -- the factorisation the
    moving-body Schur form is built on."""
    p, b = decks.tiny_case(m=40, n=32, mgridlev=3)
    sim = ob.OracleSim(p, b)
    nF_all = sim.Fx + sim.Fy
    zx, zy = np.zeros(sim.Fx), np.zeros(sim.Fy)
    E = np.empty((sim.nf, nF_all))
    for a in range(nF_all):
        hx, hy = zx.copy(), zy.copy()
        (hx if a < sim.Fx else hy)[a if a < sim.Fx else a - sim.Fx] = 1.0
        E[:, a] = sim.regT(hx, hy)
    F = np.flatnonzero(np.abs(E).max(0) > 0)
    G = np.empty((F.size, F.size))
    for c, a in enumerate(F):
        hx, hy = zx.copy(), zy.copy()
        (hx if a < sim.Fx else hy)[a if a < sim.Fx else a - sim.Fx] = 1.0
        om = np.zeros(sim.L * sim.P)
        om[: sim.P] = sim.ainv(sim.rot(hx, hy))
        _, _, qx, qy = sim.vort2flux(om)
        G[:, c] = np.concatenate([qx[: sim.Fx], qy[: sim.Fy]])[F]
    B = np.stack([ob.btimes(sim, np.eye(sim.nf)[j]) for j in range(sim.nf)], 1)
    EF = E[:, F]
    S = sim.schur_matrix()
    assert rel((EF @ G @ EF.T + B).ravel(), S.ravel()) < 1e-13
    assert np.abs(B).max() > 0  # the torsional block is present


def _upper_sym(S):
    """A(i, j) = S(min(i, j), max(i, j)): the matrix a Cholesky of the upper triangle factors."""
    U = np.triu(S)
    return U + np.triu(S, 1).T


def test_num_stat_single_level_matches_gmres():
    """On one level the a_times matrix is symmetric, so the Cholesky path
    (num_stat) and tight GMRES give the same steps."""
    kw = dict(m=64, n=64, mgridlev=1, dt=0.01, npts=50, tol=1e-13, gmres_restart=120)
    g = ob.OracleSim(*decks.cylinder_case(**kw))
    c = ob.OracleSim(*decks.cylinder_case(num_stat=1, **kw))
    for _ in range(3):
        ig, ic = g.advance(), c.advance()
        assert ic.ksp_iter == 0 and ic.fsi_iter == 0 and ig.ksp_iter > 0
        # GMRES stops at a residual of 1e-13 ||b||: f is fixed to cond(S) * 1e-13 ~ 1e-9
        assert rel(c.body()["f"], g.body()["f"]) < 1e-8
        assert rel(c.level(1)["omega"], g.level(1)["omega"]) < 1e-10


def test_num_stat_solves_the_upper_triangle():
    """On three levels the matrix is slightly non-symmetric; the num_stat
    force solves A f = rhs with A built from the upper triangle (the one
    triangle a Cholesky factorisation reads), rhs = regT(q + q0) = S f_gmres."""
    kw = dict(m=64, n=64, mgridlev=3, dt=0.01, npts=50, tol=1e-14, gmres_restart=200)
    g = ob.OracleSim(*decks.cylinder_case(**kw))
    c = ob.OracleSim(*decks.cylinder_case(num_stat=1, **kw))
    S = g.schur_matrix()
    assert np.abs(S - S.T).max() > 1e-8 * np.abs(S).max()  # the case really is non-symmetric
    g.advance(); c.advance()
    rhs = S @ g.body()["f"]
    assert rel(_upper_sym(S) @ c.body()["f"], rhs) < 1e-9


def test_gmres_restart_is_petsc_gmres30():
    """Restarted GMRES(30) on the same system needs more iterations than full GMRES."""
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-9, gmres_restart=30)
    sim = ob.OracleSim(p, b)
    rhs = sim.schur_matvec(np.random.default_rng(9).standard_normal(sim.nf))
    _, its30 = sim.gmres(rhs)
    p2, b2 = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-9, gmres_restart=200)
    _, its_full = ob.OracleSim(p2, b2).gmres(rhs)
    assert its30 > its_full
