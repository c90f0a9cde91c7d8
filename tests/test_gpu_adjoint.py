"""Discrete adjoint (ib_adjoint_run) against central finite differences of the
GPU step, whose parity with the oracle the other GPU tests establish.

J = mean over the steps of sum_b cx[b] fx_b + cy[b] fy_b.  dJ/dRe is checked
against (J(Re + h) - J(Re - h)) / 2h from fresh handles started from the same
state; dJ/domega0 along random directions against the same difference in the
initial vorticity.  Bar: relative error 1e-5 (the FD truncation error at these
steps is below 1e-8).  dJ/d(body-point coordinates) along random directions
against differences of fresh handles whose bodies (and Schur matrix) are
built at the moved points."""
import numpy as np
import pytest

from ib_parallel_amd import abi, decks

pytestmark = pytest.mark.gpu

CX = np.array([1.0])
CY = np.array([0.5])


def case(re=100.0, istart=0, m=64, n=64, mgridlev=3):
    p, b = decks.cylinder_case(m=m, n=n, mgridlev=mgridlev, dt=0.01, npts=50, tol=1e-12, num_stat=1, re=re)
    p.istart = istart
    return p, b


def sim_from(p, b, omega0, rhs_old):
    from ib_parallel_amd import GpuSim
    from ib_parallel_amd.adjoint import start_state
    s = GpuSim(p, b)
    start_state(s, omega0, rhs_old)
    return s


def J_forward(p, b, omega0, rhs_old, nsteps, cx=CX, cy=CY):
    s = sim_from(p, b, omega0, rhs_old)
    J = 0.0
    for _ in range(nsteps):
        s.advance()
        fx, fy = s.forces(0)
        J += float(np.dot(cx, fx) + np.dot(cy, fy))
    return J / nsteps, s


def state_of(s):
    lv = [s.level(k + 1) for k in range(s.L)]
    return np.concatenate([d["omega"] for d in lv]), np.concatenate([d["rhs_old"] for d in lv])


@pytest.fixture(scope="module")
def warm():
    """The flow after 15 steps from rest at Re = 100 (the AB2 branch, a shed wake forming)."""
    from ib_parallel_amd import GpuSim
    p, b = case()
    s = GpuSim(p, b)
    for _ in range(15):
        s.advance()
    w, ro = state_of(s)
    s.close()
    return 15, w, ro


def test_objective_and_end_state_match_plain_steps(warm):
    it0, w0, ro0 = warm
    p, b = case(istart=it0)
    Jf, sf = J_forward(p, b, w0, ro0, 7)
    sa = sim_from(p, b, w0, ro0)
    Ja, dre, dw = sa.adjoint(7, CX, CY, ckpt_every=3)
    assert abs(Ja - Jf) <= 1e-13 * abs(Jf)
    assert sa.itime.value == sf.itime.value == it0 + 7
    for k in range(1, p.mgridlev + 1):
        a, f = sa.level(k), sf.level(k)
        for key in ("omega", "s", "qx", "qy", "rhs_old"):
            assert np.array_equal(a[key], f[key]), (k, key)
    assert all(np.array_equal(x, y) for x, y in zip(sa.forces(0), sf.forces(0)))
    # the handle steps on from there exactly like the plain one
    sa.advance(); sf.advance()
    assert np.array_equal(sa.forces(0)[0], sf.forces(0)[0])


@pytest.mark.parametrize("nsteps,ckpt", ((1, 1), (6, 4), (9, 3), (7, None)))
def test_dre_matches_finite_differences(warm, nsteps, ckpt):
    it0, w0, ro0 = warm
    re, h = 100.0, 0.01
    p, b = case(re, istart=it0)
    _, dre, _ = sim_from(p, b, w0, ro0).adjoint(nsteps, CX, CY, ckpt_every=ckpt, omega_grad=False)
    Jp, _ = J_forward(*case(re + h, istart=it0), w0, ro0, nsteps)
    Jm, _ = J_forward(*case(re - h, istart=it0), w0, ro0, nsteps)
    fd = (Jp - Jm) / (2 * h)
    print(f"nsteps {nsteps}: dJ/dRe adjoint {dre:.12e} fd {fd:.12e}")
    assert abs(dre - fd) <= 1e-5 * abs(fd)


@pytest.mark.parametrize("nsteps,ckpt,seed", ((1, 1, 0), (5, 2, 1)))
def test_domega0_matches_finite_differences(warm, nsteps, ckpt, seed):
    it0, w0, ro0 = warm
    p, b = case(istart=it0)
    _, _, dw = sim_from(p, b, w0, ro0).adjoint(nsteps, CX, CY, ckpt_every=ckpt)
    rng = np.random.default_rng(seed)
    for trial in range(2):
        d = rng.standard_normal(w0.size)
        eps = 1e-4 * np.abs(w0).max() / np.abs(d).max()
        Jp, _ = J_forward(p, b, w0 + eps * d, ro0, nsteps)
        Jm, _ = J_forward(p, b, w0 - eps * d, ro0, nsteps)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.dot(dw, d))
        print(f"nsteps {nsteps} trial {trial}: <dJ/dw0, d> adjoint {ad:.12e} fd {fd:.12e}")
        assert abs(ad - fd) <= 1e-5 * abs(fd)


def test_gradients_from_rest_include_the_euler_step():
    """From istart = 0 the first step is Euler; omega0 is a small random
    vorticity so that the nonlinear term is exercised from the first step."""
    p, b = case()
    rng = np.random.default_rng(7)
    L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
    w0 = 0.1 * rng.standard_normal(L * P)
    ro0 = np.zeros(L * P)
    _, dre, dw = sim_from(p, b, w0, ro0).adjoint(4, CX, CY, ckpt_every=2)
    h = 0.01
    Jp, _ = J_forward(*case(100.0 + h), w0, ro0, 4)
    Jm, _ = J_forward(*case(100.0 - h), w0, ro0, 4)
    assert abs(dre - (Jp - Jm) / (2 * h)) <= 1e-5 * abs((Jp - Jm) / (2 * h))
    d = rng.standard_normal(L * P)
    eps = 1e-5
    Jp, _ = J_forward(p, b, w0 + eps * d, ro0, 4)
    Jm, _ = J_forward(p, b, w0 - eps * d, ro0, 4)
    fd = (Jp - Jm) / (2 * eps)
    assert abs(np.dot(dw, d) - fd) <= 1e-5 * abs(fd)


def test_torch_function_gives_the_adjoint_gradients(warm):
    import torch
    from ib_parallel_amd.adjoint import ForceObjective
    it0, w0, ro0 = warm
    p, b = case(istart=it0)
    re = torch.tensor(100.0, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(w0, dtype=torch.float64, requires_grad=True)
    J = ForceObjective.apply(re, w, p, b, 3, CX, CY, 2, ro0)
    gre, gw = torch.autograd.grad(3.0 * J, (re, w))
    Ja, dre, dw = sim_from(p, b, w0, ro0).adjoint(3, CX, CY, ckpt_every=2)
    assert float(J.detach()) == Ja
    assert float(gre) == pytest.approx(3.0 * dre, rel=1e-14)
    assert np.allclose(gw.numpy(), 3.0 * dw, rtol=1e-14, atol=0)


def test_unsupported_configurations_are_rejected():
    """The matrix-free GMRES operator has no matrix to transpose; positions
    on the assembled form need G; num_stat treats torsional bodies as rigid
    and is rejected with them."""
    from ib_parallel_amd import GpuSim, IBError
    p, b = decks.cylinder_case(m=64, n=64, mgridlev=2, dt=0.01, npts=50, tol=1e-10, num_stat=0, schur_mode=1)
    s = GpuSim(p, b)
    with pytest.raises(IBError) as e:
        s.adjoint(2, CX, CY)
    assert e.value.code == abi.IB_E_UNSUPPORTED
    p, b = decks.cylinder_case(m=64, n=64, mgridlev=2, dt=0.01, npts=50, tol=1e-10, num_stat=0, schur_mode=2)
    s = GpuSim(p, b)
    with pytest.raises(IBError) as e:
        s.adjoint(2, CX, CY, position_grad=True)
    assert e.value.code == abi.IB_E_UNSUPPORTED
    pt, bt = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, num_stat=1)
    st = GpuSim(pt, bt)
    with pytest.raises(IBError) as e:
        st.adjoint(2, [0.0] * st.nbodies, [1.0] * st.nbodies)
    assert e.value.code == abi.IB_E_UNSUPPORTED


def test_segmenting_does_not_change_the_gradients(warm):
    """Recomputation from checkpoints reproduces the taped forward exactly, so
    every segment length gives the same bits."""
    it0, w0, ro0 = warm
    p, b = case(istart=it0)
    ref = sim_from(p, b, w0, ro0).adjoint(7, CX, CY, ckpt_every=None)
    for K in (1, 2, 3, 7):
        got = sim_from(p, b, w0, ro0).adjoint(7, CX, CY, ckpt_every=K)
        assert got[0] == ref[0] and got[1] == ref[1], K
        assert np.array_equal(got[2], ref[2]), K


def two_body_case(xs=None, ys=None, f32=0):
    """A cylinder (40 points, off the grid lines) and a small second cylinder
    (16 points) at 64x64x3; xs, ys replace the point coordinates (deck order)."""
    t1 = 2 * np.pi * (np.arange(40) + 0.37) / 40
    t2 = 2 * np.pi * (np.arange(16) + 0.11) / 16
    x = np.concatenate([0.5 * np.cos(t1) + 0.0123, 0.2 * np.cos(t2) + 1.5071])
    y = np.concatenate([0.5 * np.sin(t1) - 0.0071, 0.2 * np.sin(t2) + 0.6033])
    if xs is not None:
        x, y = xs, ys
    p = abi.make_params(64, 64, 3, 0.01, 100.0, 4.0, 1.0, 2.0, fsi_tol=1e-12, atol=1e-12, rtol=1e-12,
                        num_stat=1, deltafnc_f32_literals=f32)
    return p, [abi.BodySpec(x[:40], y[:40]), abi.BodySpec(x[40:], y[40:])], x, y


@pytest.mark.parametrize("f32", (0, 1))
def test_dpositions_match_finite_differences(f32):
    from ib_parallel_amd import GpuSim
    p, b, x, y = two_body_case(f32=f32)
    s = GpuSim(p, b)
    for _ in range(10):
        s.advance()
    w0, ro0 = state_of(s)
    s.close()
    p.istart = 10
    cx, cy = np.array([1.0, 0.3]), np.array([0.5, -0.2])
    J, _, _, gx, gy = sim_from(p, b, w0, ro0).adjoint(4, cx, cy, ckpt_every=2, position_grad=True)
    assert gx.shape == (56,) and gy.shape == (56,)
    rng = np.random.default_rng(11)
    h = 1e-6
    for trial in range(2):
        dx, dy = rng.standard_normal(56), rng.standard_normal(56)
        Js = []
        for sg in (1.0, -1.0):
            pp, bb, _, _ = two_body_case(x + sg * h * dx, y + sg * h * dy, f32=f32)
            pp.istart = 10
            Js.append(J_forward(pp, bb, w0, ro0, 4, cx, cy)[0])
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float(np.dot(gx, dx) + np.dot(gy, dy))
        print(f"f32 {f32} trial {trial}: <dJ/dx, d> adjoint {ad:.12e} fd {fd:.12e}")
        assert abs(ad - fd) <= 1e-5 * abs(fd)


# ---- moving bodies: the FSI fixed point in the face-space form ------------

def tors_case(re=50.0, k=0.2, g0=None, dx=None, dy=None):
    """decks.tiny_case at 40x32x2: a rigid cylinder (moved off the grid lines,
    where deltafnc's REAL(4) literals jump) and a hinged torsional plate
    (face-space Schur form), tolerances 1e-12."""
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, re=re)
    nc = b[0].npts
    t = 2 * np.pi * (np.arange(nc) + 0.29) / nc
    x, y = 0.3 * np.cos(t) - 0.4 + 0.0123, 0.3 * np.sin(t) - 0.0071
    if dx is not None:
        x, y = x + dx, y + dy
    b[0] = abi.BodySpec(x, y, fam=1)
    b[1].ktheta = k
    if g0 is not None:
        b[1].gamma0 = g0
    return p, b


def tors_J(p, b, w0, nsteps, cx, cy):
    return J_forward(p, b, w0, np.zeros_like(w0), nsteps, cx, cy)[0]


def test_fsi_objective_and_end_state_match_plain_steps():
    p, b = tors_case()
    L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
    w0 = np.zeros(L * P)
    cx, cy = np.array([0.4, 1.0]), np.array([0.2, -0.7])
    Jf, sf = J_forward(p, b, w0, w0, 6, cx, cy)
    sa = sim_from(p, b, w0, w0)
    assert sa.schur_info()[0] == 3
    Ja = sa.adjoint(6, cx, cy, ckpt_every=4)[0]
    assert abs(Ja - Jf) <= 1e-13 * abs(Jf)
    for k in range(1, p.mgridlev + 1):
        a, f = sa.level(k), sf.level(k)
        for key in ("omega", "s", "qx", "qy", "rhs_old"):
            assert np.array_equal(a[key], f[key]), (k, key)
    ba, bf = sa.body(), sf.body()
    for key in ("f", "xb", "yb", "theta", "thetad", "thetadd"):
        assert np.array_equal(ba[key], bf[key]), key
    sa.advance(); sf.advance()
    assert np.array_equal(sa.body()["f"], sf.body()["f"])


@pytest.mark.parametrize("nsteps,ckpt", ((1, 1), (5, 2)))
def test_fsi_stiffness_and_pitch_match_finite_differences(nsteps, ckpt):
    p, b = tors_case()
    L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
    w0 = np.zeros(L * P)
    cx, cy = np.array([0.4, 1.0]), np.array([0.2, -0.7])
    out = sim_from(p, b, w0, w0).adjoint(nsteps, cx, cy, ckpt_every=ckpt, body_grad=True)
    dK, dg = out[3][1], out[4][1]
    assert out[3][0] == 0.0 and out[4][0] == 0.0
    # dJ/dK is small (~2e-3) against the FSI loop's 1e-12 convergence noise:
    # central differences in K scatter by up to 2.4e-5 relative at steps
    # 2.5e-4..2e-3 (measured, either DST form) and by 1e-6 at 4e-3, so the
    # difference takes steps 4e-3 and 8e-3 with Richardson extrapolation
    # (truncation O(h^4))
    def fd_k(hk):
        return (tors_J(*tors_case(k=0.2 + hk), w0, nsteps, cx, cy) -
                tors_J(*tors_case(k=0.2 - hk), w0, nsteps, cx, cy)) / (2 * hk)
    fdk = (4 * fd_k(4e-3) - fd_k(8e-3)) / 3
    g0 = b[1].gamma0
    hg = 1e-6
    fdg = (tors_J(*tors_case(g0=g0 + hg), w0, nsteps, cx, cy) -
           tors_J(*tors_case(g0=g0 - hg), w0, nsteps, cx, cy)) / (2 * hg)
    print(f"nsteps {nsteps}: dJ/dK adjoint {dK:.12e} fd {fdk:.12e}; dJ/dgamma0 adjoint {dg:.12e} fd {fdg:.12e}")
    assert abs(dK - fdk) <= 1e-5 * abs(fdk)
    assert abs(dg - fdg) <= 1e-5 * abs(fdg)


def test_fsi_gradients_match_finite_differences_of_the_oracle():
    """The same moving-body case from rest, 5 steps in 3 segments: dJ/dK,
    dJ/dgamma0 and dJ/dRe of the GPU adjoint against central differences of
    the oracle's forward steps (the CPU restatement of the reference's
    step), so the gradient is checked against the reference's algorithm
    without the GPU forward in between."""
    from oracle import binding as ob
    p, b = tors_case()
    L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
    w0 = np.zeros(L * P)
    cx, cy = np.array([0.4, 1.0]), np.array([0.2, -0.7])
    n = 5
    J, dre, _, dK, dg = sim_from(p, b, w0, w0).adjoint(n, cx, cy, ckpt_every=2, omega_grad=False, body_grad=True)

    def oracle_J(pp, bb):
        o = ob.OracleSim(pp, bb)
        acc = 0.0
        for _ in range(n):
            assert o.advance().status == abi.IB_OK
            fx, fy = o.forces(0)
            acc += float(np.dot(cx, fx) + np.dot(cy, fy))
        o.close()
        return acc / n

    assert abs(J - oracle_J(p, b)) <= 1e-9 * abs(J)

    def fd_k(hk):
        return (oracle_J(*tors_case(k=0.2 + hk)) - oracle_J(*tors_case(k=0.2 - hk))) / (2 * hk)
    fdk = (4 * fd_k(1e-3) - fd_k(2e-3)) / 3
    g0 = b[1].gamma0
    fdg = (oracle_J(*tors_case(g0=g0 + 1e-6)) - oracle_J(*tors_case(g0=g0 - 1e-6))) / 2e-6
    fdre = (oracle_J(*tors_case(re=50.0 + 1e-2)) - oracle_J(*tors_case(re=50.0 - 1e-2))) / 2e-2
    print(f"oracle differences: dJ/dK adjoint {dK[1]:.12e} fd {fdk:.12e}; dJ/dgamma0 adjoint {dg[1]:.12e} "
          f"fd {fdg:.12e}; dJ/dRe adjoint {dre:.12e} fd {fdre:.12e}")
    assert abs(dK[1] - fdk) <= 1e-5 * abs(fdk)
    assert abs(dg[1] - fdg) <= 1e-5 * abs(fdg)
    assert abs(dre - fdre) <= 1e-5 * abs(fdre)


def test_fsi_re_omega_and_rigid_positions_match_finite_differences():
    p, b = tors_case()
    L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
    rng = np.random.default_rng(5)
    w0 = 0.05 * rng.standard_normal(L * P)
    cx, cy = np.array([0.4, 1.0]), np.array([0.2, -0.7])
    n = 4
    J, dre, dw, gx, gy = sim_from(p, b, w0, np.zeros_like(w0)).adjoint(n, cx, cy, ckpt_every=3, position_grad=True)
    h = 1e-3
    fd = (tors_J(*tors_case(re=50.0 + h), w0, n, cx, cy) - tors_J(*tors_case(re=50.0 - h), w0, n, cx, cy)) / (2 * h)
    print(f"dJ/dRe adjoint {dre:.12e} fd {fd:.12e}")
    assert abs(dre - fd) <= 1e-5 * abs(fd)
    d = rng.standard_normal(L * P)
    eps = 1e-5
    fd = (tors_J(p, b, w0 + eps * d, n, cx, cy) - tors_J(p, b, w0 - eps * d, n, cx, cy)) / (2 * eps)
    print(f"<dJ/dw0, d> adjoint {np.dot(dw, d):.12e} fd {fd:.12e}")
    assert abs(np.dot(dw, d) - fd) <= 1e-5 * abs(fd)
    nc = b[0].npts
    assert np.all(gx[nc:] == 0.0) and np.all(gy[nc:] == 0.0)
    ex, ey = rng.standard_normal(nc), rng.standard_normal(nc)
    h = 1e-6
    fd = (tors_J(*tors_case(dx=h * ex, dy=h * ey), w0, n, cx, cy) -
          tors_J(*tors_case(dx=-h * ex, dy=-h * ey), w0, n, cx, cy)) / (2 * h)
    ad = float(np.dot(gx[:nc], ex) + np.dot(gy[:nc], ey))
    print(f"<dJ/dx, d> adjoint {ad:.12e} fd {fd:.12e}")
    assert abs(ad - fd) <= 1e-5 * abs(fd)


def test_fsi_hinge_and_spacing_gradients_match_finite_differences():
    """The torsional plate's hinge and point-spacing gradients
    (ib_adjoint_hinge_grad) against central differences of J: the plate
    translated by +-h (its deck points, and so its hinge, the first of
    them), and its deck points scaled about the hinge so that ds changes by
    +-h.  The hinge gradient is the sum of the points' position partials;
    ds also scales the arms (the coupling vector and the structural block).
    The plate is first moved off the grid lines (its hinge sits on them in
    tors_case, where deltafnc's REAL(4) literals jump)."""
    from ib_parallel_amd import IBError

    def shifted(dx, dy, dds=0.0):
        p2, b2 = tors_case()
        x, y = b2[1].x + 0.0123, b2[1].y - 0.0071
        ds = np.hypot(x[1] - x[0], y[1] - y[0])
        x, y = x[0] + (x - x[0]) * (1 + dds / ds), y[0] + (y - y[0]) * (1 + dds / ds)
        b2[1].x, b2[1].y = x + dx, y + dy
        return p2, b2
    p, b = shifted(0.0, 0.0)
    L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
    rng = np.random.default_rng(7)
    w0 = 0.05 * rng.standard_normal(L * P)
    cx, cy = np.array([0.4, 1.0]), np.array([0.2, -0.7])
    n = 4
    s = sim_from(p, b, w0, np.zeros_like(w0))
    s.adjoint(n, cx, cy, ckpt_every=3, omega_grad=False, position_grad=True)
    gxh, gyh, gds = s.hinge_grad()
    assert gxh[0] == 0.0 and gyh[0] == 0.0 and gds[0] == 0.0  # the rigid cylinder has no hinge
    h = 1e-6
    for ex, ey in ((1.0, 0.0), (0.0, 1.0)):
        fd = (tors_J(*shifted(h * ex, h * ey), w0, n, cx, cy) - tors_J(*shifted(-h * ex, -h * ey), w0, n, cx, cy)) \
            / (2 * h)
        ad = gxh[1] * ex + gyh[1] * ey
        print(f"dJ/dhinge ({ex}, {ey}): adjoint {ad:.12e} fd {fd:.12e}")
        assert abs(ad - fd) <= 1e-5 * abs(fd)
    fd = (tors_J(*shifted(0.0, 0.0, h), w0, n, cx, cy) - tors_J(*shifted(0.0, 0.0, -h), w0, n, cx, cy)) / (2 * h)
    print(f"dJ/dds: adjoint {gds[1]:.12e} fd {fd:.12e}")
    assert abs(gds[1] - fd) <= 1e-5 * abs(fd)
    # a run without position gradients leaves none to report
    s.adjoint(1, cx, cy, omega_grad=False)
    with pytest.raises(IBError) as e:
        s.hinge_grad()
    assert e.value.code == abi.IB_E_INVAL
    s.close()


def test_batch_adjoint_equals_each_simulations_adjoint():
    """ib_batch_adjoint runs every simulation's adjoint from its own host
    thread: the results are those of each handle's own ib_adjoint_run."""
    from ib_parallel_amd import GpuBatch
    cases = [tors_case(k=k) for k in (0.15, 0.2, 0.3)]
    cx, cy = np.array([0.4, 1.0] * 3), np.array([0.2, -0.7] * 3)
    batch = GpuBatch(cases)
    J, dre, dk, dg = batch.adjoint(4, cx, cy, ckpt_every=3)
    for i, (p, b) in enumerate(cases):
        L, P = p.mgridlev, (p.m - 1) * (p.n - 1)
        z = np.zeros(L * P)
        out = sim_from(p, b, z, z).adjoint(4, cx[2 * i:2 * i + 2], cy[2 * i:2 * i + 2], ckpt_every=3,
                                            omega_grad=False, body_grad=True)
        assert J[i] == out[0] and dre[i] == out[1], i
        assert np.array_equal(dk[2 * i:2 * i + 2], out[3]) and np.array_equal(dg[2 * i:2 * i + 2], out[4]), i
    assert all(s.itime.value == 4 for s in batch.sims)
    batch.close()


def test_assembled_gmres_form_matches_finite_differences():
    """Rigid bodies on the GMRES path with the assembled Schur matrix (the
    default form without num_stat): the adjoint solves S^T by GMRES."""
    def mk(re=100.0, istart=0):
        p, b = decks.cylinder_case(m=64, n=64, mgridlev=3, dt=0.01, npts=50, tol=1e-13, num_stat=0, re=re,
                                   schur_mode=2, gmres_restart=120)
        p.istart = istart
        return p, b
    from ib_parallel_amd import GpuSim
    p, b = mk()
    s = GpuSim(p, b)
    assert s.schur_info()[0] == 2
    for _ in range(10):
        s.advance()
    w0, ro0 = state_of(s)
    s.close()
    n = 4
    J, dre, dw = sim_from(*mk(istart=10), w0, ro0).adjoint(n, CX, CY, ckpt_every=3)
    Jf, _ = J_forward(*mk(istart=10), w0, ro0, n)
    assert abs(J - Jf) <= 1e-13 * abs(Jf)
    h = 0.01
    fd = (J_forward(*mk(100.0 + h, 10), w0, ro0, n)[0] - J_forward(*mk(100.0 - h, 10), w0, ro0, n)[0]) / (2 * h)
    print(f"assembled: dJ/dRe adjoint {dre:.12e} fd {fd:.12e}")
    assert abs(dre - fd) <= 1e-5 * abs(fd)
    d = np.random.default_rng(4).standard_normal(w0.size)
    eps = 1e-4 * np.abs(w0).max() / np.abs(d).max()
    fd = (J_forward(*mk(istart=10), w0 + eps * d, ro0, n)[0] - J_forward(*mk(istart=10), w0 - eps * d, ro0, n)[0]) / (2 * eps)
    print(f"assembled: <dJ/dw0, d> adjoint {np.dot(dw, d):.12e} fd {fd:.12e}")
    assert abs(np.dot(dw, d) - fd) <= 1e-5 * abs(fd)


def test_c3_plate_adjoint_matches_finite_differences():
    """The C3 geometry and grid (hinged plate of 101 points, 400x400x3, K =
    0.1, pitch 10 deg, Re 200, dt 1e-3; tolerances 1e-12 for clean
    differences), 200 steps from rest with a checkpoint every 20 (10
    segments, 180 steps recomputed): dJ/dK and dJ/dgamma0 of the mean lift
    against central differences.  Measured over h: the K differences
    approach the adjoint as h grows (2.1e-5 relative at h = 1e-6, where J
    moves by 7e-8 against its ~1e-12 convergence noise; 1.2e-6 at 1e-5,
    2.4e-7 at 1e-3), so h = 1e-3 (1% of K).  The pitch differences are
    closest at h = 1e-6 (9e-6; 2.9e-5 at 1e-5): a larger turn moves plate
    points across the radii where deltafnc's REAL(4) literals make the
    weights jump, and the adjoint differentiates each branch."""
    from ib_parallel_amd import GpuSim

    def make(K=0.1, g0=None):
        p, b = decks.plate_case(mgridlev=3, tol=1e-12)
        b[0].ktheta = K
        if g0 is not None:
            b[0].gamma0 = g0
        return p, b

    def mean_lift(p, b, n):
        s = GpuSim(p, b)
        acc = 0.0
        for _ in range(n):
            assert s.advance().status == abi.IB_OK
            acc += s.forces()[1][0]
        s.close()
        return acc / n

    n = 200
    p, b = make()
    g0 = b[0].gamma0
    s = GpuSim(p, b)
    J, _, _, dK, dg = s.adjoint(n, [0.0], [1.0], ckpt_every=20, omega_grad=False, body_grad=True)
    s.close()
    assert abs(J - mean_lift(p, b, n)) <= 1e-12 * abs(J)
    hK = 1e-3
    fdK = (mean_lift(*make(0.1 + hK), n) - mean_lift(*make(0.1 - hK), n)) / (2 * hK)
    hg = 1e-6
    fdg = (mean_lift(*make(g0=g0 + hg), n) - mean_lift(*make(g0=g0 - hg), n)) / (2 * hg)
    print(f"C3 plate: dJ/dK adjoint {dK[0]:.10e} fd {fdK:.10e}; dJ/dgamma0 adjoint {dg[0]:.10e} fd {fdg:.10e}")
    assert abs(dK[0] - fdK) <= 1e-5 * abs(fdK)
    assert abs(dg[0] - fdg) <= 2e-5 * abs(fdg)
