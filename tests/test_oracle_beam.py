"""Family-3 bodies (deformable plates as linear Euler-Bernoulli beams) in the
oracle.  The reference declares the family and leaves its routines stubs
(src/operators_struct.F90:104-106, 146-148, 376-379, 472-474), so there is
no reference output to match: the beam is pinned by closed forms that a
plausible mistake fails.

* Natural frequencies of the clamped-free beam: omega_k = (beta_k L)^2
  sqrt(EI / m) / L^2 with beta_1 L = 1.8751, beta_2 L = 4.6941 (a wrong
  entry of the element mass or stiffness matrix, or a free first node,
  moves them).
* Static deflection under a uniform load q: w(s) = q s^2 (6 L^2 - 4 L s +
  s^2) / (24 EI), tip q L^4 / (8 EI).  Point loads with Simpson weights
  over each element (points at the nodes and mid-elements) give the
  consistent load vector of the cubic shape functions exactly, and Hermite
  elements are then exact at the nodes: rounding only.
* The Newmark update (average acceleration): the static solution of K u =
  delta/dt Phi^T f stays put under its own load (pins the force scaling),
  and free vibration conserves 1/2 ud^T M ud + 1/2 u^T K u.
* Coupled to the flow: after each step E(q + q0) equals delta times the
  beam points' velocity n (H ud) (no slip on the moving plate).
"""
import numpy as np
import pytest

from ib_parallel_amd import abi, decks
from oracle import binding as ob


def beam_sim(npts=101, mass=0.2, k_b=0.1, **kw):
    p = abi.make_params(64, 64, 2, 1e-3, 200.0, 4.0, 1.0, 2.0, **kw)
    return p, ob.OracleSim(p, [decks.beam_body(npts=npts, mass_ratio=mass, k_b=k_b)])


@pytest.mark.parametrize("npts", (21, 101))
def test_beam_natural_frequencies(npts):
    import scipy.linalg as sl
    p, o = beam_sim(npts)
    M, K, H = o.beam_matrices()
    w2 = sl.eigh(K, M, eigvals_only=True)
    om = np.sqrt(np.sort(w2))[:2]
    L, EI, m = 1.0, 0.1, 0.2
    exact = np.array([1.875104068711961, 4.694091132974175]) ** 2 * np.sqrt(EI / m) / L ** 2
    ne = (npts - 1) // 2
    # Hermite cubics converge as h^4 in the frequencies
    tol = 2.0 * (1.0 / ne) ** 4
    assert np.all(np.abs(om / exact - 1) < tol), (om, exact)
    o.close()


def test_beam_static_deflection_under_uniform_load():
    p, o = beam_sim(101)
    M, K, H = o.beam_matrices()
    npts, ne = 101, 50
    L, EI, q = 1.0, 0.1, 0.7
    h = L / ne
    w = np.zeros(npts)  # Simpson weights: h/6 (1, 4, 1) per element
    for e in range(ne):
        w[2 * e] += h / 6
        w[2 * e + 1] += 4 * h / 6
        w[2 * e + 2] += h / 6
    F = H.T @ (q * w)
    u = np.linalg.solve(K, F)
    s = h * np.arange(1, ne + 1)
    exact = q * s ** 2 * (6 * L ** 2 - 4 * L * s + s ** 2) / (24 * EI)
    slope = q * s * (3 * L ** 2 - 3 * L * s + s ** 2) / (6 * EI)
    # exact up to rounding in the solve (cond(K) ~ (L/h)^4 ~ 6e6); a load
    # vector that is not the consistent one would be off by ~h^2 ~ 4e-4
    assert np.abs(u[0::2] - exact).max() < 2e-9 * exact.max()
    assert np.abs(u[1::2] - slope).max() < 2e-9 * slope.max()
    assert abs(u[-2] - q * L ** 4 / (8 * EI)) < 2e-9 * q * L ** 4 / (8 * EI)
    # at the nodes the shape functions pick the node's deflection (the
    # clamped first point none); mid-element they are 1/2, h/8, 1/2, -h/8
    nodes = np.zeros((ne + 1, 2 * ne))
    for k in range(1, ne + 1):
        nodes[k, 2 * (k - 1)] = 1.0
    assert np.array_equal(H[0::2], nodes)
    mid = H[3]  # between nodes 1 and 2
    assert np.allclose([mid[0], mid[1], mid[2], mid[3]], [0.5, h / 8, 0.5, -h / 8], rtol=1e-14, atol=0)
    o.close()


def test_beam_newmark_static_fixed_point_and_energy():
    p, o = beam_sim(101)
    M, K, H = o.beam_matrices()
    npts = 101
    delta, dt = p.len / p.m, p.dt
    g = -np.radians(10.0)
    nx, ny = -np.sin(g), np.cos(g)  # the plate's normal (axis at angle -10 deg)
    rng = np.random.default_rng(3)
    fn = rng.standard_normal(npts) * 1e-3  # normal point forces
    fb = np.concatenate([nx * fn, ny * fn])
    us = np.linalg.solve(K, delta / dt * (H.T @ fn))
    z = np.zeros_like(us)
    u, ud, udd = o.beam_newmark(0, 5, fb, us, z, z)
    assert np.abs(u - us).max() < 1e-12 * np.abs(us).max()
    assert np.abs(ud).max() < 1e-9 * np.abs(us).max()
    # free vibration from the static shape: energy is conserved
    u0, ud0 = us, z
    udd0 = -np.linalg.solve(M, K @ u0)
    E0 = 0.5 * ud0 @ M @ ud0 + 0.5 * u0 @ K @ u0
    u, ud, udd = o.beam_newmark(0, 400, np.zeros(2 * npts), u0, ud0, udd0)
    E1 = 0.5 * ud @ M @ ud + 0.5 * u @ K @ u
    # rounding drift 9e-11 over 400 steps; a dissipative or mis-weighted
    # update loses or gains O(omega dt) per step, per cents here
    assert abs(E1 - E0) < 1e-9 * E0
    assert np.abs(u - u0).max() > 1e-3 * np.abs(u0).max()  # it did move
    o.close()


def test_beam_step_enforces_no_slip():
    p, b = decks.tiny_case(m=40, n=32, mgridlev=2, tol=1e-12, gmres_restart=80, beam=True)
    o = ob.OracleSim(p, b)
    d = p.len / p.m
    n0, n1 = b[0].npts, b[1].npts
    g = b[1].gamma0
    nx, ny = -np.sin(g), np.cos(g)
    M, K, H = o.beam_matrices()
    for _ in range(3):
        assert o.advance().status == abi.IB_OK
        lv = o.level(1)
        e = o.regT(lv["qx"] + d, lv["qy"])
        assert np.abs(e[:2 * n0]).max() < 1e-11 * d  # the rigid cylinder
        vn = H @ o.beam(0)["ud"]  # normal velocity of the beam's points
        assert np.abs(vn).max() > 1e-3
        ex, ey = e[2 * n0:2 * n0 + n1], e[2 * n0 + n1:2 * n0 + 2 * n1]
        assert np.abs(ex - d * nx * vn).max() < 1e-10 * np.abs(d * vn).max()
        assert np.abs(ey - d * ny * vn).max() < 1e-10 * np.abs(d * vn).max()
    o.close()


def test_beam_deck_checks():
    p = abi.make_params(64, 64, 2, 1e-3, 200.0, 4.0, 1.0, 2.0)
    from ib_parallel_amd.sim import IBError
    bad = decks.beam_body(npts=100)  # even: no whole elements
    with pytest.raises(IBError):
        ob.OracleSim(p, [bad])
    p2 = abi.make_params(64, 64, 2, 1e-3, 200.0, 4.0, 1.0, 2.0, num_stat=1)
    with pytest.raises(IBError):
        ob.OracleSim(p2, [decks.beam_body(npts=21)])


def test_beam_schur_term():
    """b_times2 on the beam's entries: 2 delta^2/dt^2 Phi K_eff^-1 Phi^T with
    Phi = n H and K_eff = 4/dt^2 M + 2/dt C + K, built here from the pinned
    M, K and H (the FSI fixed point does not depend on this term, only its
    convergence does, so no-slip does not pin it)."""
    p, o = beam_sim(21)
    M, K, H = o.beam_matrices()
    delta, dt = p.len / p.m, p.dt
    g = -np.radians(10.0)
    n = np.array([-np.sin(g), np.cos(g)])
    Phi = np.vstack([n[0] * H, n[1] * H])
    Keff = 4 / dt ** 2 * M + K
    B = 2 * delta ** 2 / dt ** 2 * Phi @ np.linalg.solve(Keff, Phi.T)
    x = np.random.default_rng(1).standard_normal(o.nf)
    y = ob.btimes(o, x)
    assert np.abs(y - B @ x).max() < 1e-12 * np.abs(B @ x).max()
    o.close()
