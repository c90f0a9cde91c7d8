"""Oracle pressure post-processing (calculate_pressure,
src/operators_pressure.F90:110-437).

ddti inverts the 5-point cell Laplacian with zero Dirichlet data (DD^T of
the cell-centred grid); the pressure coarsify sums four fine cells onto the
coarse cell at the same place (a linear field lands on the coarse cell
centre, a constant c becomes 4c); a uniform flow with no body has pressure
coefficient 0 on every level (boundary 1/2|U|^2 propagates to a constant,
then p - 1/2|u|^2 = 0)."""
import numpy as np
import pytest

from conftest import rel
from ib_parallel_amd import abi, decks
from oracle import binding as ob


def lap_cells(p):
    q = np.pad(p, 1)
    return q[1:-1, 2:] + q[1:-1, :-2] + q[2:, 1:-1] + q[:-2, 1:-1] - 4.0 * p


@pytest.mark.parametrize("m,n", ((40, 32), (64, 48)))
def test_ddti_inverts_cell_laplacian(m, n):
    sim = ob.OracleSim(abi.make_params(m, n, 2, 0.01, 100.0, 4.0, 1.6, 1.5), [])
    r = np.random.default_rng(m).standard_normal((n, m))
    p = sim.ddti(r)
    assert rel(lap_cells(p).ravel(), r.ravel()) < 1e-12


def test_coarsify_pressure_places_cells():
    m, n, ln, ox, oy = 40, 32, 4.0, 1.6, 1.5
    sim = ob.OracleSim(abi.make_params(m, n, 2, 0.01, 100.0, ln, ox, oy), [])
    d = ln / m
    i = np.arange(1, m + 1)
    j = np.arange(1, n + 1)
    xc = (i - 0.5) * d - ox
    yc = (j - 0.5) * d - oy
    # coarse level: same centre, twice the extent
    xc2 = (i - 0.5) * 2 * d - ox - ln / 2
    yc2 = (j - 0.5) * 2 * d - oy - n * d / 2
    ids = np.arange(m // 4 + 2, 3 * m // 4)  # coarse cells written: m/4+2 .. 3m/4-1
    jds = np.arange(n // 4 + 2, 3 * n // 4)
    const = sim.coarsify_pressure(np.full((n, m), 3.0))
    assert np.all(const[np.ix_(jds - 1, ids - 1)] == 12.0)
    written = np.zeros((n, m), bool)
    written[np.ix_(jds - 1, ids - 1)] = True
    assert np.all(const[~written] == 0.0)
    lin = sim.coarsify_pressure(np.add.outer(2.0 * yc, xc))
    want = 4.0 * np.add.outer(2.0 * yc2[jds - 1], xc2[ids - 1])
    assert np.abs(lin[np.ix_(jds - 1, ids - 1)] - want).max() < 1e-12


@pytest.mark.parametrize("L", (1, 3))
def test_pressure_of_uniform_flow_is_zero(L):
    sim = ob.OracleSim(abi.make_params(48, 40, L, 0.01, 100.0, 4.0, 1.6, 1.5), [])
    assert np.abs(sim.pressure()).max() < 1e-12
    for _ in range(2):
        sim.advance()
    assert np.abs(sim.pressure()).max() < 1e-12


def test_pressure_cylinder_front_is_positive():
    """After a few steps of an impulsively started cylinder the stagnation
    region ahead of the body is a high-pressure region and the flanks are
    low (a sanity check of signs and scaling, not parity)."""
    p, b = decks.cylinder_case(m=64, n=64, mgridlev=2, dt=0.01, npts=50, tol=1e-10, gmres_restart=120)
    sim = ob.OracleSim(p, b)
    for _ in range(3):
        sim.advance()
    cp = sim.pressure()[0]
    d = 4.0 / 64
    xc = (np.arange(64) + 0.5) * d - 1.0
    yc = (np.arange(64) + 0.5) * d - 2.0
    front = cp[np.argmin(np.abs(yc)), np.argmin(np.abs(xc + 0.6))]
    flank = cp[np.argmin(np.abs(yc - 0.6)), np.argmin(np.abs(xc))]
    assert front > 0.5 and flank < 0.0
