"""Oracle DST-I pinned against scipy golden vectors, brute force and S^2 = 2(N+1)."""
import numpy as np
import pytest

from conftest import rel
from oracle import binding as ob

SIZES = (1, 2, 3, 7, 15, 63, 199, 255, 267, 399, 511, 1023)


@pytest.mark.parametrize("N", SIZES)
def test_dst1d_golden(golden, N):
    x = golden[f"d1_in_{N}"]
    assert rel(ob.dst1d(x), golden[f"d1_out_{N}"]) < 2e-15 * max(1.0, np.log2(N + 1))


@pytest.mark.parametrize("N", (1, 5, 12, 33))
def test_dst1d_bruteforce(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N)
    j = np.arange(1, N + 1)
    S = 2 * np.sin(np.pi * np.outer(j, j) / (N + 1))  # FFTW RODFT00 definition
    assert rel(ob.dst1d(x), S @ x) < 1e-14


@pytest.mark.parametrize("N", (15, 267, 1023))
def test_dst1d_involution(N):
    x = np.random.default_rng(1).standard_normal(N)
    assert rel(ob.dst1d(ob.dst1d(x)) / (2 * (N + 1)), x) < 1e-14


@pytest.mark.parametrize("shape", ((7, 5), (31, 23), (63, 47), (99, 67)))
def test_dst2d_golden(golden, shape):
    from ib_parallel_amd import abi
    nx, ny = shape
    m, n = nx + 1, ny + 1
    if m % 4 or n % 4 or m < 8 or n < 8:
        pytest.skip("grid sizes of the solver are multiples of 4")
    p = abi.make_params(m, n, 1, 0.01, 100.0, 1.0, 0.5, 0.5)
    sim = ob.OracleSim(p, [])
    x = golden[f"d2_in_{nx}x{ny}"]
    assert rel(sim.dst2d(x.ravel()), golden[f"d2_out_{nx}x{ny}"].ravel()) < 1e-14
