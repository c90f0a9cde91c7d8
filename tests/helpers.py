"""Shared geometry helpers for the tests (grid coordinates of the reference)."""
import numpy as np


def level_coords(sim, k):
    """x, y of nodes (i=2..m, j=2..n) of level k (1-based) as (n-1, m-1) arrays.

    Level k spans 2^(k-1) times the finest domain about the same centre; node
    i of level k sits at x = c_x + (i-1-m/2) * delta * 2^(k-1)
    (src/variables.F90:339-340 uses the same construction for faces).
    """
    p = sim.params
    d = p.len / p.m
    cx = p.m / 2 * d - p.offsetx
    cy = p.n / 2 * d - p.offsety
    h = d * 2 ** (k - 1)
    i = np.arange(2, p.m + 1)
    j = np.arange(2, p.n + 1)
    X = cx + (i - 1 - p.m / 2) * h
    Y = cy + (j - 1 - p.n / 2) * h
    return np.meshgrid(X, Y)


def curl_dense(m, n, s_ext):
    """C applied to a node field with boundary (s_ext is (n+1, m+1), i fastest)."""
    qx = s_ext[1:, :] - s_ext[:-1, :]          # qx(i,j) = s(i,j+1) - s(i,j), shape (n, m+1)
    qy = s_ext[:, :-1] - s_ext[:, 1:]          # qy(i,j) = s(i,j) - s(i+1,j), shape (n+1, m)
    return qx, qy


def lap5(s, m, n):
    """5-point C^T C s with zero Dirichlet boundary, s shape (n-1, m-1)."""
    S = np.pad(s, 1)
    return 4 * S[1:-1, 1:-1] - S[:-2, 1:-1] - S[2:, 1:-1] - S[1:-1, :-2] - S[1:-1, 2:]
