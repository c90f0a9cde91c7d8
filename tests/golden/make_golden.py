"""Writes tests/golden/dst_golden.npz: DST-I known answers from scipy.

scipy.fft.dst(type=1) is the same transform as FFTW's RODFT00 (unnormalised,
Y_k = 2 sum_j X_j sin(pi (j+1)(k+1)/(N+1))), which src/myfft.F90:260 plans
for every spectral solve.  Inputs are seeded normal vectors; inputs and
outputs are both stored so the fixture does not depend on the RNG stream.

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for N in (1, 2, 3, 7, 15, 63, 199, 255, 267, 399, 511, 1023):
        x = rng.standard_normal((2, N))
        out[f"d1_in_{N}"] = x
        out[f"d1_out_{N}"] = sf.dst(x, type=1, axis=-1)
    for (nx, ny) in ((7, 5), (31, 23), (63, 47), (99, 67)):
        # node arrays are (m-1) x (n-1) Fortran order = numpy (ny, nx) C order
        x = rng.standard_normal((ny, nx))
        out[f"d2_in_{nx}x{ny}"] = x
        out[f"d2_out_{nx}x{ny}"] = sf.dstn(x, type=1)
    np.savez_compressed(os.path.join(HERE, "dst_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "dst_golden.npz"))


if __name__ == "__main__":
    main()
