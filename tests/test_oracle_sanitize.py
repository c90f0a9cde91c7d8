"""The oracle's C++ under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds liboracle_san.so from the same sources with
-fsanitize=address,undefined (UB not recoverable).  A child process preloads
the two runtimes (Python itself is not instrumented), loads that build
through IB_ORACLE_LIB and runs small cases of every path the oracle has:
the matrix-free GMRES step with a torsional body, num_stat, the beam of
family 3, a sub-domain deck, restart state, the pressure and the operator
entry points.  Any report aborts the child, and its exit status fails the
test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

CHILD = r"""
import ctypes
# both runtimes are in the process (an instrumented library would not load otherwise)
ctypes.CDLL(None).__asan_init
ctypes.CDLL(None).__ubsan_handle_add_overflow
import numpy as np
from ib_parallel_amd import abi, decks
from oracle import binding as ob
assert ob.LIB.endswith("liboracle_san.so"), ob.LIB

def steps(o, n):
    for _ in range(n):
        assert o.advance().status == abi.IB_OK

# matrix-free GMRES, rigid + torsional bodies, two levels
p, b = decks.tiny_case(m=32, n=24, mgridlev=2, tol=1e-10, gmres_restart=40)
o = ob.OracleSim(p, b)
steps(o, 2)
rng = np.random.default_rng(0)
o.ctci(rng.standard_normal(o.P)); o.ainv(rng.standard_normal(o.P))
o.vort2flux(rng.standard_normal(o.L * o.P))
o.schur_matvec(rng.standard_normal(o.nf)); o.schur_matrix()
o.reg(rng.standard_normal(o.nf)); o.forces(); o.pressure()
lv = o.level(1); o.set_level(1, **lv); o.set_body(f=o.body()["f"])
o.close()
# num_stat (precomputed matrix + Cholesky)
p, b = decks.cylinder_case(m=48, n=48, mgridlev=1, dt=0.01, npts=36, num_stat=1)
o = ob.OracleSim(p, b); steps(o, 2); o.close()
# the beam of family 3
p, b = decks.tiny_case(m=32, n=24, mgridlev=2, tol=1e-10, gmres_restart=40, beam=True)
o = ob.OracleSim(p, b); steps(o, 2)
M, K, H = o.beam_matrices()
nb = b[1].npts
o.beam_newmark(0, 3, np.zeros(2 * nb), np.zeros(M.shape[0]), np.zeros(M.shape[0]), np.zeros(M.shape[0]))
o.close()
# a sub-domain deck (rigid points included, one torsional body)
p, b = decks.tiny_case(m=48, n=40, mgridlev=2, tol=1e-10, gmres_restart=40)
q = abi.IBParams.from_buffer_copy(p); q.sub_domain = 1
sd = abi.make_subdomain(-0.8, 1.0, -0.45, 0.4, 4.0 / 48, full_rectangular=0, droptol=1e-300)
o = ob.OracleSim(q, b, subdomain=sd); o.subdomain_info(); o.subdomain_matrix(); steps(o, 1); o.close()
# the 1-D and 2-D DST entry points
ob.dst1d(rng.standard_normal(37))
p = abi.make_params(24, 20, 1, 0.01, 100.0, 1.0, 0.5, 0.5)
o = ob.OracleSim(p, []); o.dst2d(rng.standard_normal(o.P)); o.close()
print("sanitized oracle: all paths ran clean", flush=True)
"""


def runtime(name):
    try:
        path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True,
                              check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_runs_clean_under_asan_and_ubsan():
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan}:{ubsan}",
        "IB_ORACLE_LIB": os.path.join(ORACLE, "liboracle_san.so"),
        # Python's own allocations are not ours to judge; OpenMP threads at 2
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "OMP_NUM_THREADS": "2",
        "PYTHONPATH": ROOT + os.pathsep + env.get("PYTHONPATH", ""),
    })
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=600)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "all paths ran clean" in r.stdout, tail
