"""Readers for the reference's input decks and the named workloads.

* ``read_ib_inp``: namelist &READ_PARAMETERS of input/ib.inp
  (src/parameters.F90:51-78).
* ``read_body``: one input/body.NNN.inp (src/grid.F90:252-288): npts, fam,
  [fam 2: gamma0 (deg), alpha0 (deg), itheta, ktheta, ctheta], then x y per line.
  Family 3 (a beam; the reference reads no parameters for its stub) takes the
  same five lines: pitch, alpha0, mass per unit length, EI, damping.
* ``read_case``: ib.inp + body.001.inp, body.002.inp, ... until one is missing.
* ``cylinder_case``, ``plate_case``, ``tiny_case``: synthetic workloads with the
  shapes of the named configurations (circular cylinder 200x200x3; hinged
  plate 400x400x4), generated here because the reference ships no such decks.
"""
import math
import os
import re

import numpy as np

from .abi import BodySpec, make_params

_NML_KEYS = ["istart", "istop", "isave", "irestart", "m", "n", "dt", "re", "fsi_tol", "atol", "rtol", "len",
             "offsetx", "offsety", "mgridlev", "compute_pressure", "dimen", "num_stat", "motion_prescribed",
             "sub_domain", "sub_domain_full_rectangular", "sub_domain_precomputed", "sdxi", "sdxe", "sdyi",
             "sdye", "delta_sd"]


def _fortran_value(tok):
    t = tok.strip().rstrip(",").strip()
    u = t.upper()
    if u in ("T", ".TRUE.", "TRUE"):
        return True
    if u in ("F", ".FALSE.", "FALSE"):
        return False
    u = u.replace("D", "E")
    if re.fullmatch(r"[+-]?\d+", u):
        return int(u)
    return float(u)


def read_ib_inp(path):
    """Parse the namelist group READ_PARAMETERS; keys are lower-cased."""
    txt = open(path).read()
    m = re.search(r"&\s*read_parameters(.*?)/", txt, re.IGNORECASE | re.DOTALL)
    if not m:
        raise ValueError(f"{path}: no &READ_PARAMETERS group")
    body = m.group(1)
    out = {}
    for key, val in re.findall(r"([A-Za-z_][A-Za-z0-9_]*)\s*=\s*([^=\n]+?)(?=,?\s*(?:[A-Za-z_][A-Za-z0-9_]*\s*=|$))",
                               body, re.DOTALL):
        k = key.lower()
        if k not in _NML_KEYS:
            raise ValueError(f"{path}: unknown namelist key {key}")
        out[k] = _fortran_value(val)
    return out


def read_body(path):
    vals = open(path).read().split()
    npts = int(vals[0]); fam = int(vals[1]); pos = 2
    gamma0 = alpha0 = itheta = ktheta = ctheta = 0.0
    if fam in (2, 3):
        gamma0 = float(vals[2].replace("D", "E")) * math.pi / 180
        alpha0 = float(vals[3].replace("D", "E")) * math.pi / 180
        itheta, ktheta, ctheta = (float(v.replace("D", "E")) for v in vals[4:7])
        pos = 7
    xy = np.array([float(v.replace("D", "E")) for v in vals[pos:pos + 2 * npts]]).reshape(npts, 2)
    return BodySpec(xy[:, 0], xy[:, 1], fam, gamma0, alpha0, itheta, ktheta, ctheta)


def read_case(input_dir, **overrides):
    nml = read_ib_inp(os.path.join(input_dir, "ib.inp"))
    nml.update(overrides)
    bodies = []
    i = 1
    while os.path.exists(os.path.join(input_dir, f"body.{i:03d}.inp")):
        bodies.append(read_body(os.path.join(input_dir, f"body.{i:03d}.inp")))
        i += 1
    p = make_params(nml["m"], nml["n"], nml["mgridlev"], nml["dt"], nml["re"], nml["len"], nml["offsetx"],
                    nml["offsety"], fsi_tol=nml.get("fsi_tol", 1e-7), atol=nml.get("atol", 1e-7),
                    rtol=nml.get("rtol", 1e-7), istart=nml.get("istart", 0),
                    num_stat=int(bool(nml.get("num_stat", False))), sub_domain=int(bool(nml.get("sub_domain", False))))
    return p, bodies, nml


def subdomain_of(nml):
    """The sub-domain parameters of a deck (abi.IBSubdomain), or None when
    sub_domain = F; interp_points, droptol and the sparsity window take the
    reference's values."""
    from .abi import make_subdomain
    if not nml.get("sub_domain", False):
        return None
    return make_subdomain(nml["sdxi"], nml["sdxe"], nml["sdyi"], nml["sdye"], nml["delta_sd"],
                          full_rectangular=int(bool(nml.get("sub_domain_full_rectangular", False))),
                          precomputed=int(bool(nml.get("sub_domain_precomputed", False))))


def write_body(path, b):
    with open(path, "w") as fh:
        fh.write(f"{b.npts}\n{b.fam}\n")
        if b.fam in (2, 3):
            fh.write(f"{b.gamma0 * 180 / math.pi:.16f}\n{b.alpha0 * 180 / math.pi:.16f}\n")
            fh.write(f"{b.itheta:.16e}\n{b.ktheta:.16e}\n{b.ctheta:.16e}\n")
        for x, y in zip(b.x, b.y):
            fh.write(f"{x:.16f}   {y:.16f}\n")


def write_ib_inp(path, p, istop=10, isave=10, compute_pressure=False, subdomain=None):
    with open(path, "w") as fh:
        fh.write("&READ_PARAMETERS\n")
        fh.write(f"ISTART = {p.istart},\nISTOP = {istop},\nISAVE = {isave},\nIRESTART = 0,\n")
        fh.write(f"M = {p.m},\nN = {p.n},\nDT = {p.dt!r},\nRe = {p.re!r},\n")
        fh.write(f"FSI_TOL = {p.fsi_tol!r},\nATOL = {p.atol!r},\nRTOL = {p.rtol!r},\n")
        fh.write(f"LEN = {p.len!r},\nOFFSETX = {p.offsetx!r},\nOFFSETY = {p.offsety!r},\nMGRIDLEV = {p.mgridlev},\n")
        ns = "T" if getattr(p, "num_stat", 0) else "F"
        cpf = "T" if compute_pressure else "F"
        fh.write(f"COMPUTE_PRESSURE = {cpf},\ndimen = 2,\nnum_stat={ns},\nmotion_prescribed=F,\n")
        if subdomain is None:
            fh.write("sub_domain = F,\n/\n")
        else:
            sd = subdomain
            fr = "T" if sd.full_rectangular else "F"
            pc = "T" if sd.precomputed else "F"
            fh.write(f"sub_domain = T,\nsub_domain_full_rectangular = {fr},\nsub_domain_precomputed = {pc},\n")
            fh.write(f"sdxi = {sd.sdxi!r},\nsdxe = {sd.sdxe!r},\nsdyi = {sd.sdyi!r},\nsdye = {sd.sdye!r},\n")
            fh.write(f"delta_sd = {sd.delta_sd!r}\n/\n")


# --------------------------------------------------------------------------
# named synthetic workloads
# --------------------------------------------------------------------------

def cylinder_body(npts=157, radius=0.5):
    t = 2 * np.pi * np.arange(npts) / npts
    return BodySpec(radius * np.cos(t), radius * np.sin(t), fam=1)


def cylinder_case(m=200, n=200, mgridlev=3, dt=0.004, re=100.0, npts=157, tol=1e-7, **kw):
    """Rigid circular cylinder D=1, Re=100, finest domain [-1,3]x[-2,2] (dx = 4/m)."""
    p = make_params(m, n, mgridlev, dt, re, 4.0, 1.0, 2.0, fsi_tol=tol, atol=tol, rtol=tol, **kw)
    return p, [cylinder_body(npts)]


def plate_body(npts=101, length=1.0, pitch_deg=10.0, mass_ratio=0.2, k_theta=0.1, c_theta=0.0):
    """Flat plate pinned at its leading edge (origin) as a torsional body.

    Points j = 0..npts-1 at j*ds along angle gamma0 = -pitch (trailing edge
    below the hinge, flow in +x).  Rotational inertia about the hinge of a thin
    plate with mass ratio M = rho_s h / (rho_f L): I = M L^3 / 3 (unit fluid
    density); the spring stiffness is taken as k_theta directly.
    """
    g = -math.radians(pitch_deg)
    ds = length / (npts - 1)
    r = ds * np.arange(npts)
    return BodySpec(r * math.cos(g), r * math.sin(g), fam=2, gamma0=g, alpha0=0.0,
                    itheta=mass_ratio * length ** 3 / 3.0, ktheta=k_theta, ctheta=c_theta)


def beam_body(npts=101, length=1.0, pitch_deg=10.0, mass_ratio=0.2, k_b=0.1, damping=0.0):
    """Flexible flat plate clamped at its leading edge (origin) as a family-3
    body: a linear Euler-Bernoulli beam of (npts - 1) / 2 Hermite elements.

    Points j = 0..npts-1 at j*ds along angle -pitch (trailing edge below the
    leading edge, flow in +x).  Mass per unit length = mass ratio (rho_s h /
    (rho_f L) with L = 1 and unit fluid density), bending stiffness EI = K_B
    (EI / (rho_f U^2 L^3)), damping C = damping * M.  The deck lines of
    family 2 carry them: itheta = mass per length, ktheta = EI, ctheta =
    damping.
    """
    g = -math.radians(pitch_deg)
    ds = length / (npts - 1)
    r = ds * np.arange(npts)
    return BodySpec(r * math.cos(g), r * math.sin(g), fam=3, gamma0=g, alpha0=0.0,
                    itheta=mass_ratio * length, ktheta=k_b, ctheta=damping)


def beam_plate_case(m=400, n=400, mgridlev=4, dt=0.001, re=200.0, npts=101, tol=1e-10, gmres_restart=250, **kw):
    """The flexible-plate FSI workload with the plate as an Euler-Bernoulli
    beam (50 elements, 101 points, mass ratio 0.2, K_B = 0.1, pitch 10 deg,
    Re 200) at 400x400 cells, 4 levels, dx = 0.01, dt = 1e-3, tolerance 1e-10."""
    p = make_params(m, n, mgridlev, dt, re, 4.0, 1.0, 2.0, fsi_tol=tol, atol=tol, rtol=tol,
                    gmres_restart=gmres_restart, **kw)
    return p, [beam_body(npts)]


def plate_case(m=400, n=400, mgridlev=4, dt=0.001, re=200.0, npts=101, tol=1e-10, gmres_restart=250, **kw):
    """Hinged flexible-plate FSI workload: 400x400 cells, 4 levels, dx=0.01.

    GMRES restart 250 >= nf = 202 (full GMRES; the reference takes it from
    -ksp_gmres_restart through KSPSetFromOptions, src/variables.F90:571): with
    the default restart 30 the 1e-10 tolerance is unreachable for this
    operator (condition number ~2e7) and every solve stops at max_it.
    """
    p = make_params(m, n, mgridlev, dt, re, 4.0, 1.0, 2.0, fsi_tol=tol, atol=tol, rtol=tol,
                    gmres_restart=gmres_restart, **kw)
    return p, [plate_body(npts)]


def ellipse_arclength(npts, chord=1.0, thickness=0.12, x0=0.0, y0=0.0):
    """npts points equispaced in arc length on an ellipse with the given chord
    (x extent) and thickness (y extent), leading edge at (x0, y0)."""
    a, b = 0.5 * chord, 0.5 * thickness
    t = np.linspace(0.0, 2.0 * np.pi, 20001)
    x, y = a * np.cos(t), b * np.sin(t)
    s = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(x), np.diff(y)))])
    tt = np.interp(np.linspace(0.0, s[-1], npts, endpoint=False), s, t)
    return x0 + a + a * np.cos(tt), y0 + b * np.sin(tt)


def airfoil_flap_case(m=1024, n=512, mgridlev=5, dt=0.0005, re=1000.0, n_airfoil=300, n_flap=100,
                      k_theta=0.05, mass_ratio=0.5, flap_deg=10.0, tol=1e-10, gmres_restart=1000, **kw):
    """Two-body system: rigid elliptic airfoil (chord 1, thickness 0.12,
    leading edge at the origin) and a flat flap (chord 0.3) hinged 0.01
    behind its trailing edge on a torsional spring, deflected flap_deg
    downward.  dx = 0.004 at 1024x512 (finest domain [-1, 3.096] x
    [-1.024, 1.024]); the flap's inertia about the hinge is
    mass_ratio * c^3 / 3.  The flap's points are 0.76 dx apart, which makes
    the Schur system ill conditioned (cond ~1e8 at a quarter of this size):
    GMRES(250) stagnates above 1e-10 at nf = 800, so the restart length
    defaults to 1000 (full GMRES)."""
    length = 0.004 * m
    p = make_params(m, n, mgridlev, dt, re, length, 1.0, 0.004 * n / 2, fsi_tol=tol, atol=tol, rtol=tol,
                    gmres_restart=gmres_restart, **kw)
    xa, ya = ellipse_arclength(n_airfoil)
    chord = 0.3
    g = -math.radians(flap_deg)
    r = chord / (n_flap - 1) * np.arange(n_flap)
    hx, hy = 1.01, 0.0
    flap = BodySpec(hx + r * math.cos(g), hy + r * math.sin(g), fam=2, gamma0=g, alpha0=0.0,
                    itheta=mass_ratio * chord ** 3 / 3.0, ktheta=k_theta, ctheta=0.0)
    return p, [BodySpec(xa, ya, fam=1), flap]


def tiny_case(m=32, n=24, mgridlev=2, torsional=True, tol=1e-12, dt=0.01, re=50.0, beam=False, **kw):
    """Small grid for oracle-vs-GPU parity: a coarse cylinder plus a hinged
    plate (or, with beam=True, a clamped flexible plate of family 3)."""
    p = make_params(m, n, mgridlev, dt, re, 4.0, 1.6, 1.5, fsi_tol=tol, atol=tol, rtol=tol, **kw)
    dx = 4.0 / m
    nc = max(8, int(round(np.pi * 0.6 / dx)))
    t = 2 * np.pi * np.arange(nc) / nc
    bodies = [BodySpec(0.3 * np.cos(t) - 0.4, 0.3 * np.sin(t), fam=1)]
    if beam:
        npl = max(5, int(round(0.6 / dx)) + 1)
        npl += 1 - npl % 2  # odd: two points per element
        g = -math.radians(15.0)
        r = np.linspace(0, 0.6, npl)
        bodies.append(BodySpec(0.2 + r * math.cos(g), 0.05 + r * math.sin(g), fam=3, gamma0=g,
                               itheta=0.05, ktheta=0.02, ctheta=0.0))
    elif torsional:
        npl = max(4, int(round(0.6 / dx)) + 1)
        g = -math.radians(15.0)
        r = np.linspace(0, 0.6, npl)
        bodies.append(BodySpec(0.2 + r * math.cos(g), 0.05 + r * math.sin(g), fam=2, gamma0=g,
                               itheta=0.05, ktheta=0.2, ctheta=0.0))
    return p, bodies


def ellipse_batch_cases(count=64, seed=20240611, m_range=(256, 512), levels=(3, 5), npts_range=(100, 400),
                        aspect=(1.0, 4.0), aoa_deg=(0.0, 20.0), re_range=(100.0, 400.0), cfl=0.2, num_stat=1,
                        tol=1e-10, **kw):
    """Independent rigid ellipses for the batched form, drawn from a PCG64
    stream: aspect ratio, angle of attack and Re uniform in their ranges; m
    and n uniform over the multiples of 4 in m_range; mgridlev uniform in
    levels; the body point count uniform in npts_range.  Each ellipse has
    chord 1 and thickness 1/aspect, centred in its finest domain and pitched
    nose-up by the angle of attack; its points are equispaced in arc length,
    so the grid spacing is set to the point spacing (dx = perimeter / npts)
    and dt = cfl * dx.  Returns a list of (params, [body]) pairs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cases = []
    for _ in range(count):
        ar = rng.uniform(*aspect)
        aoa = math.radians(rng.uniform(*aoa_deg))
        re = rng.uniform(*re_range)
        m = 4 * int(rng.integers(m_range[0] // 4, m_range[1] // 4 + 1))
        n = 4 * int(rng.integers(m_range[0] // 4, m_range[1] // 4 + 1))
        lev = int(rng.integers(levels[0], levels[1] + 1))
        npts = int(rng.integers(npts_range[0], npts_range[1] + 1))
        x, y = ellipse_arclength(npts, chord=1.0, thickness=1.0 / ar, x0=-0.5)
        perim = float(np.sum(np.hypot(np.diff(np.append(x, x[0])), np.diff(np.append(y, y[0])))))
        dx = perim / npts
        c, s = math.cos(-aoa), math.sin(-aoa)
        xb, yb = c * x - s * y, s * x + c * y
        length = m * dx
        p = make_params(m, n, lev, cfl * dx, re, length, 0.5 * length, 0.5 * n * dx, fsi_tol=tol, atol=tol,
                        rtol=tol, num_stat=num_stat, **kw)
        cases.append((p, [BodySpec(xb, yb, fam=1)]))
    return cases
