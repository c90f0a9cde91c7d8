"""Readers for the dataRR_IIIIIII.var snapshot files (write_variables,
src/variables.F90:709-803; the reference's own reader is output/getdata.m)
and the pressureRR_IIIIIII.var files (write_pressure,
src/operators_pressure.F90:857-880).

Fortran unformatted sequential: every record is framed by 4-byte lengths;
integers are int32 (PetscInt), reals float64.
"""
import numpy as np


def _records(path):
    raw = open(path, "rb").read()
    pos = 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, "<i4", 1, pos)[0])
        yield raw[pos + 4:pos + 4 + n]
        assert int(np.frombuffer(raw, "<i4", 1, pos + 4 + n)[0]) == n, "corrupt record marker"
        pos += 8 + n


def read_var(path):
    rec = list(_records(path))
    i4 = lambda b: np.frombuffer(b, "<i4")
    f8 = lambda b: np.frombuffer(b, "<f8")
    m, n, L, nb = (int(v) for v in i4(rec[0]))
    re, dt, len_, offx, offy = f8(rec[1])
    out = {"m": m, "n": n, "mgridlev": L, "nb": nb, "re": re, "dt": dt, "len": len_,
           "offsetx": offx, "offsety": offy, "levels": []}
    r = 5
    for _ in range(L):
        names = ("omega", "s", "qx", "q0px", "q0rx", "qy", "q0py", "q0ry")
        out["levels"].append({k: f8(rec[r + i]).copy() for i, k in enumerate(names)})
        r += 8
    out["rhs_old"] = f8(rec[r]).copy(); r += 1
    out["rot_angle"], out["rox"], out["roy"] = f8(rec[r]); r += 1
    out["xfm"], out["xfi"], out["xfe"] = (int(v) for v in i4(rec[r])); r += 1
    out["f"] = f8(rec[r]).copy(); out["xb"] = f8(rec[r + 1]).copy(); out["yb"] = f8(rec[r + 2]).copy(); r += 3
    out["mt_body"] = int(i4(rec[r])[0]); r += 1
    if out["mt_body"] > 0:
        th = f8(rec[r]); nt = out["mt_body"]
        out["theta"], out["thetad"], out["thetadd"] = th[:nt].copy(), th[nt:2 * nt].copy(), th[2 * nt:].copy()
        r += 1
    out["f_rdst"] = f8(rec[r]).copy()
    return out


def read_pressure(path):
    """{"m", "n", "levels": [array (n, m) per level]} of a pressure .var file."""
    rec = list(_records(path))
    xpi, xpe, xpm, ypi, ype, ypm = (int(v) for v in np.frombuffer(rec[0], "<i4"))
    lv = [np.frombuffer(r, "<f8").reshape(ypm, xpm).copy() for r in rec[1:]]
    return {"m": xpm, "n": ypm, "levels": lv}
