"""Rebuilds the inputs of a golden case (tests/golden/manifest.json) for the oracle
and for the product, so every parity test reads the same configuration the
reference was run with."""
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_MANIFEST = None


def manifest():
    global _MANIFEST
    if _MANIFEST is None:
        with open(os.path.join(GOLDEN, "manifest.json")) as f:
            _MANIFEST = json.load(f)
    return _MANIFEST


def args(name):
    return manifest()["cases"][name]["args"]


def vec(a, key, default=None):
    if key not in a:
        return default
    return [float(x) for x in a[key].split(",")]


def _rotate_xyz(v, r):
    # TVector3D::RotateSelfXYZ, as the reference's field constructors apply it to the field
    x, y, z = v
    for axis, ang in enumerate(r):
        s, c = math.sin(ang), math.cos(ang)
        if axis == 0:
            y, z = c * y - s * z, s * y + c * z
        elif axis == 1:
            z, x = c * z - s * x, s * z + c * x
        else:
            x, y = c * x - s * y, s * x + c * y
    return [x, y, z]


def undulators(a):
    """Field elements of a case in the harness's order: und1..4, gauss1..2, box1..2.
    Gaussian / box entries carry kind 1 / 2, the raw field (rot_field: rotated as the
    constructors do), width, center and rotations."""
    out = []
    for k in ("und1", "und2", "und3", "und4"):
        u = vec(a, k)
        if not u:
            continue
        out.append(dict(field=u[0:3], period=u[3:6], nperiods=int(u[6]), phase=u[7] if len(u) > 7 else 0.0,
                        taper=u[8] if len(u) > 8 else 0.0, center=u[9:12] if len(u) > 11 else [0.0, 0.0, 0.0]))
    for kind, keys, lo in ((1, ("gauss1", "gauss2"), 6), (2, ("box1", "box2"), 3)):
        for k in keys:
            g = vec(a, k)
            if not g or len(g) < lo:
                continue
            w = g[3:6] if len(g) >= 6 else [0.0, 0.0, 0.0]
            c = g[6:9] if len(g) >= 9 else [0.0, 0.0, 0.0]
            r = g[9:12] if len(g) >= 12 else [0.0, 0.0, 0.0]
            out.append(dict(kind=kind, raw_field=g[0:3], field=_rotate_xyz(g[0:3], r), width=w, center=c,
                            rotations=r))
    return out


def field_maps(a):
    """Field maps of a case in the harness's order (gridfile1, gridfile2, then gridmap), as
    keyword arguments of sr.add_bfield_file / sr.add_bfield_interpolated, with the file
    paths resolved under tests/golden."""
    out = []
    for k in ("1", "2"):
        if "gridfile" + k not in a:
            continue
        out.append(("file", dict(ifile=os.path.join(GOLDEN, a["gridfile" + k]), iformat=a.get("gridfmt" + k, "OSCARS"),
                                 rotations=vec(a, "gridrot" + k), translation=vec(a, "gridtrans" + k),
                                 scale=vec(a, "gridscale" + k))))
    if "gridmap" in a:
        mapping = []
        for e in a["gridmap"].split(";"):
            p, f = e.split(":", 1)
            mapping.append([float(p), os.path.join(GOLDEN, f)])
        out.append(("interpolated", dict(mapping=mapping, iformat=a.get("gridmapfmt", "OSCARS"),
                                         parameter=float(a["gridparam"]), rotations=vec(a, "gridmaprot"),
                                         translation=vec(a, "gridmaptrans"), scale=vec(a, "gridmapscale"))))
    return out


def add_field_maps(osr, a):
    for kind, kw in field_maps(a):
        if kind == "file":
            osr.add_bfield_file(**kw)
        else:
            osr.add_bfield_interpolated(**kw)


def pole(a):
    u = vec(a, "poles")
    if not u:
        return None
    return dict(field=u[0:3], period=u[3:6], nperiods=int(u[6]))


def beam(a):
    return dict(energy=float(a.get("energy", 3)), x0=vec(a, "x0", [0, 0, 0]), d0=vec(a, "d0", [0, 0, 1]),
                current=float(a.get("current", 0.5)))


def ct(a):
    c = vec(a, "ct", [0, 3])
    return c[0], c[1], int(float(a["npoints"]))


def rect(a):
    n = vec(a, "nx")
    return dict(plane=a.get("plane", "XY"), nx1=int(n[0]), nx2=int(n[1]), w1=vec(a, "width")[0], w2=vec(a, "width")[1],
                rot=vec(a, "rot", [0, 0, 0]), trans=vec(a, "trans", [0, 0, 30]), normal=int(float(a.get("normal", 0))))


def points(a):
    p = np.array(vec(a, "points")).reshape(-1, 6)
    return p[:, :3].copy(), p[:, 3:].copy()


def oracle_setup(O, name, npoints=None):
    """Oracle fields, ideal particle, trajectory and dt for a case."""
    a = args(name)
    F = O.FieldSet(undulators(a), pole(a), None)
    b = beam(a)
    p = O.ideal_electron(b["energy"], b["x0"], b["d0"], b["current"])
    c0, c1, n = ct(a)
    t, dt = O.trajectory(F, p, c0, c1, npoints or n)
    return F, p, t, dt
