"""Readers for the reference's fixture formats and the comparison
tolerances of the reference's tools/compare.py (without pynbody)."""
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(*parts):
    return os.path.join(GOLDEN, *parts)


def parse_paramfile(path):
    """Commands of a paramfile as (name, [args]) with comments stripped."""
    cmds = []
    with open(path) as f:
        for ln in f:
            ln = ln.split("#")[0].strip()
            if not ln:
                continue
            parts = ln.split()
            cmds.append((parts[0].lower(), parts[1:]))
    return cmds


def cosmology_from_paramfile(path):
    """Cosmology keyword arguments set by a paramfile (reference defaults
    otherwise, genetIC/src/ic.hpp:196-206)."""
    c = dict(Om=0.3, Ol=0.0, Ob=0.0, hubble=0.701, s8=0.0, ns=0.96, z=0.0)
    names = {"om": "Om", "ol": "Ol", "ob": "Ob", "hubble": "hubble", "s8": "s8", "ns": "ns", "zin": "z"}
    for name, args in parse_paramfile(path):
        if name in names:
            c[names[name]] = float(args[0])
    return c


def fortran_blocks(raw):
    out = []
    p = 0
    while p < len(raw):
        n = struct.unpack("<i", raw[p:p + 4])[0]
        out.append(raw[p + 4:p + 4 + n])
        p += n + 8
    return out


def read_gadget(path):
    """Single-file gadget2/3 snapshot with fixed masses: header fields, pos, vel, ids."""
    b = fortran_blocks(open(path, "rb").read())
    hdr = b[0]
    npart = struct.unpack("<6i", hdr[:24])
    mass = struct.unpack("<6d", hdr[24:72])
    time, redshift = struct.unpack("<2d", hdr[72:88])
    pos = np.frombuffer(b[1], dtype="<f4").reshape(-1, 3)
    vel = np.frombuffer(b[2], dtype="<f4").reshape(-1, 3)
    ids = np.frombuffer(b[3], dtype="<u8")
    return dict(npart=npart, mass=mass, time=time, redshift=redshift, header=hdr, pos=pos, vel=vel, ids=ids)


TIPSY_DARK = np.dtype([("mass", "<f4"), ("pos", "<f4", 3), ("vel", "<f4", 3), ("eps", "<f4"), ("phi", "<f4")])


def read_tipsy(path):
    raw = open(path, "rb").read()
    a, n, ndim, ngas, ndark, nstar = struct.unpack("<diiiii", raw[:28])
    dark = np.frombuffer(raw[32 + 48 * ngas:32 + 48 * ngas + 36 * ndark], dtype=TIPSY_DARK)
    return dict(a=a, n=n, ngas=ngas, ndark=ndark, nstar=nstar, dark=dark)


def read_ps(path):
    return np.loadtxt(path)


def assert_decimal(actual, desired, decimal, what=""):
    """numpy.testing.assert_almost_equal semantics used by compare.py:
    abs(desired-actual) < 1.5 * 10**(-decimal)."""
    err = np.max(np.abs(np.asarray(actual, dtype=np.float64) - np.asarray(desired, dtype=np.float64)))
    assert err < 1.5 * 10.0 ** (-decimal), "%s: max abs error %g exceeds decimal=%d" % (what, err, decimal)
