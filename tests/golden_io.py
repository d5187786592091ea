"""Readers for the reference's output files kept under tests/golden/ and the reference's own comparison rule.

diff_numeric (tests/2exec/diff_numeric.awk:18-61 of the reference) fails a pair of numbers when |a-b| > abs_tol AND
the relative difference > rel_tol; comp2exec (tests/2exec/comp2exec:147-240) uses, as -log10: CrossSec-* abs 1e-14 /
rel 1e-8, mueller abs 1e-9 / rel 1e-5, IntField/DipPol abs 1e-12 / rel 1e-6.
"""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def path(*p):
    return os.path.join(GOLD, *p)


def crosssec(name, pol="Y"):
    out = {}
    with open(path(name, "CrossSec-" + pol)) as f:
        for line in f:
            k, v = line.split("=")
            out[k.strip()] = float(v)
    return out


def mueller(name):
    return np.loadtxt(path(name, "mueller"), skiprows=1)


def dippol(name, pol="Y"):
    """DipPol-* columns x y z |P|^2 Px.r Px.i Py.r Py.i Pz.r Pz.i -> complex P[3*nd] (file order = dipole order)."""
    a = np.loadtxt(path(name, "DipPol-" + pol), skiprows=1)
    return (a[:, 4::2] + 1j * a[:, 5::2]).reshape(-1)


def geometry(name):
    """geom file -> positions uint16[nd,3], material int32[nd] (0-based)."""
    rows, mats = [], []
    with open(path("geom", name + ".geom")) as f:
        for line in f:
            if not line.strip() or line.startswith("#") or line.startswith("Nmat"):
                continue
            v = [int(t) for t in line.split()]
            rows.append(v[:3])
            mats.append(v[3] - 1 if len(v) > 3 else 0)
    return np.array(rows, np.uint16), np.array(mats, np.int32)


def log_re(name):
    vals = []
    with open(path(name, "log_RE")) as f:
        for line in f:
            if line.startswith("RE_"):
                vals.append(float(line.split("=")[1].split()[0]))
    return np.array(vals)


def numeric_close(a, b, abs_tol, rel_tol):
    """diff_numeric's rule, elementwise: pass unless |a-b| > abs_tol and |a-b|/max(|a|,|b|) > rel_tol."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(a - b)
    scale = np.maximum(np.abs(a), np.abs(b))
    bad = (d > abs_tol) & (d > rel_tol * np.where(scale > 0, scale, 1))
    return not bad.any(), np.argwhere(bad)
