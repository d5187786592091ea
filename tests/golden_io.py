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


def mueller_close(mu, ref, s11):
    """Mueller matrices (rows of 16 elements, or the file layout with a leading theta column) against the reference's:
    rel 1e-5 as in comp2exec, and an absolute floor of 1e-9 relative to the largest S11 instead of comp2exec's bare
    1e-9. Elements that vanish by symmetry (S34/S43 in the forward direction of a sphere) carry the solver's noise,
    about S11 times the residual tolerance (141.5 x 1e-10 for the 16^3 sphere); comp2exec's fixed 1e-9 only holds
    between two runs with the same iteration rounding."""
    return numeric_close(mu, ref, 1e-9 * float(np.abs(s11).max()), 1e-5)


def adj8_fd(pol):
    """Central differences of the reference's Cext and Cabs in Re(m) and Im(m) of each finite-difference voxel of the
    adj8 fixtures: {voxel: [[dCext/dRe m, dCext/dIm m], [dCabs/dRe m, dCabs/dIm m]]}."""
    from cases import ADJ_DELTA, adj8_materials
    _, _, vox = adj8_materials(512)
    out = {}
    for j in vox:
        c = {t: crosssec("adj8/adj8_v%d_%s" % (j, t), pol) for t in ("rp", "rm", "ip", "im")}
        g = np.zeros((2, 2))
        for row, q in enumerate(("Cext", "Cabs")):
            g[row, 0] = (c["rp"][q] - c["rm"][q]) / (2 * ADJ_DELTA)
            g[row, 1] = (c["ip"][q] - c["im"][q]) / (2 * ADJ_DELTA)
        out[int(j)] = g
    return out


def adj8_fd_resolution(c):
    """Resolution of a central difference of a quantity of size c printed to 10 significant digits (one unit in the
    last place of each of the two values, over 2*delta)."""
    from cases import ADJ_DELTA
    ulp = 10.0 ** (np.floor(np.log10(abs(c))) - 9)
    return ulp / ADJ_DELTA


def deps_to_dm(g, m):
    """(dC/dRe eps, dC/dIm eps) -> (dC/dRe m, dC/dIm m) with d eps = 2 m dm."""
    return 2 * (g[0] * m.real + g[1] * m.imag), 2 * (g[1] * m.real - g[0] * m.imag)
