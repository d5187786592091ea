"""Pins the CPU oracle (oracle/adda_oracle.c) against the reference's golden outputs, known answers and the
mathematics. Runs on the CPU only."""
import math
import os

import numpy as np
import pytest

import golden_io as G
from cases import Case, make_shape, read_shape, oracle, ptr, PROP_Z, POL_X, POL_Y


def test_fftfit_rules():
    lib = oracle()
    assert lib.or_fftfit(32, 1, 0) == 32                # sample log: box 16 -> grid 32
    assert lib.or_fftfit(104, 1, 0) == 104              # FFTW rule allows one factor 13
    assert lib.or_fftfit(104, 1, 1) == 108              # 2/3/5 rule
    assert lib.or_fftfit(256, 1, 1) == 256 and lib.or_fftfit(640, 1, 1) == 640
    assert lib.or_fftfit(14, 1, 0) == 14 and lib.or_fftfit(14, 1, 1) == 16
    for x in range(2, 300):
        y = lib.or_fftfit(x, 1, 1)
        r = y
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        assert y >= x and y % 2 == 0 and r == 1
        assert all(not (z % 2 == 0 and _smooth235(z)) for z in range(x, y))


def _smooth235(z):
    for p in (2, 3, 5):
        while z % p == 0:
            z //= p
    return z == 1


def test_green_poi_static_limit_and_parity():
    """k -> 0 gives the static dipole tensor (3 qq - I)/r^3; G is even in r with xy,xz,yz odd per reflected axis."""
    lib = oracle()
    res = (G_ := np.zeros(12))  # noqa: F841
    out = np.zeros(12)
    d = 0.37
    for (i, j, k) in [(1, 0, 0), (2, -1, 3), (-4, 5, 1), (0, 0, -7)]:
        lib.or_green_poi(i, j, k, 1e-9, d, ptr(out))
        g = out[0::2] + 1j * out[1::2]
        r = math.sqrt(i * i + j * j + k * k)
        q = np.array([i, j, k]) / r
        qq = np.outer(q, q)
        stat = (3 * qq - np.eye(3)) / (r * d) ** 3
        full = np.array([[g[0], g[1], g[2]], [g[1], g[3], g[4]], [g[2], g[4], g[5]]])
        assert np.allclose(full.real, stat, rtol=1e-9, atol=1e-12 / d ** 3)
        lib.or_green_poi(i, -j, k, 1.0, d, ptr(out))
        gm = out[0::2] + 1j * out[1::2]
        lib.or_green_poi(i, j, k, 1.0, d, ptr(out))
        g1 = out[0::2] + 1j * out[1::2]
        sign = np.array([1, -1, 1, 1, -1, 1]) if j != 0 else np.ones(6)
        assert np.allclose(gm, sign * g1, rtol=1e-15, atol=0)


def test_green_poi_far_field():
    """kr >> 1: G -> k^2 e^{ikr}/r (I - qq) (radiation zone of the dipole field)."""
    lib = oracle()
    out = np.zeros(12)
    d, i, j, k = 50.0, 3, 4, 12  # r = 13 d = 650, kr = 650
    lib.or_green_poi(i, j, k, 1.0, d, ptr(out))
    g = out[0::2] + 1j * out[1::2]
    r = 13 * d
    q = np.array([i, j, k]) / 13
    ref = np.exp(1j * r) / r * (np.eye(3) - np.outer(q, q))
    full = np.array([[g[0], g[1], g[2]], [g[1], g[3], g[4]], [g[2], g[4], g[5]]])
    assert np.abs(full - ref).max() < 3.0 / r * 1e-2


@pytest.mark.parametrize("n", list(range(1, 41)) + [48, 60, 64, 90, 96, 100, 108, 120, 128])
def test_fft_vs_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        a = x.copy()
        oracle().or_fft1d(ptr(a), n, 1, 1, n, sign)
        k = np.arange(n)
        ref = np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x
        assert np.abs(a - ref).max() <= 1e-12 * max(1, np.abs(ref).max())


def _small_cases():
    out = []
    pos, b = make_shape("ellipsoid", 8, 0.7, 0.5)
    out.append(Case(pos, b, 1.3 + 0.2j))
    pos, b = make_shape("box", 6, 1.0, 0.5)
    rng = np.random.default_rng(5)
    out.append(Case(pos, b, [1.5, 1.2 + 0.3j], mat=rng.integers(0, 2, len(pos))))
    pos, b = make_shape("sphere", 6)
    out.append(Case(pos, b, 2.0 + 0.1j, reduced_fft=False))
    return out


@pytest.mark.parametrize("ci", range(3))
def test_fft_matvec_equals_direct_sum(ci):
    """The FFT convolution with the reduced D-matrix is exactly the direct sum that ADDA's SPARSE MatVec computes."""
    c = _small_cases()[ci]
    cc, ccs = c.cc(PROP_Z, POL_Y)
    d = c.dmatrix()
    rng = np.random.default_rng(ci)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    for her in (False, True):
        y1, n1 = c.matvec_fft(d, ccs, x, her)
        y2, n2 = c.matvec_direct(ccs, x, her)
        assert np.abs(y1 - y2).max() < 1e-12 * np.abs(y2).max()
        assert abs(n1 - n2) < 1e-12 * n2


def test_reduced_equals_full_dmatrix():
    pos, b = make_shape("ellipsoid", 10, 0.8, 0.6)
    c1, c2 = Case(pos, b, 1.5 + 0.05j), Case(pos, b, 1.5 + 0.05j, reduced_fft=False)
    cc, ccs = c1.cc(PROP_Z, POL_Y)
    x = np.random.default_rng(9).standard_normal(3 * c1.nd) + 0.5j
    y1, _ = c1.matvec_fft(c1.dmatrix(), ccs, x)
    y2, _ = c2.matvec_fft(c2.dmatrix(), ccs, x)
    assert np.abs(y1 - y2).max() < 1e-12 * np.abs(y1).max()


def test_matvec_symmetry_and_her():
    """A is complex symmetric (x.Ay == y.Ax, bilinear) and her gives conj(A conj x) = A^H x."""
    c = _small_cases()[0]
    cc, ccs = c.cc(PROP_Z, POL_Y)
    d = c.dmatrix()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    Ay, _ = c.matvec_fft(d, ccs, y)
    Ax, _ = c.matvec_fft(d, ccs, x)
    assert abs(np.sum(x * Ay) - np.sum(y * Ax)) < 1e-12 * abs(np.sum(x * Ay))
    AHx, _ = c.matvec_fft(d, ccs, x, her=True)
    assert abs(np.vdot(AHx, y) - np.vdot(x, Ay)) < 1e-12 * abs(np.vdot(x, Ay))


def test_couple_constant_kat():
    """sample/run000_sphere_g16_m1.5/log:36: CoupleConstant:0.005259037197+1.843854148e-005i (LDR, m=1.5, dpl=15)."""
    c = G.path("c1_qmr_e5", "log_RE")
    with open(c) as f:
        line = [l for l in f if l.startswith("CoupleConstant")][0]
    cc_ref = complex(line.split(":")[1].strip().replace("i", "j"))
    case = Case(*make_shape("sphere", 16), 1.5, dpl=15)
    cc, _ = case.cc(PROP_Z, POL_Y)
    assert abs(cc[0, 0] - (0.005259037197 + 1.843854148e-05j)) < 1e-12
    assert abs(cc[0, 0] - cc_ref) < 1e-12
    cm, _ = case.cc(PROP_Z, POL_Y, polrel=1)
    m2 = 1.5 ** 2
    assert abs(cm[0, 0] - 3 / (4 * math.pi) * case.gridspace ** 3 * (m2 - 1) / (m2 + 2)) < 1e-16


def _solve_both(case, method, eps, prop, pY, pX, ntheta=None, polrel=0):
    """Both incident polarizations; with ntheta, the yz-plane amplitudes over ADDA's range extended to 360 deg
    (-ntheta N gives nTheta = N+1, src/param.c:1285-1290; extension to 2*(nTheta-1) points for an asymmetric
    particle, src/calculator.c:671-674)."""
    d = case.dmatrix()
    res = {}
    for name, pol in (("Y", pY), ("X", pX)):
        cc, ccs = case.cc(prop, pol, polrel)
        r = case.solve(d, ccs, case.einc(prop, pol), method=method, eps=eps)
        assert r["rc"] == 0
        res[name] = dict(p=r["p"], cext=case.cext(r["p"], prop, pol), cabs=case.cabs(r["p"], cc), niter=r["niter"],
                         hist=r["hist"], nmatvec=r["nmatvec"])
        if ntheta:
            res[name]["epl"] = case.eplane(r["p"], prop, pY, pX, ntheta, 2 * (ntheta - 1))
    return res


def _c1():
    pos, mat = G.geometry("sphere16")
    p, box = read_shape(pos)
    return Case(p, box, 1.5, dpl=15, rule=0)


def test_sample_golden_eps5():
    """Shipped sample (ADDA v1.3b4, FFTW): 22 iterations, Cext = 135.0449047, RE_001..RE_018 history."""
    c = _c1()
    r = _solve_both(c, 0, 1e-5, PROP_Z, POL_Y, POL_X)
    assert r["Y"]["niter"] == 22
    assert abs(r["Y"]["cext"] - 135.0449047) / 135.0449047 < 1e-8
    ref = np.array([1.0, 8.4752662637E-01, 8.2113292044E-01, 4.2639057157E-01, 3.0639031825E-01, 2.2448283848E-01,
                    2.2740255145E-01, 1.5952149116E-01, 6.1602401378E-02, 3.5443250605E-02, 2.9898244023E-02])
    assert np.allclose(r["Y"]["hist"][:11], ref, rtol=1e-8)
    assert np.allclose(r["Y"]["hist"][:19], G.log_re("c1_qmr_e5")[:19], rtol=1e-5)


@pytest.mark.parametrize("method,name", [(0, "c1_qmr"), (1, "c1_bicgstab"), (2, "c1_cgnr"), (3, "c1_bicg"),
                                         (4, "c1_csym"), (5, "c1_bcgs2"), (6, "c1_qmr2")])
def test_c1_solvers_vs_sparse_reference(method, name):
    c = _c1()
    r = _solve_both(c, method, 1e-10, PROP_Z, POL_Y, POL_X, ntheta=181 if method == 0 else None)
    for pol in ("Y", "X"):
        g = G.crosssec(name, pol)
        assert G.numeric_close(r[pol]["cext"], g["Cext"], 1e-14, 1e-8)[0], (r[pol]["cext"], g["Cext"])
        assert abs(r[pol]["cabs"]) < 1e-12 and abs(g["Cabs"]) < 1e-12
    if method == 0:
        mu = c.mueller(r["X"]["epl"], r["Y"]["epl"], 360)
        gm = G.mueller("c1_qmr")
        ok, bad = G.numeric_close(mu, gm[:, 1:], 1e-9, 1e-5)
        assert ok, bad[:5]


@pytest.mark.parametrize("method,name,nmv_per_it", [(5, "c1_bcgs2", 4), (6, "c1_qmr2", 1)])
def test_c1_bcgs2_qmr2_history_vs_sparse_reference(method, name, nmv_per_it):
    """BCGS2 (src/iterative.c:342-488) and QMR_CS_2 (1039-1143): iteration counts, matvec counts and the residual
    history of the reference's log (both polarizations, Y first), to the summation-order drift of a
    direct-sum vs FFT matvec."""
    c = _c1()
    ref = G.log_re(name)
    r = _solve_both(c, method, 1e-10, PROP_Z, POL_Y, POL_X)
    nY = r["Y"]["niter"]
    assert len(ref) == nY + 1 + r["X"]["niter"] + 1
    # the first half of each history to the log's print resolution; later iterations drift with the summation
    # order (QMR_CS_2 reaches a few percent by 1e-10), so the whole history only loosely
    for h, g in ((r["Y"]["hist"], ref[:nY + 1]), (r["X"]["hist"], ref[nY + 1:])):
        k = (len(h) + 1) // 2
        assert np.allclose(h[:k], g[:k], rtol=1e-7)
        assert np.allclose(h, g, rtol=0.15)
    with open(G.path(name, "log_RE")) as f:
        total = int([ln for ln in f if "matrix-vector" in ln][0].split(":")[1])
    # TotalMatVec counts CalcInitField's matvec too; with x_0 = 0 that matvec stands in for the first iteration's
    assert total == sum(r[p]["nmatvec"] for p in ("Y", "X")) == sum(nmv_per_it * r[p]["niter"] for p in ("Y", "X"))


def test_orientation_absorbing_ellipsoid():
    """-orient 30 40 50, m = 1.33+0.05i, LDR: cross sections, Mueller (19 angles), polarization P per dipole."""
    import adda_rot
    pos, _ = G.geometry("ell12")
    p, box = read_shape(pos)
    c = Case(p, box, 1.33 + 0.05j, dpl=12, rule=0)
    prop, pY, pX = (adda_rot.rotate_zyz(30, 40, 50, v) for v in (PROP_Z, POL_Y, POL_X))
    r = _solve_both(c, 0, 1e-10, prop, pY, pX, ntheta=20)
    for pol in ("Y", "X"):
        g = G.crosssec("ell12_orient", pol)
        assert G.numeric_close([r[pol]["cext"], r[pol]["cabs"]], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0]
        ok, bad = G.numeric_close(np.c_[r[pol]["p"].real, r[pol]["p"].imag],
                                  np.c_[G.dippol("ell12_orient", pol).real, G.dippol("ell12_orient", pol).imag],
                                  1e-12, 1e-6)
        assert ok, bad[:5]
    mu = c.mueller(r["X"]["epl"], r["Y"]["epl"], 38)
    assert G.numeric_close(mu, G.mueller("ell12_orient")[:, 1:], 1e-9, 1e-5)[0]


@pytest.mark.parametrize("io", range(4))
def test_batched_workload_orientations(io):
    """The batched workload's physics (ellipsoid 1:0.8:0.5, m = 1.33+0.001i, dpl 13.3) at four of its seeded
    orientations: oracle cross sections vs the reference's for each orientation and polarization."""
    import adda_rot
    from cases import orientations
    a, b, g_ = orientations(4)[io]
    pos, _ = G.geometry("ell12")
    p, box = read_shape(pos)
    c = Case(p, box, 1.33 + 0.001j, dpl=13.3, rule=0)
    prop, pY, pX = (adda_rot.rotate_zyz(a, b, g_, v) for v in (PROP_Z, POL_Y, POL_X))
    r = _solve_both(c, 0, 1e-10, prop, pY, pX)
    for pol in ("Y", "X"):
        g = G.crosssec("ell12_c3o%d" % io, pol)
        assert G.numeric_close([r[pol]["cext"], r[pol]["cabs"]], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0]


def test_cm_polarizability_sphere8():
    pos, _ = G.geometry("sphere8")
    p, box = read_shape(pos)
    c = Case(p, box, 1.1 + 0.1j, dpl=10, rule=0)
    r = _solve_both(c, 0, 1e-10, PROP_Z, POL_Y, POL_X, ntheta=20, polrel=1)
    for pol in ("Y", "X"):
        g = G.crosssec("sphere8_cm", pol)
        assert G.numeric_close([r[pol]["cext"], r[pol]["cabs"]], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0]
    gp = G.dippol("sphere8_cm", "Y")
    assert G.numeric_close(np.c_[r["Y"]["p"].real, r["Y"]["p"].imag], np.c_[gp.real, gp.imag], 1e-12, 1e-6)[0]


def test_multimaterial_box_bicgstab():
    pos, mat = G.geometry("box8mm")
    p, box = read_shape(pos)
    c = Case(p, box, [1.2 + 0.01j, 1.5 + 0.1j, 1.8], mat=mat, dpl=14, rule=0)
    r = _solve_both(c, 1, 1e-10, PROP_Z, POL_Y, POL_X)
    for pol in ("Y", "X"):
        g = G.crosssec("box8_multimat", pol)
        assert G.numeric_close([r[pol]["cext"], r[pol]["cabs"]], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0]
    gp = G.dippol("box8_multimat", "Y")
    assert G.numeric_close(np.c_[r["Y"]["p"].real, r["Y"]["p"].imag], np.c_[gp.real, gp.imag], 1e-12, 1e-6)[0]


# ---------------------------------------------------------------------------------------------------------------------
# adjoint gradient of Cext and Cabs with respect to each dipole's permittivity

@pytest.mark.parametrize("polrel", [0, 1, 2])
def test_couple_constant_deps(polrel):
    """d(cc)/d(eps) against central differences of CoupleConstant in the real and imaginary directions of eps."""
    lib = oracle()
    from cases import OrCplx
    prop, pol = np.array([0.0, 0.6, 0.8]), np.array([1.0, 0.0, 0.0])
    for m in (1.5, 1.33 + 0.001j, 1.8 + 0.05j, 1.2 + 0.3j):
        eps = m * m
        d = lib.or_couple_constant_deps(OrCplx(m.real, m.imag), polrel, 1.0, 0.35, ptr(prop), ptr(pol))
        d = complex(d.re, d.im)
        h = 1e-6
        def cc(e):
            z = np.sqrt(e)
            r = lib.or_couple_constant(OrCplx(z.real, z.imag), polrel, 1.0, 0.35, ptr(prop), ptr(pol))
            return complex(r.re, r.im)
        fr = (cc(eps + h) - cc(eps - h)) / (2 * h)
        fi = (cc(eps + 1j * h) - cc(eps - 1j * h)) / (2j * h)
        assert abs(fr - d) < 1e-7 * abs(d) and abs(fi - d) < 1e-7 * abs(d), (m, d, fr, fi)


def _adj_case(box=5):
    from cases import per_voxel_m
    pos, _ = make_shape("box", box)
    p, b = read_shape(pos)
    nd = len(p)
    return Case(p, b, per_voxel_m(nd), dpl=18, mat=np.arange(nd, dtype=np.int32), rule=1)


@pytest.mark.parametrize("pol", ["Y", "X"])
def test_oracle_adjoint_vs_central_differences(pol):
    """Per-voxel permittivities (box 5^3): adjoint gradient vs central differences (step 1e-6) of forward solves."""
    c = _adj_case()
    d = c.dmatrix()
    vec = POL_Y if pol == "Y" else POL_X
    ge, ga, _, _ = c.adjoint_grad(d, PROP_Z, vec)
    scale = np.abs(np.c_[ge, ga]).max()
    for j in np.random.default_rng(20261205).choice(c.nd, 6, replace=False):
        fd = c.fd_grad(d, PROP_Z, vec, int(j))
        err = np.abs(fd - np.array([ge[j], ga[j]])).max()
        assert err < 1e-7 * scale, (j, fd, ge[j], ga[j])


@pytest.mark.parametrize("pol", ["Y", "X"])
def test_oracle_adjoint_vs_reference_differences(pol):
    """15-material variant (box 8^3): the oracle's adjoint gradient, converted to d/dRe(m) and d/dIm(m) of the voxel
    (d eps = 2 m dm), vs central differences of the reference's own Cext and Cabs (voxel moved alone to material 15,
    m +- 1e-3). Bar: 1e-4 relative plus the resolution of those differences, whose inputs the reference prints to 10
    significant digits."""
    from cases import adj8_materials
    pos, _ = G.geometry("adj8")
    p, b = read_shape(pos)
    m, mat, vox = adj8_materials(len(p))
    c = Case(p, b, m[mat], dpl=18, mat=np.arange(len(p), dtype=np.int32), rule=0)
    vec = POL_Y if pol == "Y" else POL_X
    ge, ga, cext, cabs = c.adjoint_grad(c.dmatrix(), PROP_Z, vec)
    base = G.crosssec("adj8/adj8_base", pol)
    assert abs(cext - base["Cext"]) < 1e-8 * base["Cext"] and abs(cabs - base["Cabs"]) < 1e-8 * base["Cabs"]
    fd = G.adj8_fd(pol)
    for j in vox:
        for row, g in ((0, ge[j]), (1, ga[j])):
            dre, dim = G.deps_to_dm(g, m[mat[j]])
            ref = fd[int(j)][row]
            tol = G.adj8_fd_resolution(base["Cext" if row == 0 else "Cabs"])
            assert abs(dre - ref[0]) < 1e-4 * abs(ref[0]) + tol and abs(dim - ref[1]) < 1e-4 * abs(ref[1]) + tol, \
                (j, row, dre, dim, ref, tol)


# ---------------------------------------------------------------------------------------------------------------------
# particle above a substrate (-surf, image-dipole reflection)

def _surf_frame(prop):
    """prop normalized as -prop does (src/param.c:1372-1382) and the two incident polarizations of
    src/param.c:2116-2131."""
    prop = np.asarray(prop, float) / np.linalg.norm(prop)
    if abs(prop[2]) == 1:
        return prop, np.array([math.copysign(1.0, prop[2]), 0, 0]), np.array([0, 1.0, 0])
    t = math.sqrt(1 - prop[2] ** 2)
    return prop, np.array([prop[0] * prop[2] / t, prop[1] * prop[2] / t, -t]), np.array([-prop[1] / t, prop[0] / t, 0])


# name: geometry, materials, dpl, (hsub, msub or None = perfect reflector), prop, multi-material
SURF_RUNS = {
    "surf_img_above": ("sphere8", 1.5 + 0.01j, 15, (2.0, 1.5 + 0.1j), (0, 0, -1), False),
    "surf_img_below": ("sphere8", 1.5 + 0.01j, 15, (2.0, 1.5 + 0.1j), (0, 0, 1), False),
    "surf_inf_above": ("sphere8", 1.5 + 0.01j, 15, (2.0, None), (0, 0, -1), False),
    "surf_img_oblique": ("box8mm", [1.2 + 0.01j, 1.5 + 0.1j, 1.8], 14, (1.7, 2.0 + 0.5j), (0.5, 0.3, -1), True),
    "surf_img_tir": ("sphere8", 1.5 + 0.01j, 15, (1.9, 1.5 + 0j), (0.8, 0, 0.6), False),
}


def surf_case(name):
    geo, m, dpl, surf, prop, mm = SURF_RUNS[name]
    pos, mat = G.geometry(geo)
    p, box = read_shape(pos)
    return Case(p, box, m, mat=mat if mm else None, dpl=dpl, rule=0, surface=surf), _surf_frame(prop)


def test_refl_img_is_the_perfect_conductor_image():
    """With rcn = -1 the reflected term applied as cReflMatrVec is the free-space field of the image dipole
    p' = (-px, -py, pz) placed at the mirror height: GR.v = G(i, j, ksum + shift).(-v0, -v1, v2)."""
    lib = oracle()
    from cases import OrCplx
    rng = np.random.default_rng(5)
    for (i, j, k, sh) in ((0, 0, 0, 6.0), (2, -1, 3, 4.0), (-3, 2, 1, 7.0), (1, 1, 5, 2.0)):
        r, g = (OrCplx * 6)(), (OrCplx * 6)()
        lib.or_refl_img(i, j, k, sh, OrCplx(-1.0, 0.0), 1.0, 0.4, r)
        lib.or_green_poi(i, j, int(k + sh), 1.0, 0.4, g)
        R = np.array([complex(t.re, t.im) for t in r])
        Gm = np.array([complex(t.re, t.im) for t in g])
        Gfull = np.array([[Gm[0], Gm[1], Gm[2]], [Gm[1], Gm[3], Gm[4]], [Gm[2], Gm[4], Gm[5]]])
        v = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        refl = np.array([R[0] * v[0] + R[1] * v[1] + R[2] * v[2], R[1] * v[0] + R[3] * v[1] + R[4] * v[2],
                         -R[2] * v[0] - R[4] * v[1] + R[5] * v[2]])
        img = Gfull @ np.array([-v[0], -v[1], v[2]])
        assert np.abs(refl - img).max() < 1e-13 * np.abs(img).max()
    # a finite substrate scales the same tensor by (1 - m^2)/(1 + m^2)
    m = 1.5 + 0.1j
    rcn = (1 - m * m) / (1 + m * m)
    r1, r2 = (OrCplx * 6)(), (OrCplx * 6)()
    lib.or_refl_img(1, 2, 3, 3.5, OrCplx(rcn.real, rcn.imag), 1.0, 0.4, r1)
    lib.or_refl_img(1, 2, 3, 3.5, OrCplx(1.0, 0.0), 1.0, 0.4, r2)
    a = np.array([complex(t.re, t.im) for t in r1])
    b = np.array([complex(t.re, t.im) for t in r2])
    assert np.abs(a - rcn * b).max() < 1e-15 * np.abs(a).max()


@pytest.mark.parametrize("reduced", [True, False])
def test_surface_fft_matvec_equals_direct_sum(reduced):
    """The R-matrix correlation (z transformed backward) is exactly the direct sum over source images, the dipole's
    own image included; the operator stays complex symmetric (R_ji = R_ij^T)."""
    pos, b = make_shape("ellipsoid", 8, 0.8, 0.6)
    c = Case(pos, b, 1.5 + 0.05j, reduced_fft=reduced, surface=(1.2, 1.5 + 0.1j))
    cc, ccs = c.cc(PROP_Z, POL_Y)
    d = c.dmatrix()
    rng = np.random.default_rng(1)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    for her in (False, True):
        y1, n1 = c.matvec_fft(d, ccs, x, her)
        y2, n2 = c.matvec_direct(ccs, x, her)
        assert np.abs(y1 - y2).max() < 1e-12 * np.abs(y2).max()
        assert abs(n1 - n2) < 1e-12 * n2
    Ay, _ = c.matvec_fft(d, ccs, y)
    Ax, _ = c.matvec_fft(d, ccs, x)
    assert abs(np.sum(x * Ay) - np.sum(y * Ax)) < 1e-12 * abs(np.sum(x * Ay))
    # the reflected term is really there: without the substrate the result differs
    c0 = Case(pos, b, 1.5 + 0.05j, reduced_fft=reduced)
    y0, _ = c0.matvec_fft(c0.dmatrix(), ccs, x)
    assert np.abs(y0 - y1).max() > 1e-3 * np.abs(y1).max()


@pytest.mark.parametrize("name", list(SURF_RUNS))
def test_surface_vs_sparse_reference(name):
    """Converged Cext, Cabs (general ExtCross with inc_scale) and dipole polarizations against the reference's SPARSE
    build with -surf ... -int_surf img (tests/golden/surf_*/cmd.txt)."""
    c, (prop, px, py) = surf_case(name)
    d = c.dmatrix()
    for pol, inc, wx in (("Y", py, 0), ("X", px, 1)):
        cc, ccs = c.cc(prop, inc)
        e, sc = c.einc_surf(prop, px, py, wx)
        r = c.solve(d, ccs, e, method=0, eps=1e-10)
        assert r["rc"] == 0
        g = G.crosssec(name, pol)
        assert G.numeric_close([c.cext_gen(r["p"], e) * sc, c.cabs(r["p"], cc) * sc], [g["Cext"], g["Cabs"]],
                               1e-14, 1e-8)[0]
        if os.path.exists(G.path(name, "DipPol-" + pol)):
            gp = G.dippol(name, pol)
            assert G.numeric_close(np.c_[r["p"].real, r["p"].imag], np.c_[gp.real, gp.imag], 1e-12, 1e-6)[0]
