"""GPU parity of the particle-above-substrate path (-surf, image-dipole reflection): the R-matrix built on the device
(InitRmatrix, src/fft.c:819-891), the matvec with the reflected term fused into P3 (src/matvec.c:249,288-302), the
substrate's plane wave (src/GenerateB.c:258-322) and converged solves against the reference's SPARSE -surf runs.

Bars: R-matrix and FP64 matvec relative error <= 1e-12 against the oracle; FP32 matvec <= 2e-5; Cext, Cabs and DipPol
under comp2exec's rules (CrossSec abs 1e-14 / rel 1e-8, DipPol abs 1e-12 / rel 1e-6) against tests/golden/surf_*.
"""
import os
import subprocess

import numpy as np
import pytest

import golden_io as G
from cases import Case, make_shape, read_shape, PROP_Z, POL_Y
from test_oracle import SURF_RUNS, surf_case

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _op(adda, c, fp32=False):
    op = adda.Operator(c.pos, c.box, c.grid, c.mat, len(c.m), c.wavenum, c.gridspace, c.reduced_fft, fp32=fp32)
    hsub, msub = c.surface
    rcn = -1.0 if msub is None else (1 - msub * msub) / (1 + msub * msub)
    op.set_surface(c.zsum_shift, rcn)
    return op


SMALL = [
    ("ellipsoid", 8, 1.5 + 0.05j, 0.8, 0.6, (1.2, 1.5 + 0.1j)),
    ("sphere", 16, 1.5, 1.0, 1.0, (4.0, 2.0 + 0.5j)),            # grid 32^3
    ("box", 10, 1.2 + 0.05j, 0.7, 0.4, (1.0, None)),              # box 10x8x4 -> grid 20x16x8, perfect reflector
    ("ellipsoid", 24, 1.5 + 0.01j, 1.0, 0.5, (3.0, 1.33 + 0.0j)), # grid 48x48x24
]


def _small(i):
    shape, box, m, ay, az, surf = SMALL[i]
    pos, b = make_shape(shape, box, ay, az)
    c = Case(pos, b, m, surface=surf)
    assert surf[0] > (b[2] - 1) / 2 * c.gridspace  # every dipole above the substrate
    return c


@pytest.mark.parametrize("i", range(len(SMALL)))
def test_rmatrix_matches_oracle(hipdev, i):
    c = _small(i)
    op = _op(hipdev, c)
    try:
        r_gpu = op.rmatrix()
        r_or = c.rmatrix()
        assert r_gpu.shape == r_or.shape
        assert _rel(r_gpu, r_or) < 1e-12
    finally:
        op.close()


@pytest.mark.parametrize("i", range(len(SMALL)))
@pytest.mark.parametrize("her", [False, True])
def test_surface_matvec_fp64_vs_oracle(hipdev, i, her):
    """FP64 matvec with the reflected term against the oracle's FFT matvec (which transforms the reflected source with
    z backward, as the reference does)."""
    c = _small(i)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c)
    try:
        op.set_cc_sqrt(ccs)
        rng = np.random.default_rng(11 + i)
        x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
        y, ip = op.matvec(x, her=her)
        y_ref, ip_ref = c.matvec_fft(c.dmatrix(), ccs, x, her=her)
        assert _rel(y, y_ref) < 1e-12
        assert abs(ip - ip_ref) / ip_ref < 1e-12
    finally:
        op.close()


def test_surface_matvec_vs_direct_sum(hipdev):
    """The GPU matvec equals the direct sum over the dipoles and their images (ADDA's SPARSE MatVec with -surf)."""
    c = _small(0)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c)
    try:
        op.set_cc_sqrt(ccs)
        x = np.random.default_rng(4).standard_normal(3 * c.nd) + 0.3j
        y, _ = op.matvec(x)
        assert _rel(y, c.matvec_direct(ccs, x)[0]) < 1e-12
    finally:
        op.close()


@pytest.mark.parametrize("i", range(3))
def test_surface_matvec_fp32(hipdev, i):
    c = _small(i)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    try:
        op.set_cc_sqrt(ccs)
        x = np.random.default_rng(3).standard_normal(3 * c.nd) + 1j * np.random.default_rng(4).standard_normal(3 * c.nd)
        y, _ = op.matvec(x, prec=hipdev.F32)
        y_ref, _ = c.matvec_fft(c.dmatrix(), ccs, x)
        assert _rel(y, y_ref) < 2e-5
        op.set_matvec_mode(1)  # FP64 vectors around FP32 transforms
        y, _ = op.matvec(x)
        assert _rel(y, y_ref) < 2e-5
    finally:
        op.close()


@pytest.mark.parametrize("cfg", ["x=8.4 y=8.4 z=8.4 LY=1 TC=4 LP=1", "x=4.8 y=4.4.2 z=2.4.4 LY=8 TC=32 LP=2",
                                 "x=2.2.2.4 y=16.2 z=8.2.2 LY=2 TC=8 LP=4", "x=8.4 y=8.4 z=16.2 LY=3 TC=8 LP=1 XP=0",
                                 "x=16.2 y=8.4 z=8.4 LY=5 TC=16 LP=3 PT=1",
                                 "x=8.4 y=8.4 z=8.4 LY=4 TC=16 LP=1 XP=0 PT=1"])
def test_surface_configs_agree(hipdev, cfg):
    """Every P3 tile the substrate's kernel accepts (LP, column pairs on and off, thread counts) gives the same product
    (grid 32^3); the register and global-twiddle P3 variants, which lack the reflected term, are refused."""
    c = _small(1)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    try:
        op.set_cc_sqrt(ccs)
        x = np.random.default_rng(7).standard_normal(3 * c.nd) + 1j * np.random.default_rng(8).standard_normal(3 * c.nd)
        y_ref, _ = c.matvec_fft(c.dmatrix(), ccs, x)
        for prec in (hipdev.F64, hipdev.F32):
            op.set_config(cfg, prec)
        y1, _ = op.matvec(x)
        y132, _ = op.matvec(x, prec=hipdev.F32)
        assert _rel(y1, y_ref) < 1e-12
        assert _rel(y132, y_ref) < 2e-5
        for bad in ("x=8.4 y=8.4 z=8.4 LY=4 TC=8 LP=1 R3=1", "x=8.4 y=8.4 z=8.4 LY=4 TC=8 LP=1 TG=1"):
            with pytest.raises(RuntimeError):
                op.set_config(bad, hipdev.F64)
    finally:
        op.close()


def test_surface_properties_large(hipdev):
    """At a size the oracle does not reach (sphere box 64, grid 128^3, 137k dipoles): A stays complex symmetric
    (x.Ay = y.Ax, R_ji = R_ij^T); the product is linear in the reflection coefficient (A(2r) - A(r) = A(r) - A(0));
    rcn = 0 gives the free-space matvec."""
    pos, b = make_shape("sphere", 64)
    c = Case(pos, b, 1.5 + 0.02j, dpl=20, surface=(12.0, 1.5 + 0.1j))
    cc, ccs = c.cc(PROP_Z, POL_Y)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    free = hipdev.Operator(c.pos, c.box, c.grid, c.mat, 1, c.wavenum, c.gridspace)
    op = _op(hipdev, c)
    try:
        free.set_cc_sqrt(ccs)
        op.set_cc_sqrt(ccs)
        Ax, Ay = op.matvec(x)[0], op.matvec(y)[0]
        assert abs(np.sum(x * Ay) - np.sum(y * Ax)) < 1e-12 * abs(np.sum(x * Ay))
        r = (1 - (1.5 + 0.1j) ** 2) / (1 + (1.5 + 0.1j) ** 2)
        op.set_surface(c.zsum_shift, 0.0)
        A0 = op.matvec(x)[0]
        assert _rel(A0, free.matvec(x)[0]) < 1e-14
        op.set_surface(c.zsum_shift, 2 * r)
        A2 = op.matvec(x)[0]
        assert np.abs((A2 - Ax) - (Ax - A0)).max() < 1e-12 * np.abs(Ax - A0).max()
    finally:
        op.close()
        free.close()


def _surf_problem(adda, name, fp32=False):
    c, (prop, px, py) = surf_case(name)
    pr = adda.Problem(c.pos, c.box, c.m, mat=c.mat, dpl=c.dpl, grid=c.grid, surface=c.surface, fp32=fp32)
    return pr, prop, px, py


@pytest.mark.parametrize("name", list(SURF_RUNS))
def test_surface_solve_vs_sparse_reference(hipdev, name):
    """Converged QMR solves (FP64) near the substrate: Cext, Cabs and DipPol against the reference's SPARSE -surf runs
    (from above and below, a perfect reflector, oblique multi-material, evanescent transmission)."""
    pr, prop, px, py = _surf_problem(hipdev, name)
    try:
        for pol, inc in (("Y", py), ("X", px)):
            pr.set_polarization(prop, inc)
            st = pr.solve("qmr", eps=1e-10)
            assert st["rc"] == 0
            g = G.crosssec(name, pol)
            assert G.numeric_close([pr.cext(), pr.cabs()], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0], pol
            if os.path.exists(G.path(name, "DipPol-" + pol)):
                gp = G.dippol(name, pol)
                P = pr.polarization()
                assert G.numeric_close(np.c_[P.real, P.imag], np.c_[gp.real, gp.imag], 1e-12, 1e-6)[0], pol
    finally:
        pr.close()


@pytest.mark.parametrize("name", ["surf_img_above", "surf_img_tir"])
def test_surface_mixed_precision(hipdev, name):
    """FP32 inner iterations (FP32 D- and R-matrices) inside FP64 defect correction reach the reference's values."""
    pr, prop, px, py = _surf_problem(hipdev, name, fp32=True)
    try:
        pr.set_polarization(prop, py)
        st = pr.solve("bicgstab", eps=1e-10, precision=hipdev.PREC_MIXED)
        assert st["rc"] == 0 and st["true_resid"] <= 1e-10
        g = G.crosssec(name, "Y")
        assert G.numeric_close([pr.cext(), pr.cabs()], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0]
    finally:
        pr.close()


def test_surface_einc_vs_oracle(hipdev):
    """The substrate's incident field on the device equals the oracle's GenerateB restatement (from above with the
    reflected wave, from below with the evanescent transmitted wave), both polarizations."""
    for name in ("surf_img_oblique", "surf_img_tir", "surf_inf_above"):
        c, (prop, px, py) = surf_case(name)
        pr = hipdev.Problem(c.pos, c.box, c.m, mat=c.mat, dpl=c.dpl, grid=c.grid, surface=c.surface)
        try:
            for inc, wx in ((py, 0), (px, 1)):
                pr.set_polarization(prop, inc)
                e_ref, sc = c.einc_surf(prop, px, py, wx)
                e = pr.op.download(pr.einc)
                assert _rel(e, e_ref) < 1e-13, name
                assert pr.inc_scale == sc
        finally:
            pr.close()


def test_cli_surface(hipdev, tmp_path):
    """adda_hip -surf ... -int_surf img -prop ... reproduces the reference's CrossSec-Y/X and DipPol-Y (oblique
    incidence on the three-material box above an absorbing substrate)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "a-dda_amd", "adda_hip")
    run = tmp_path / "run"
    cmd = [exe, "-shape", "read", G.path("geom", "box8mm.geom"), "-dpl", "14", "-m", "1.2", "0.01", "1.5", "0.1",
           "1.8", "0", "-int_surf", "img", "-iter", "qmr", "-eps", "10", "-store_dip_pol", "-surf", "1.7", "2.0",
           "0.5", "-prop", "0.5", "0.3", "-1", "-dir", str(run)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    for pol in ("Y", "X"):
        a = G.crosssec("surf_img_oblique", pol)
        with open(run / ("CrossSec-" + pol)) as f:
            b = {l.split("=")[0].strip(): float(l.split("=")[1]) for l in f}
        assert G.numeric_close([b["Cext"], b["Cabs"]], [a["Cext"], a["Cabs"]], 1e-14, 1e-8)[0], pol
        ours = np.loadtxt(run / ("DipPol-" + pol), skiprows=1)
        ref = np.loadtxt(G.path("surf_img_oblique", "DipPol-" + pol), skiprows=1)
        assert G.numeric_close(ours, ref, 1e-12, 1e-6)[0], pol
    assert not os.path.exists(run / "mueller")
