"""Full-size parity against the oracle of the bench workload -- sphere on a 128^3 box (1,099,136 dipoles), FFT grid
256^3, m = 1.5+0.01i, size parameter 40 (-size 80 with the volume correction, src/make_particle.c:2192-2196) -- and
of the 64-RHS batched workload.

The oracle's FP64 restatement builds its own D-matrix (InitDmatrix, src/fft.c:918-1416) and runs its own FFT matvec
(src/matvec.c:128-425) on one seeded argument, in about a minute of host time. Against it:
  - the GPU's D-matrix, expanded to ADDA's layout, to 1e-12 relative (max-norm);
  - the FP64 matvec and its ||y||^2 to 1e-12;
  - the FP32 matvec and the FP32-transform matvec with FP64 vectors to 2e-5 (the FP32 bar of the small cases).
"""
import numpy as np
import pytest

from cases import Case, make_shape, PROP_Z, POL_Y

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_grid256_dmatrix_and_matvec_vs_oracle(hipdev):
    pos, box = make_shape("sphere", 128)
    nd = len(pos)
    assert nd == 1099136
    dpl = 2 * np.pi * (nd / (np.pi / 6)) ** (1 / 3) / 80.0
    c = Case(pos, box, 1.5 + 0.01j, dpl=dpl)
    assert c.grid == (256, 256, 256)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    rng = np.random.default_rng(256)
    x = rng.standard_normal(3 * nd) + 1j * rng.standard_normal(3 * nd)
    op = hipdev.Operator(c.pos, c.box, c.grid, c.mat, len(c.m), c.wavenum, c.gridspace, c.reduced_fft, fp32=True)
    try:
        op.set_cc_sqrt(ccs)
        y64, ip64 = op.matvec(x)
        y32, _ = op.matvec(x, prec=hipdev.F32)
        op.set_matvec_mode(1)
        yv, _ = op.matvec(x)
        op.set_matvec_mode(0)
        d_gpu = op.dmatrix()
    finally:
        op.close()
    d_or = c.dmatrix()
    assert d_gpu.shape == d_or.shape
    assert _rel(d_gpu, d_or) < 1e-12
    del d_gpu
    y_or, ip_or = c.matvec_fft(d_or, ccs, x)
    assert _rel(y64, y_or) < 1e-12
    assert abs(ip64 - ip_or) / ip_or < 1e-12
    assert _rel(y32, y_or) < 2e-5
    assert _rel(yv, y_or) < 2e-5


def test_batched_workload_matvec_vs_oracle(hipdev):
    """The batched workload at full size: ellipsoid box 64x52x32 (semi-axes 1:0.8:0.5, 54,928 dipoles) on the fast
    grid 128x128x64, m = 1.33+0.001i, default dpl 10|m|, 32 seeded orientations x 2 polarizations = 64 RHS, each with
    its own coupling constants (InitCC per orientation, src/calculator.c:266-289). One batched FP64 matvec of all 64
    RHS; three of them against the oracle's FP64 matvec (1e-12) with the oracle's coupling constants, and the batched
    FP32 matvec of the same RHS within 2e-5."""
    from adda_rot import rotate_zyz
    from cases import orientations, POL_X
    pos, box = make_shape("ellipsoid", 64, 0.8, 0.5)
    m = 1.33 + 0.001j
    props, pols = [], []
    for a, b, g in orientations(32):
        prop, pY, pX = (rotate_zyz(a, b, g, v) for v in (PROP_Z, POL_Y, POL_X))
        props += [prop, prop]
        pols += [pY, pX]
    grid = hipdev.fast_grid(box)
    assert len(pos) == 54928 and grid == (128, 128, 64)
    K = 64
    bp = hipdev.BatchProblem(pos, box, m, np.array(props), np.array(pols), fp32=True, grid=grid)
    c = Case(pos, box, m, dpl=bp.dpl, grid=grid)
    assert abs(c.gridspace - bp.gridspace) <= 1e-15 * c.gridspace
    rng = np.random.default_rng(64)
    xs = [rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd) for _ in range(K)]
    check = (0, 21, 63)
    op = bp.op
    try:
        yb = {}
        for prec in (hipdev.F64, hipdev.F32):
            a, r = op.alloc_batch(K, prec), op.alloc_batch(K, prec)
            for k in range(K):
                op.upload(op.rhs(a, k, prec), xs[k], prec)
            op.matvec_batch(a, r, K, prec)
            yb[prec] = {k: op.download(op.rhs(r, k, prec), prec) for k in check}
            op.free(a)
            op.free(r)
    finally:
        bp.close()
    d = c.dmatrix()
    for k in check:
        cc, ccs = c.cc(props[k], pols[k])
        assert np.allclose(bp.cc_sqrt[k], np.reshape(ccs, bp.cc_sqrt[k].shape), rtol=1e-13, atol=0)
        y_or, _ = c.matvec_fft(d, ccs, xs[k])
        assert _rel(yb[hipdev.F64][k], y_or) < 1e-12, k
        assert _rel(yb[hipdev.F32][k], y_or) < 2e-5, k


def test_grid256_converged_solves(hipdev):
    """The bench workload solved to ||b - Ax||/||b|| <= 1e-10 (true FP64 residual) by the plain FP64 BiCGStab and
    QMR_CS and by the FP32-mixed forms (FP32 inner iterations, FP64 defect correction): BiCGStab, and QMR_CS with its
    automatic inner method. Every mixed solve ends below 1e-10 on the true residual, and its Cext is within 1e-6 of
    the FP64 BiCGStab solve's (the 1e-10 solves agree far closer, about 1e-10)."""
    pos, mat, box = hipdev.make_particle("sphere", 128)
    nd = len(pos)
    dpl = 2 * np.pi * (nd / (np.pi / 6)) ** (1 / 3) / 80.0
    pr = hipdev.Problem(pos, box, 1.5 + 0.01j, dpl=dpl, fp32=True)
    try:
        pr.set_polarization(PROP_Z, POL_Y)
        r64 = pr.solve("bicgstab", eps=1e-10, precision=hipdev.PREC_FP64, recalc_resid=True)
        c64 = pr.cext()
        assert r64["rc"] == 0 and r64["true_resid"] <= 1e-10 * 1.05
        q64 = pr.solve("qmr", eps=1e-10, precision=hipdev.PREC_FP64, recalc_resid=True)
        assert q64["rc"] == 0 and abs(pr.cext() - c64) / c64 < 1e-8
        for method in ("bicgstab", "qmr"):
            r = pr.solve(method, eps=1e-10, precision=hipdev.PREC_MIXED)
            assert r["rc"] == 0 and r["true_resid"] <= 1e-10, (method, r)
            assert abs(pr.cext() - c64) / c64 < 1e-6, method
    finally:
        pr.close()
