"""Full-size parity of the bench workload against the oracle: sphere on a 128^3 box (1,099,136 dipoles), FFT grid
256^3, m = 1.5+0.01i, size parameter 40 (-size 80 with the volume correction, src/make_particle.c:2192-2196).

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
