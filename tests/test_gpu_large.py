"""Full-size checks of the large single-particle configuration: sphere on a 320^3 box (17,158,672 dipoles, padded FFT
grid 640^3 = 2^7*5 per axis, radix-5 plans), m = 1.33+0.001i, size parameter 100.

The oracle cannot run at this size (one FP64 matvec would take minutes on the host), so these are size-independent
properties of the operator A = I + S.D.S, each of which a wrong FFT index, mirror sign or tile boundary would break:
  - complex symmetry: y^T (A x) == x^T (A y) to 1e-12 relative (QMR_CS relies on it);
  - her: A^H x == conj(A conj(x)) to 1e-12 (A symmetric, so A^H = conj(A));
  - the compile-time plans the operator picks for length 640 agree with a run-time plan of another radix schedule to
    1e-12 (two independent FFT paths);
  - linearity of the batched form: A(x + 2i y) == A x + 2i A y to 1e-12;
  - FP32 transforms (FP32 vectors, and FP64 vectors around FP32 transforms) within 2e-5 of FP64.
The plans themselves are checked against the oracle at length 640 in test_gpu_matvec.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def large(hipdev):
    pos, _, box = hipdev.make_particle("sphere", 320)
    nd = len(pos)
    lam = 2 * np.pi
    dpl = lam * (nd / (np.pi / 6)) ** (1 / 3) / 200.0  # -size 200: size parameter 100
    pr = hipdev.Problem(pos, box, 1.33 + 0.001j, dpl=dpl, fp32=True)
    assert pr.grid == (640, 640, 640) and nd == 17158672
    pr.set_polarization(np.array([0.0, 0, 1]), np.array([0.0, 1, 0]))
    yield pr
    pr.close()


def test_large_symmetry_her_and_plans(hipdev, large):
    op, n = large.op, 3 * large.nd
    rng = np.random.default_rng(640)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a, b, ra, rb = op.alloc(), op.alloc(), op.alloc(), op.alloc()
    try:
        op.upload(a, x)
        op.upload(b, y)
        op.matvec_dev(a, ra)
        op.matvec_dev(b, rb)
        s1, _ = op.dot(b, ra, hipdev.DOT_BILIN)
        s2, _ = op.dot(a, rb, hipdev.DOT_BILIN)
        assert abs(s1 - s2) / abs(s1) < 1e-12, (s1, s2)
        ax = op.download(ra)
        yh, _ = op.matvec(x, her=True)
        yc, _ = op.matvec(x.conj())
        assert _rel(yh, yc.conj()) < 1e-12
        tuned = op.get_config(hipdev.F64)
        op.set_config("x=8.2.8.5 y=2.4.16.5 z=4.2.16.5 LY=2 TC=4 LP=1 XP=0", hipdev.F64)  # run-time plans
        y_dyn, _ = op.matvec(x)
        op.set_config(tuned, hipdev.F64)
        assert _rel(y_dyn, ax) < 1e-12
        for mode, prec in ((0, hipdev.F32), (1, hipdev.F64)):
            op.set_matvec_mode(mode)
            y32, _ = op.matvec(x, prec=prec)
            op.set_matvec_mode(0)
            assert _rel(y32, ax) < 2e-5, (mode, prec)
    finally:
        for v in (a, b, ra, rb):
            op.free(v)


def test_large_batched_linearity(hipdev, large):
    op, n, K = large.op, 3 * large.nd, 3
    rng = np.random.default_rng(641)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) - 1j * rng.standard_normal(n)
    op.reserve(K)
    op.set_cc_sqrt_batch([large.cc_sqrt] * K)
    a, r = op.alloc_batch(K), op.alloc_batch(K)
    try:
        for k, v in enumerate((x, y, x + 2j * y)):
            op.upload(op.rhs(a, k), v)
        op.matvec_batch(a, r, K)
        ax, ay, az = (op.download(op.rhs(r, k)) for k in range(K))
        assert _rel(az, ax + 2j * ay) < 1e-12
    finally:
        op.free(a)
        op.free(r)
