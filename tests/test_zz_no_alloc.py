"""adda_hip_dot_noalloc, the dot product that allocates nothing (partial sums in a static device array, results in a
device buffer of the caller), against adda_hip_dot_batch through the C ABI: bit for bit, every mode, FP64 and FP32
vectors, two calls in a row -- the second on a smaller operator, whose fewer blocks must not pick up the first call's
partial sums left in the array -- and the size limit, which fails before anything is launched."""
import ctypes

import numpy as np
import pytest

from cases import Case, make_shape

pytestmark = pytest.mark.gpu


def _op(A, box, K):
    pos, b = make_shape("sphere", box)
    c = Case(pos, b, 1.5 + 0.01j)
    op = A.Operator(c.pos, c.box, c.grid, c.mat, len(c.m), c.wavenum, c.gridspace, c.reduced_fft)
    op.reserve(K)
    return op


def _check(A, op, K, seed):
    L = A.hip()
    rng = np.random.default_rng(seed)
    out = op.alloc()  # a device buffer of 3*nd complex, far more than the 3K doubles written
    try:
        for prec in (A.F64, A.F32):
            va, vb = op.alloc_batch(K, prec), op.alloc_batch(K, prec)
            for k in range(K):
                n = 3 * op.nd
                op.upload(op.rhs(va, k, prec), rng.standard_normal(n) + 1j * rng.standard_normal(n), prec)
                op.upload(op.rhs(vb, k, prec), rng.standard_normal(n) - 1j * rng.standard_normal(n), prec)
            for mode in (A.DOT_CONJ, A.DOT_BILIN, A.DOT_SELF_BILIN, A.DOT_NORM2):
                op.upload(out, np.full(3 * op.nd, np.nan + 1j * np.nan))  # nothing stale may survive
                b = vb if mode in (A.DOT_CONJ, A.DOT_BILIN) else None
                A.check(L.adda_hip_dot_noalloc(op.h, prec, K, va, b, mode, out))
                got = op.download(out).view(np.float64)[:3 * K].reshape(K, 3)
                res = (A.Cplx * K)()
                n2 = (ctypes.c_double * K)()
                A.check(L.adda_hip_dot_batch(op.h, prec, K, None, va, b, mode, res, n2))
                for k in range(K):
                    assert got[k, 0] == res[k].re and got[k, 1] == res[k].im, (prec, mode, k)
                    if mode in (A.DOT_SELF_BILIN, A.DOT_NORM2):
                        assert got[k, 2] == n2[k], (prec, mode, k)
                    else:
                        assert got[k, 2] == 0.0
            op.free(va)
            op.free(vb)
    finally:
        op.free(out)


def test_dot_noalloc_matches_dot_batch(hipdev):
    A = hipdev
    big = _op(A, 24, 3)  # 3*nd = 21 k elements: 83 blocks of partial sums
    try:
        _check(A, big, 3, 1)
    finally:
        big.close()
    small = _op(A, 8, 1)  # 4 blocks: the slots the first call filled beyond them stay unread
    try:
        _check(A, small, 1, 2)
    finally:
        small.close()


def test_dot_noalloc_size_limit(hipdev):
    A = hipdev
    op = _op(A, 80, 17)  # 2048 blocks x 3 rows x 17 RHS > the 3*2048*16 doubles of the array
    try:
        va = op.alloc_batch(17)
        out = op.alloc()
        rc = A.hip().adda_hip_dot_noalloc(op.h, A.F64, 17, va, None, A.DOT_NORM2, out)
        assert rc != 0 and b"scratch" in A.hip().adda_hip_last_error()
        A.check(A.hip().adda_hip_dot_noalloc(op.h, A.F64, 16, va, None, A.DOT_NORM2, out))  # 16 fit
        op.free(va)
        op.free(out)
    finally:
        op.close()
