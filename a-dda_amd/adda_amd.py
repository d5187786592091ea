"""Python view of the MI355X ADDA backend: ctypes bindings of include/adda_hip.h (libadda_hip.so) and
include/adda.h (libadda.so, the C host side: particle builder, polarizability, Krylov solvers, cross sections).

The host-language product is C; this module only binds it for the test harness, bench.py and the torch adjoint shim.
Every call fails loudly (RuntimeError with the library's "file:line: message") -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F64, F32 = 0, 1
DOT_CONJ, DOT_BILIN, DOT_SELF_BILIN, DOT_NORM2 = 0, 1, 2, 3
CONJ_A, CONJ_B = 1, 2


class Cplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @staticmethod
    def of(z):
        z = complex(z)
        return Cplx(z.real, z.imag)

    def value(self):
        return complex(self.re, self.im)


_hip = None


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def hip():
    """Load libadda_hip.so (built in-tree by __graft_entry__.build()); raise if it is missing."""
    global _hip
    if _hip is not None:
        return _hip
    path = os.path.join(HERE, "libadda_hip.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build the HIP extension first (__graft_entry__.build())")
    L = ctypes.CDLL(path)
    P, I, D, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    L.adda_hip_last_error.restype = ctypes.c_char_p
    L.adda_hip_device_count.restype = I
    L.adda_hip_fftfit.argtypes = [I, I]
    L.adda_hip_fftfit.restype = I
    L.adda_hip_op_create.argtypes = [I, I, I, I, I, I, I, S, P, P, I, D, D, I, I]
    L.adda_hip_op_create.restype = P
    L.adda_hip_op_destroy.argtypes = [P]
    L.adda_hip_op_destroy.restype = None
    L.adda_hip_set_cc_sqrt.argtypes = [P, P]
    L.adda_hip_get_dmatrix.argtypes = [P, P]
    L.adda_hip_dmatrix_size.argtypes = [P]
    L.adda_hip_dmatrix_size.restype = S
    L.adda_hip_vec_alloc.argtypes = [P, I]
    L.adda_hip_vec_alloc.restype = P
    L.adda_hip_vec_free.argtypes = [P, P]
    L.adda_hip_vec_upload.argtypes = [P, I, P, P]
    L.adda_hip_vec_download.argtypes = [P, I, P, P]
    L.adda_hip_vec_convert.argtypes = [P, P, I, P, I]
    L.adda_hip_nrows.argtypes = [P]
    L.adda_hip_nrows.restype = S
    L.adda_hip_matvec.argtypes = [P, I, P, P, P, I]
    L.adda_hip_matvec_host.argtypes = [P, P, P, P, I]
    L.adda_hip_lincomb.argtypes = [P, I, P, P, Cplx, P, Cplx, P, Cplx, I, P]
    L.adda_hip_dot.argtypes = [P, I, P, P, I, P, P]
    L.adda_hip_mult_mat.argtypes = [P, I, P, P]
    L.adda_hip_sync.argtypes = [P]
    _hip = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError(hip().adda_hip_last_error().decode())


def device_count():
    return hip().adda_hip_device_count()


class Operator:
    """One particle's FFT interaction operator on one GPU (D-matrix built in FP64 on the device)."""

    def __init__(self, pos, box, grid, mat, nmat, wavenum, gridspace, reduced_fft=True, fp32=False, device=0):
        L = hip()
        self.pos = np.ascontiguousarray(pos, np.uint16)
        self.mat = np.ascontiguousarray(mat, np.int32)
        self.nd = len(self.pos)
        self.box, self.grid, self.nmat = tuple(box), tuple(grid), int(nmat)
        h = L.adda_hip_op_create(device, *self.box, *self.grid, self.nd, _p(self.pos), _p(self.mat), self.nmat,
                                 wavenum, gridspace, int(reduced_fft), 1 if fp32 else 0)
        if not h:
            raise RuntimeError(L.adda_hip_last_error().decode())
        self.h = h
        self._vecs = []

    def close(self):
        if getattr(self, "h", None):
            hip().adda_hip_op_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_cc_sqrt(self, cc_sqrt):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(cc_sqrt, np.complex128).reshape(-1, 3),
                                                 (self.nmat, 3)))
        check(hip().adda_hip_set_cc_sqrt(self.h, _p(a)))

    def dmatrix(self):
        L = hip()
        d = np.empty(L.adda_hip_dmatrix_size(self.h), np.complex128)
        check(L.adda_hip_get_dmatrix(self.h, _p(d)))
        return d

    # device vectors
    def alloc(self, prec=F64):
        v = hip().adda_hip_vec_alloc(self.h, prec)
        if not v:
            raise RuntimeError(hip().adda_hip_last_error().decode())
        return v

    def free(self, v):
        check(hip().adda_hip_vec_free(self.h, v))

    def upload(self, v, x, prec=F64):
        x = np.ascontiguousarray(x, np.complex128)
        check(hip().adda_hip_vec_upload(self.h, prec, v, _p(x)))

    def download(self, v, prec=F64):
        out = np.empty(3 * self.nd, np.complex128)
        check(hip().adda_hip_vec_download(self.h, prec, _p(out), v))
        return out

    def matvec_dev(self, a, r, prec=F64, her=False, inprod=False):
        ip = ctypes.c_double()
        check(hip().adda_hip_matvec(self.h, prec, a, r, ctypes.byref(ip) if inprod else None, int(her)))
        return ip.value if inprod else None

    def matvec(self, x, her=False, prec=F64):
        """Host convenience: y = A x (or A^H x) with the requested device precision; returns (y, ||y||^2)."""
        a, r = self.alloc(prec), self.alloc(prec)
        try:
            self.upload(a, x, prec)
            ip = self.matvec_dev(a, r, prec, her, inprod=True)
            return self.download(r, prec), ip
        finally:
            self.free(a)
            self.free(r)

    def lincomb(self, out, a=None, ca=0, b=None, cb=0, c=None, cc=0, flags=0, prec=F64, norm=False):
        n2 = ctypes.c_double()
        check(hip().adda_hip_lincomb(self.h, prec, out, a, Cplx.of(ca), b, Cplx.of(cb), c, Cplx.of(cc), flags,
                                     ctypes.byref(n2) if norm else None))
        return n2.value if norm else None

    def dot(self, a, b=None, mode=DOT_CONJ, prec=F64):
        r, n2 = Cplx(), ctypes.c_double()
        check(hip().adda_hip_dot(self.h, prec, a, b, mode, ctypes.byref(r), ctypes.byref(n2)))
        return r.value(), n2.value

    def mult_mat(self, out, inp, prec=F64):
        check(hip().adda_hip_mult_mat(self.h, prec, out, inp))

    def sync(self):
        check(hip().adda_hip_sync(self.h))
