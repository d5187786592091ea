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
    L.adda_hip_fastgrid.argtypes = [I, I]
    L.adda_hip_fastgrid.restype = I
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
    L.adda_hip_plane_wave.argtypes = [P, D, D, P, P, P]
    L.adda_hip_field_sums.argtypes = [P, I, P, D, P, I, P]
    L.adda_hip_weighted_norm2.argtypes = [P, I, P, P, P]
    L.adda_hip_profile_matvec.argtypes = [P, I, P, P, I, P, P]
    L.adda_hip_matvec_bytes.argtypes = [P, I, I]
    L.adda_hip_matvec_bytes.restype = D
    L.adda_hip_get_config.argtypes = [P, I, ctypes.c_char_p, S]
    L.adda_hip_set_config.argtypes = [P, I, ctypes.c_char_p]
    L.adda_hip_set_matvec_mode.argtypes = [P, I]
    L.adda_hip_set_launch_list.argtypes = [P, I, P]
    L.adda_hip_autotune.argtypes = [P, I, I, ctypes.c_char_p, ctypes.c_char_p, S]
    L.adda_hip_dk_create.argtypes = [P, I]
    L.adda_hip_dk_create.restype = P
    L.adda_hip_dk_free.argtypes = [P, P]
    L.adda_hip_dk_set.argtypes = [P, P, I, P, I]
    L.adda_hip_dk_get.argtypes = [P, P, I, P, I]
    L.adda_hip_dk_history.argtypes = [P, P, P, I]
    L.adda_hip_dk_qmr_iter.argtypes = [P, I, P, I, I, P, P, P, P, P, P, P]
    L.adda_hip_dk_bicgstab_iter.argtypes = [P, I, P, I, I, P, P, P, P, P, P, P]
    L.adda_hip_dk_cgnr_iter.argtypes = [P, I, P, I, P, P, P, P]
    # batched multi-RHS
    L.adda_hip_reserve_batch.argtypes = [P, I]
    L.adda_hip_batch_stride.argtypes = [P]
    L.adda_hip_batch_stride.restype = S
    L.adda_hip_set_cc_sqrt_batch.argtypes = [P, I, P]
    L.adda_hip_vec_alloc_batch.argtypes = [P, I, I]
    L.adda_hip_vec_alloc_batch.restype = P
    L.adda_hip_vec_convert_batch.argtypes = [P, I, P, I, P, I]
    L.adda_hip_matvec_batch.argtypes = [P, I, I, P, P, P, P, I]
    L.adda_hip_lincomb_batch.argtypes = [P, I, I, P, P, P, P, P, P, P, P, I, P]
    L.adda_hip_dot_batch.argtypes = [P, I, I, P, P, P, I, P, P]
    L.adda_hip_dot_noalloc.argtypes = [P, I, I, P, P, I, P]
    L.adda_hip_mult_mat_batch.argtypes = [P, I, I, P, P]
    L.adda_hip_profile_matvec_batch.argtypes = [P, I, I, P, P, I, P, P]
    L.adda_hip_matvec_bytes_batch.argtypes = [P, I, I, I]
    L.adda_hip_matvec_bytes_batch.restype = D
    L.adda_hip_dk_create_batch.argtypes = [P, I, I]
    L.adda_hip_dk_create_batch.restype = P
    L.adda_hip_autotune_batch.argtypes = [P, I, I, I, ctypes.c_char_p, ctypes.c_char_p, S]
    L.adda_hip_capture_begin.argtypes = [P]
    L.adda_hip_capture_end.argtypes = [P]
    L.adda_hip_capture_end.restype = P
    L.adda_hip_graph_launch.argtypes = [P, P]
    L.adda_hip_graph_free.argtypes = [P, P]
    L.adda_hip_weighted_conj.argtypes = [P, P, P, P]
    L.adda_hip_adjoint_grad.argtypes = [P, P, P, P, P, D, D, P, P]
    L.adda_hip_vec_download_part.argtypes = [P, I, P, P, S, S]
    L.adda_hip_set_surface.argtypes = [P, D, Cplx]
    L.adda_hip_get_rmatrix.argtypes = [P, P]
    L.adda_hip_rmatrix_size.argtypes = [P]
    L.adda_hip_rmatrix_size.restype = S
    L.adda_hip_plane_waves.argtypes = [P, D, D, I, P, P, P]
    L.adda_hip_hbm_rates.argtypes = [I, S, I, P]
    _hip = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError(hip().adda_hip_last_error().decode())


def device_count():
    return hip().adda_hip_device_count()


def hbm_rates(device=0, nbytes=1 << 30, nrep=20):
    """GB/s of a plain device copy (read + write bytes), a read-only pass, a fill and a nontemporal copy of nbytes (one
    16-byte element per thread, HIP events)."""
    out = np.zeros(4)
    check(hip().adda_hip_hbm_rates(device, nbytes, nrep, _p(out)))
    return {"bytes": nbytes, "copy_GBps": round(out[0], 1), "read_GBps": round(out[1], 1), "fill_GBps": round(out[2], 1),
            "copy_nt_GBps": round(out[3], 1)}


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

    def set_surface(self, zsum_shift, rcn):
        """Image-dipole substrate term (adda_hip_set_surface): every later matvec adds it."""
        check(hip().adda_hip_set_surface(self.h, zsum_shift, Cplx.of(rcn)))

    def rmatrix(self):
        L = hip()
        r = np.empty(L.adda_hip_rmatrix_size(self.h), np.complex128)
        check(L.adda_hip_get_rmatrix(self.h, _p(r)))
        return r

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

    def matvec_host(self, x, her=False):
        """The drop-in MatVec on host vectors (adda_hip_matvec_host: copies in and out over PCIe each call); returns
        (y, ||y||^2)."""
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty_like(x)
        ip = ctypes.c_double()
        check(hip().adda_hip_matvec_host(self.h, _p(x), _p(y), ctypes.byref(ip), int(her)))
        return y, ip.value

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

    def profile_matvec(self, a, r, prec=F64, nrep=10):
        """Average per-pass and total matvec time (ms), HIP events on the operator's stream."""
        ms = np.zeros(5)
        tot = ctypes.c_double()
        check(hip().adda_hip_profile_matvec(self.h, prec, a, r, nrep, _p(ms), ctypes.byref(tot)))
        return ms, tot.value

    def get_config(self, prec=F64):
        b = ctypes.create_string_buffer(256)
        check(hip().adda_hip_get_config(self.h, prec, b, 256))
        return b.value.decode()

    def set_config(self, cfg, prec=F64):
        check(hip().adda_hip_set_config(self.h, prec, cfg.encode()))

    def set_matvec_mode(self, mode):
        """1: FP64 matvecs run their transforms in FP32 (FP64 vectors in and out); 0: in FP64."""
        check(hip().adda_hip_set_matvec_mode(self.h, mode))

    def set_launch_list(self, K, run=None):
        """Batched matvecs of K RHS launch only the RHS k with run[k] true (None clears the list)."""
        if run is None:
            check(hip().adda_hip_set_launch_list(self.h, K, None))
            return
        r = np.ascontiguousarray(run, dtype=np.int32)
        check(hip().adda_hip_set_launch_list(self.h, K, r.ctypes.data_as(ctypes.c_void_p)))

    def autotune(self, prec=F64, reps=3, cache=None):
        """Pick the fastest radix schedules and tiles per pass (HIP-event timing); returns the report string."""
        b = ctypes.create_string_buffer(512)
        check(hip().adda_hip_autotune(self.h, prec, reps, cache.encode() if cache else None, b, 512))
        return b.value.decode()

    def matvec_bytes(self, prec=F64, pass_=5):
        return hip().adda_hip_matvec_bytes(self.h, prec, pass_)

    # batched vectors: K right-hand sides at stride `stride` complex elements
    def reserve(self, K):
        check(hip().adda_hip_reserve_batch(self.h, K))

    @property
    def stride(self):
        return hip().adda_hip_batch_stride(self.h)

    def alloc_batch(self, K, prec=F64):
        v = hip().adda_hip_vec_alloc_batch(self.h, prec, K)
        if not v:
            raise RuntimeError(hip().adda_hip_last_error().decode())
        return v

    def rhs(self, v, k, prec=F64):
        """Device address of RHS k of a batched vector."""
        return v + k * self.stride * (16 if prec == F64 else 8)

    def set_cc_sqrt_batch(self, cc_sqrt):
        a = np.ascontiguousarray(np.asarray(cc_sqrt, np.complex128).reshape(-1, self.nmat, 3))
        check(hip().adda_hip_set_cc_sqrt_batch(self.h, a.shape[0], _p(a)))

    def matvec_batch(self, a, r, K, prec=F64, her=False, inprod=False, gate=None):
        ip = np.zeros(K)
        check(hip().adda_hip_matvec_batch(self.h, prec, K, gate, a, r, _p(ip) if inprod else None, int(her)))
        return ip if inprod else None

    def lincomb_batch(self, K, out, a=None, ca=0, b=None, cb=0, c=None, cc=0, flags=0, prec=F64, norm=False, gate=None):
        co = [np.ascontiguousarray(np.broadcast_to(np.asarray(z, np.complex128), (K,))) for z in (ca, cb, cc)]
        n2 = np.zeros(K)
        check(hip().adda_hip_lincomb_batch(self.h, prec, K, gate, out, a, _p(co[0]), b, _p(co[1]), c, _p(co[2]),
                                           flags, _p(n2) if norm else None))
        return n2 if norm else None

    def dot_batch(self, K, a, b=None, mode=DOT_CONJ, prec=F64, gate=None):
        r, n2 = np.zeros(K, np.complex128), np.zeros(K)
        check(hip().adda_hip_dot_batch(self.h, prec, K, gate, a, b, mode, _p(r), _p(n2)))
        return r, n2

    def profile_matvec_batch(self, a, r, K, prec=F64, nrep=10):
        ms = np.zeros(5)
        tot = ctypes.c_double()
        check(hip().adda_hip_profile_matvec_batch(self.h, prec, K, a, r, nrep, _p(ms), ctypes.byref(tot)))
        return ms, tot.value

    def matvec_bytes_batch(self, K, prec=F64, pass_=5):
        return hip().adda_hip_matvec_bytes_batch(self.h, prec, K, pass_)

    def autotune_batch(self, K, prec=F64, reps=3, cache=None):
        b = ctypes.create_string_buffer(512)
        check(hip().adda_hip_autotune_batch(self.h, prec, K, reps, cache.encode() if cache else None, b, 512))
        return b.value.decode()


# ---------------------------------------------------------------------------------------------------------------------
# C host layer (include/adda.h)

IT_QMR, IT_BICGSTAB, IT_CGNR, IT_BICG, IT_CSYM, IT_BCGS2, IT_QMR2 = 0, 1, 2, 3, 4, 5, 6
CHP_NONE, CHP_NORMAL, CHP_REGULAR, CHP_ALWAYS = 0, 1, 2, 3
ITER_NAMES = {"qmr": IT_QMR, "bicgstab": IT_BICGSTAB, "cgnr": IT_CGNR, "bicg": IT_BICG, "csym": IT_CSYM,
              "bcgs2": IT_BCGS2, "qmr2": IT_QMR2}
IF_AUTO, IF_ZERO, IF_INC = 0, 1, 2
POL_LDR, POL_CM, POL_RRC = 0, 1, 2
PREC_FP64, PREC_MIXED, PREC_MIXED_V64 = 0, 1, 2


class Particle(ctypes.Structure):
    _fields_ = [("boxX", ctypes.c_int), ("boxY", ctypes.c_int), ("boxZ", ctypes.c_int), ("nd", ctypes.c_size_t),
                ("position", ctypes.POINTER(ctypes.c_ushort)), ("material", ctypes.POINTER(ctypes.c_int)),
                ("nmat", ctypes.c_int)]


class SolverOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("eps", ctypes.c_double), ("maxiter", ctypes.c_int),
                ("init_field", ctypes.c_int), ("recalc_resid", ctypes.c_int), ("precision", ctypes.c_int),
                ("inner_eps", ctypes.c_double), ("max_outer", ctypes.c_int), ("verbose", ctypes.c_int),
                ("host_scalars", ctypes.c_int), ("max_inner", ctypes.c_int), ("graph", ctypes.c_int),
                ("chp_type", ctypes.c_int), ("chp_time", ctypes.c_double), ("chp_dir", ctypes.c_char_p),
                ("chp_load", ctypes.c_int), ("run_dir", ctypes.c_char_p), ("inner_method", ctypes.c_int)]


def _inner(method):
    """SolverOpts.inner_method of a method name (None: automatic choice)."""
    return 0 if method is None else ITER_NAMES.get(method, method) + 1


class ChpHead(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int), ("counter", ctypes.c_int), ("inprodR", ctypes.c_double),
                ("prev_err", ctypes.c_double), ("resid_scale", ctypes.c_double)]


def chp_layout(method):
    """(index in the reference's solver table, own scalars in doubles, own vectors) of an iterative method."""
    v = [ctypes.c_int() for _ in range(3)]
    hcheck(host().adda_chp_layout(ITER_NAMES.get(method, method), *[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def chp_read(chp_dir, method, nrows):
    """Reads <chp_dir>/chp.0 (ADDA's iteration checkpoint): head dict, scalars (doubles), vectors [3+n][nrows]."""
    _, ns, nv = chp_layout(method)
    h = ChpHead()
    sc = np.zeros(max(ns, 1))
    vecs = np.zeros((3 + nv, nrows), np.complex128)
    ptrs = (ctypes.c_void_p * (3 + nv))(*[v.ctypes.data for v in vecs])
    hcheck(host().adda_chp_read(chp_dir.encode(), ITER_NAMES.get(method, method), nrows, ctypes.byref(h), _p(sc),
                                ctypes.cast(ptrs, ctypes.c_void_p)))
    head = {f: getattr(h, f) for f, _ in ChpHead._fields_}
    return head, sc[:ns], vecs


def chp_write(chp_dir, method, head, scal, vecs, run_dir=None):
    """Writes <chp_dir>/chp.0 and chp.log in ADDA's format (the inverse of chp_read)."""
    _, ns, nv = chp_layout(method)
    vecs = [np.ascontiguousarray(v, np.complex128) for v in vecs]
    assert len(vecs) == 3 + nv and len(scal) == ns
    h = ChpHead(**head)
    sc = np.ascontiguousarray(np.r_[scal, 0.0], np.float64)
    ptrs = (ctypes.c_void_p * (3 + nv))(*[v.ctypes.data for v in vecs])
    hcheck(host().adda_chp_write(chp_dir.encode(), run_dir.encode() if run_dir else None,
                                 ITER_NAMES.get(method, method), len(vecs[0]), ctypes.byref(h), _p(sc),
                                 ctypes.cast(ptrs, ctypes.c_void_p)))


class SolveStats(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_int), ("niter", ctypes.c_int), ("nmatvec", ctypes.c_int),
                ("nmatvec_fp64", ctypes.c_int), ("nouter", ctypes.c_int), ("resid", ctypes.c_double),
                ("true_resid", ctypes.c_double), ("chp_saved", ctypes.c_int)]


_host = None


class Surface(ctypes.Structure):
    _fields_ = [("hsub", ctypes.c_double), ("msub", Cplx), ("inf", ctypes.c_int)]


def incident_frame(prop):
    """prop normalized (-prop, src/param.c:1372-1382) and the incident polarizations incPolX, incPolY of
    src/param.c:2116-2131."""
    prop = np.asarray(prop, np.float64)
    prop = prop / np.linalg.norm(prop)
    if prop[0] == 0 and prop[1] == 0:
        return prop, np.array([np.copysign(1.0, prop[2]), 0, 0]), np.array([0.0, 1.0, 0])
    t = np.sqrt(1 - prop[2] ** 2)
    return (prop, np.array([prop[0] * prop[2] / t, prop[1] * prop[2] / t, -t]),
            np.array([-prop[1] / t, prop[0] / t, 0.0]))


def host():
    """Load libadda.so (C host layer); raise if missing."""
    global _host
    if _host is not None:
        return _host
    hip()
    path = os.path.join(HERE, "libadda.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build first (__graft_entry__.build())")
    L = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.adda_last_error.restype = ctypes.c_char_p
    L.adda_make_particle.argtypes = [ctypes.c_char_p, I, D, D, ctypes.POINTER(Particle)]
    L.adda_read_particle.argtypes = [ctypes.c_char_p, ctypes.POINTER(Particle)]
    L.adda_free_particle.argtypes = [ctypes.POINTER(Particle)]
    L.adda_free_particle.restype = None
    L.adda_volume_ratio.argtypes = [ctypes.c_char_p, D, D]
    L.adda_volume_ratio.restype = D
    L.adda_couple_constant.argtypes = [Cplx, I, D, D, P, P]
    L.adda_couple_constant.restype = Cplx
    L.adda_generate_b.argtypes = [P, P, D, D, P, P, P]
    L.adda_iterative_solver.argtypes = [P, ctypes.POINTER(SolverOpts), P, P, P, P, I, ctypes.POINTER(SolveStats)]
    L.adda_chp_layout.argtypes = [I, P, P, P]
    L.adda_chp_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, I, ctypes.c_size_t, ctypes.POINTER(ChpHead), P, P]
    L.adda_chp_read.argtypes = [ctypes.c_char_p, I, ctypes.c_size_t, ctypes.POINTER(ChpHead), P, P]
    L.adda_iterative_solver_batch.argtypes = [P, ctypes.POINTER(SolverOpts), I, P, P, P, P, I, P]
    L.adda_couple_constant_deps.argtypes = [Cplx, I, D, D, P, P]
    L.adda_couple_constant_deps.restype = Cplx
    L.adda_adjoint_gradient.argtypes = [P, P, ctypes.POINTER(SolverOpts), P, I, D, D, P, P, P, P, P, P, P]
    L.adda_calc_field.argtypes = [P, P, D, D, P, P, I, P]
    L.adda_ext_cross.argtypes = [P, P, D, D, P, P, P, P]
    L.adda_abs_cross.argtypes = [P, P, D, P, P, P]
    L.adda_eplane_yz.argtypes = [P, P, D, D, P, P, P, P, I, I, P]
    L.adda_ext_cross_batch.argtypes = [P, P, D, D, P, I, P, P, P]
    L.adda_abs_cross_batch.argtypes = [P, P, D, P, P, I, P]
    L.adda_eplane_yz_batch.argtypes = [P, P, D, D, P, I, P, P, P, I, I, P]
    L.adda_mueller_yz.argtypes = [P, P, I, P]
    L.adda_mueller_yz.restype = None
    L.adda_set_surface.argtypes = [P, P, D, ctypes.POINTER(Surface)]
    L.adda_generate_b_surf.argtypes = [P, P, ctypes.POINTER(Surface), D, D, P, P, P, I, P, P]
    L.adda_ext_cross_gen.argtypes = [P, D, D, P, P, P]
    _host = L
    return L


def hcheck(rc):
    if rc < 0:
        raise RuntimeError(host().adda_last_error().decode())
    return rc


def fast_grid(box):
    """FFT grid chosen for speed (adda_hip_fastgrid of 2*box along each axis) instead of fftFit's."""
    return tuple(hip().adda_hip_fastgrid(2 * int(b), 1) for b in box)


def make_particle(shape, grid, aspect_y=1.0, aspect_z=1.0):
    """C particle builder -> (position uint16[nd,3], material int32[nd], box)."""
    L = host()
    p = Particle()
    hcheck(L.adda_make_particle(shape.encode(), grid, aspect_y, aspect_z, ctypes.byref(p)))
    try:
        pos = np.ctypeslib.as_array(p.position, shape=(p.nd * 3,)).reshape(-1, 3).copy()
        mat = np.ctypeslib.as_array(p.material, shape=(p.nd,)).copy()
        return pos, mat, (p.boxX, p.boxY, p.boxZ)
    finally:
        L.adda_free_particle(ctypes.byref(p))


def couple_constant(m, polrel, wavenum, gridspace, prop, incpol):
    prop = np.ascontiguousarray(prop, np.float64)
    incpol = np.ascontiguousarray(incpol, np.float64)
    return host().adda_couple_constant(Cplx.of(m), polrel, wavenum, gridspace, _p(prop), _p(incpol)).value()


class Problem:
    """A particle + operator + the incident-field / solve / scattering pipeline of one orientation (the Calculator
    flow of src/calculator.c:293-344) on the GPU."""

    def __init__(self, pos, box, m, mat=None, dpl=None, lam=2 * np.pi, reduced_fft=True, fp32=False, device=0,
                 polrel=POL_LDR, grid=None, surface=None):
        """surface: None (free space) or (hsub, msub) with msub None for a perfectly reflecting substrate
        ('-surf h inf'): the particle's centre hsub above a substrate, image-dipole reflection."""
        self.pos = np.ascontiguousarray(pos, np.uint16)
        self.box = tuple(int(b) for b in box)
        self.nd = len(self.pos)
        self.m = np.atleast_1d(np.asarray(m, np.complex128))
        self.mat = np.zeros(self.nd, np.int32) if mat is None else np.ascontiguousarray(mat, np.int32)
        self.nmat = len(self.m)
        if dpl is None:
            dpl = 10 * float(np.max(np.abs(self.m)))
        self.dpl, self.lam = dpl, lam
        self.wavenum = 2 * np.pi / lam
        self.gridspace = lam / dpl
        self.polrel = polrel
        self.grid = tuple(grid) if grid else tuple(hip().adda_hip_fftfit(2 * b, 1) for b in self.box)
        self.op = Operator(self.pos, self.box, self.grid, self.mat, self.nmat, self.wavenum, self.gridspace,
                           reduced_fft, fp32=fp32, device=device)
        self._p = Particle(self.box[0], self.box[1], self.box[2], self.nd,
                           self.pos.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort)),
                           self.mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self.nmat)
        self.einc = self.op.alloc()
        self.x = self.op.alloc()
        self.pvec = self.op.alloc()
        self.surface = surface
        self.inc_scale = 1.0
        if surface is not None:
            hsub, msub = surface
            self._surf = Surface(hsub, Cplx.of(0 if msub is None else msub), int(msub is None))
            hcheck(host().adda_set_surface(self.op.h, ctypes.byref(self._p), self.gridspace, ctypes.byref(self._surf)))

    def close(self):
        if self.op is not None:
            for v in (self.einc, self.x, self.pvec):
                self.op.free(v)
            self.op.close()
            self.op = None

    def set_materials(self, m):
        """New refractive indices for the same material table size (takes effect at the next set_polarization)."""
        m = np.atleast_1d(np.asarray(m, np.complex128))
        if len(m) != self.nmat:
            raise ValueError(f"{len(m)} refractive indices for {self.nmat} materials")
        self.m = m

    def set_polarization(self, prop, incpol):
        """InitCC (cc, cc_sqrt for every material) + GenerateB for this incident polarization."""
        self.prop = np.ascontiguousarray(prop, np.float64)
        self.incpol = np.ascontiguousarray(incpol, np.float64)
        cc = np.array([couple_constant(mi, self.polrel, self.wavenum, self.gridspace, self.prop, self.incpol)
                       for mi in self.m])
        self.cc = np.repeat(cc[:, None], 3, axis=1)
        self.cc_sqrt = np.sqrt(self.cc)
        self.op.set_cc_sqrt(self.cc_sqrt)
        if self.surface is None:
            hcheck(host().adda_generate_b(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace,
                                          _p(self.prop), _p(self.incpol), self.einc))
            return
        # near the substrate the beam is one of the two polarizations of prop's incident frame
        _, px, py = incident_frame(self.prop)
        if np.allclose(self.incpol, py, atol=1e-12):
            which = 0
        elif np.allclose(self.incpol, px, atol=1e-12):
            which = 1
        else:
            raise ValueError("near a substrate incpol must be incPolY or incPolX of incident_frame(prop)")
        sc = ctypes.c_double()
        hcheck(host().adda_generate_b_surf(self.op.h, ctypes.byref(self._p), ctypes.byref(self._surf), self.wavenum,
                                           self.gridspace, _p(self.prop), _p(np.ascontiguousarray(px)),
                                           _p(np.ascontiguousarray(py)), which, self.einc, ctypes.byref(sc)))
        self.inc_scale = sc.value

    def solve(self, method="qmr", eps=1e-5, maxiter=None, init_field=IF_AUTO, precision=PREC_FP64, inner_eps=0.0,
              max_outer=0, recalc_resid=False, hist_len=20000, verbose=False, host_scalars=False, max_inner=0,
              graph=False, chp_type=CHP_NONE, chp_time=0.0, chp_dir=None, chp_load=False, inner_method=None):
        """IterativeSolver; chp_*: ADDA's iteration checkpoints (FP64, host-form recurrences); inner_method: the
        FP32 inner solver of the mixed forms (None: automatic), see include/adda.h."""
        if maxiter is None:
            maxiter = 3 * self.box[0] * self.box[1] * self.box[2]
        o = SolverOpts(ITER_NAMES.get(method, method), eps, maxiter, init_field, int(recalc_resid), precision,
                       inner_eps, max_outer, int(verbose), int(host_scalars), int(max_inner), int(graph),
                       chp_type, chp_time, chp_dir.encode() if chp_dir else None, int(chp_load), None,
                       _inner(inner_method))
        st = SolveStats()
        hist = np.zeros(hist_len)
        rc = host().adda_iterative_solver(self.op.h, ctypes.byref(o), self.einc, self.x, self.pvec, _p(hist),
                                          hist_len, ctypes.byref(st))
        hcheck(rc)
        n = min(st.niter + 1, hist_len)
        return dict(rc=st.rc, niter=st.niter, nmatvec=st.nmatvec, nmatvec_fp64=st.nmatvec_fp64, nouter=st.nouter,
                    resid=st.resid, true_resid=st.true_resid, hist=hist[:n].copy())

    def adjoint_gradient(self, method="qmr", eps=1e-10, maxiter=None, precision=PREC_FP64, inner_eps=0.0,
                         max_outer=0, graph=True, max_inner=0, inner_method=None):
        """d(Cext, Cabs)/d(eps_j) for every dipole j of the current solution (call solve() first): two adjoint
        solves batched together, then per-dipole assembly on the GPU. Returns gext, gabs [nd][2] (d/dRe eps_j,
        d/dIm eps_j) and the stats of the two adjoint solves."""
        if self.surface is not None:
            raise NotImplementedError("adjoint gradients near a substrate are not implemented")
        if maxiter is None:
            maxiter = 3 * self.box[0] * self.box[1] * self.box[2]
        o = SolverOpts(ITER_NAMES.get(method, method), eps, maxiter, IF_AUTO, 0, precision, inner_eps, max_outer, 0, 0,
                       int(max_inner), int(graph), inner_method=_inner(inner_method))
        ge, ga = np.empty(2 * self.nd), np.empty(2 * self.nd)
        st = (SolveStats * 2)()
        m = np.ascontiguousarray(self.m)
        hcheck(host().adda_adjoint_gradient(self.op.h, ctypes.byref(self._p), ctypes.byref(o), _p(m), self.polrel,
                                            self.wavenum, self.gridspace, _p(self.prop), _p(self.incpol), self.einc,
                                            self.pvec, _p(ge), _p(ga), st))
        stats = [dict(rc=s.rc, niter=s.niter, nmatvec=s.nmatvec, nouter=s.nouter, true_resid=s.true_resid) for s in st]
        return ge.reshape(-1, 2), ga.reshape(-1, 2), stats

    def cext(self):
        out = ctypes.c_double()
        if self.surface is not None:
            hcheck(host().adda_ext_cross_gen(self.op.h, self.wavenum, self.inc_scale, self.pvec, self.einc,
                                             ctypes.byref(out)))
            return out.value
        hcheck(host().adda_ext_cross(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace, self.pvec,
                                     _p(self.prop), _p(self.incpol), ctypes.byref(out)))
        return out.value

    def cabs(self):
        out = ctypes.c_double()
        cc = np.ascontiguousarray(self.cc)
        hcheck(host().adda_abs_cross(self.op.h, ctypes.byref(self._p), self.wavenum, _p(cc), self.pvec,
                                     ctypes.byref(out)))
        return out.value * self.inc_scale

    def eplane(self, par, per, ntheta, npoints=None):
        if self.surface is not None:
            raise NotImplementedError("the scattered field near a substrate (CalcFieldSurf) is not implemented")
        npoints = ntheta if npoints is None else npoints
        out = np.empty(2 * npoints, np.complex128)
        par = np.ascontiguousarray(par, np.float64)
        per = np.ascontiguousarray(per, np.float64)
        hcheck(host().adda_eplane_yz(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace, self.pvec,
                                     _p(self.prop), _p(par), _p(per), ntheta, npoints, _p(out)))
        return out

    def polarization(self):
        return self.op.download(self.pvec)

    def field_x(self):
        return self.op.download(self.x)


class BatchProblem:
    """K right-hand sides of one particle -- incident directions and polarizations, e.g. the orientations x
    polarizations of an orientation average (src/crosssec.c:82-138) -- solved together on one operator: one D-matrix,
    per-RHS cc_sqrt tables (LDR depends on prop and polarization), batched device vectors, one batched launch per
    kernel for all RHS still iterating."""

    def __init__(self, pos, box, m, props, pols, mat=None, dpl=None, lam=2 * np.pi, reduced_fft=True, fp32=False,
                 device=0, polrel=POL_LDR, grid=None):
        self.pos = np.ascontiguousarray(pos, np.uint16)
        self.box = tuple(int(b) for b in box)
        self.nd = len(self.pos)
        self.m = np.atleast_1d(np.asarray(m, np.complex128))
        self.mat = np.zeros(self.nd, np.int32) if mat is None else np.ascontiguousarray(mat, np.int32)
        self.nmat = len(self.m)
        if dpl is None:
            dpl = 10 * float(np.max(np.abs(self.m)))
        self.dpl, self.lam = dpl, lam
        self.wavenum = 2 * np.pi / lam
        self.gridspace = lam / dpl
        self.polrel = polrel
        self.props = np.ascontiguousarray(props, np.float64).reshape(-1, 3)
        self.pols = np.ascontiguousarray(pols, np.float64).reshape(-1, 3)
        self.K = K = len(self.props)
        self.grid = tuple(grid) if grid else tuple(hip().adda_hip_fftfit(2 * b, 1) for b in self.box)
        self.op = Operator(self.pos, self.box, self.grid, self.mat, self.nmat, self.wavenum, self.gridspace,
                           reduced_fft, fp32=fp32, device=device)
        self.op.reserve(K)
        self._p = Particle(self.box[0], self.box[1], self.box[2], self.nd,
                           self.pos.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort)),
                           self.mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self.nmat)
        self.cc = np.empty((K, self.nmat, 3), np.complex128)
        for k in range(K):
            for i, mi in enumerate(self.m):
                self.cc[k, i, :] = couple_constant(mi, polrel, self.wavenum, self.gridspace, self.props[k], self.pols[k])
        self.cc_sqrt = np.sqrt(self.cc)
        self.op.set_cc_sqrt_batch(self.cc_sqrt)
        self.einc = self.op.alloc_batch(K)
        self.x = self.op.alloc_batch(K)
        self.pvec = self.op.alloc_batch(K)
        for k in range(K):
            hcheck(host().adda_generate_b(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace,
                                          _p(self.props[k]), _p(self.pols[k]), self.op.rhs(self.einc, k)))

    def close(self):
        if self.op is not None:
            for v in (self.einc, self.x, self.pvec):
                self.op.free(v)
            self.op.close()
            self.op = None

    def solve(self, method="qmr", eps=1e-5, maxiter=None, init_field=IF_AUTO, precision=PREC_FP64, inner_eps=0.0,
              max_outer=0, recalc_resid=False, hist_len=0, graph=True, max_inner=0, inner_method=None,
              raise_on_error=True):
        """adda_iterative_solver_batch; per-RHS results. raise_on_error=False returns them even when some RHS failed
        (their rc < 0), instead of raising the error of the first failing one."""
        if maxiter is None:
            maxiter = 3 * self.box[0] * self.box[1] * self.box[2]
        o = SolverOpts(ITER_NAMES.get(method, method), eps, maxiter, init_field, int(recalc_resid), precision,
                       inner_eps, max_outer, 0, 0, int(max_inner), int(graph), inner_method=_inner(inner_method))
        st = (SolveStats * self.K)()
        hist = np.zeros((self.K, max(hist_len, 1)))
        rc = host().adda_iterative_solver_batch(self.op.h, ctypes.byref(o), self.K, self.einc, self.x, self.pvec,
                                                _p(hist) if hist_len else None, hist_len, st)
        if raise_on_error:
            hcheck(rc)
        out = [dict(rc=s.rc, niter=s.niter, nmatvec=s.nmatvec, nmatvec_fp64=s.nmatvec_fp64, nouter=s.nouter,
                    resid=s.resid, true_resid=s.true_resid) for s in st]
        if hist_len:
            for k, d in enumerate(out):
                d["hist"] = hist[k, :min(d["niter"] + 1, hist_len)].copy()
        return out

    def cext(self, k):
        out = ctypes.c_double()
        hcheck(host().adda_ext_cross(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace,
                                     self.op.rhs(self.pvec, k), _p(self.props[k]), _p(self.pols[k]), ctypes.byref(out)))
        return out.value

    def cabs(self, k):
        out = ctypes.c_double()
        cc = np.ascontiguousarray(self.cc[k])
        hcheck(host().adda_abs_cross(self.op.h, ctypes.byref(self._p), self.wavenum, _p(cc), self.op.rhs(self.pvec, k),
                                     ctypes.byref(out)))
        return out.value

    def cext_all(self):
        """Cext of every RHS from one batched evaluation (adda_ext_cross_batch)."""
        out = np.empty(self.K)
        hcheck(host().adda_ext_cross_batch(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace, self.pvec,
                                           self.K, _p(self.props), _p(self.pols), _p(out)))
        return out

    def cabs_all(self):
        """Cabs of every RHS from one batched evaluation (adda_abs_cross_batch)."""
        out = np.empty(self.K)
        cc = np.ascontiguousarray(self.cc)
        hcheck(host().adda_abs_cross_batch(self.op.h, ctypes.byref(self._p), self.wavenum, _p(cc), self.pvec, self.K,
                                           _p(out)))
        return out

    def eplane(self, k, par, per, ntheta, npoints=None):
        """Scattered field of RHS k in the plane spanned by its propagation direction and par (CalcEplaneYZ):
        [npoints, 2] complex, columns E.per and E.par, theta_i = i*180/(ntheta-1)."""
        npoints = npoints or ntheta
        out = np.empty((npoints, 2), np.complex128)
        hcheck(host().adda_eplane_yz(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace,
                                     self.op.rhs(self.pvec, k), _p(self.props[k]), _p(np.asarray(par, np.float64)),
                                     _p(np.asarray(per, np.float64)), ntheta, npoints, _p(out)))
        return out

    def eplane_all(self, pars, pers, ntheta, npoints=None):
        """eplane for every RHS from one batched evaluation (adda_eplane_yz_batch): [K, npoints, 2] complex;
        pars, pers: [K, 3], the scattering-plane vectors of each RHS's beam."""
        npoints = npoints or ntheta
        pars = np.ascontiguousarray(pars, np.float64).reshape(self.K, 3)
        pers = np.ascontiguousarray(pers, np.float64).reshape(self.K, 3)
        out = np.empty((self.K, npoints, 2), np.complex128)
        hcheck(host().adda_eplane_yz_batch(self.op.h, ctypes.byref(self._p), self.wavenum, self.gridspace, self.pvec,
                                           self.K, _p(self.props), _p(pars), _p(pers), ntheta, npoints, _p(out)))
        return out

    def field_x(self, k):
        return self.op.download(self.op.rhs(self.x, k))

    def polarization(self, k):
        return self.op.download(self.op.rhs(self.pvec, k))


def mueller_yz(eX, eY, ntheta):
    out = np.empty((ntheta, 16))
    host().adda_mueller_yz(_p(np.ascontiguousarray(eX)), _p(np.ascontiguousarray(eY)), ntheta, _p(out))
    return out


def rotate_zyz(alpha, beta, gamma, v):
    """InitRotation (src/crosssec.c:82-118): beta_matr . v for zyz Euler angles in degrees."""
    a, b, g = np.radians([alpha, beta, gamma])
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    m = np.array([[ca * cb * cg - sa * sg, sa * cb * cg + ca * sg, -sb * cg],
                  [-ca * cb * sg - sa * cg, -sa * cb * sg + ca * cg, sb * sg],
                  [ca * sb, sa * sb, cb]])
    return m @ np.asarray(v, np.float64)
