"""Test-side case builders: ADDA particle rules restated with numpy, plus ctypes access to the oracle library.

Geometry follows MakeParticle (src/make_particle.c:1952-2060): with jagged=1 the dipole centre of cell (i,j,k) is
x_r = (i - box/2 + 0.5)/boxX (all axes scaled by boxX); SH_SPHERE keeps x_r^2+y_r^2+z_r^2 <= 0.25 (2136-2137),
SH_ELLIPSOID keeps x^2 + y^2/aY^2 + z^2/aZ^2 <= 0.25 (invsqY/invsqZ, 1635-1637), SH_BOX keeps |y|<=aY/2, |z|<=aZ/2.
Box sizes follow FitBox/FitBox_yz (1283-1307). Dipoles are ordered z-major, then y, then x (the loop order).
"""
import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TWO_PI = 2 * math.pi


class OrCplx(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


class OrSurface(ctypes.Structure):
    _fields_ = [("zsum_shift", ctypes.c_double), ("rcn", OrCplx), ("rmat", ctypes.c_void_p)]


class OrGrid(ctypes.Structure):
    _fields_ = [("boxX", ctypes.c_int), ("boxY", ctypes.c_int), ("boxZ", ctypes.c_int),
                ("gridX", ctypes.c_int), ("gridY", ctypes.c_int), ("gridZ", ctypes.c_int),
                ("nd", ctypes.c_size_t), ("pos", ctypes.c_void_p), ("mat", ctypes.c_void_p),
                ("wavenum", ctypes.c_double), ("gridspace", ctypes.c_double), ("reduced_fft", ctypes.c_int),
                ("surf", ctypes.POINTER(OrSurface))]


def fit_box(box):
    """FitBox with jagged=1 (src/make_particle.c:1284-1291): smallest even number >= box."""
    return 2 * ((box + 1) // 2)


def fit_box_yz(size):
    """FitBox_yz with jagged=1 (src/make_particle.c:1295-1307)."""
    return 2 * int(math.floor((size + 1) / 2))


def make_shape(shape, box, aspect_y=1.0, aspect_z=1.0):
    """Return (position uint16[nd,3], (boxX, boxY, boxZ)) for 'sphere', 'ellipsoid', 'box'."""
    bx = fit_box(box)
    if shape == "sphere":
        by = bz = bx
    else:
        by = fit_box(fit_box_yz(aspect_y * bx))
        bz = fit_box(fit_box_yz(aspect_z * bx))
    k, j, i = np.meshgrid(np.arange(bz), np.arange(by), np.arange(bx), indexing="ij")
    xr = (i - bx // 2 + 0.5) / bx
    yr = (j - by // 2 + 0.5) / bx
    zr = (k - bz // 2 + 0.5) / bx
    if shape == "sphere":
        keep = xr * xr + yr * yr + zr * zr <= 0.25
    elif shape == "ellipsoid":
        keep = xr * xr + yr * yr / (aspect_y ** 2) + zr * zr / (aspect_z ** 2) <= 0.25
    elif shape == "box":
        keep = (np.abs(yr) <= aspect_y / 2) & (np.abs(zr) <= aspect_z / 2)
    else:
        raise ValueError(shape)
    pos = np.stack([i[keep], j[keep], k[keep]], axis=1).astype(np.uint16)
    return np.ascontiguousarray(pos), (bx, by, bz)


def read_shape(pos):
    """What ADDA makes of a '-shape read' file holding these dipoles (InitDipFile/ReadDipFile,
    src/make_particle.c:1150-1260): positions shifted to start at 0, box = FitBox(max-min+1) per axis."""
    pos = np.asarray(pos, np.int64)
    mn = pos.min(axis=0)
    p = (pos - mn).astype(np.uint16)
    box = tuple(fit_box(int(e)) for e in (p.max(axis=0) + 1))
    return np.ascontiguousarray(p), box


_lib = None


def oracle():
    global _lib
    if _lib is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.or_fftfit.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.or_fftfit.restype = ctypes.c_int
        lib.or_green_poi.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [P]
        lib.or_fft1d.argtypes = [P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
        lib.or_dmatrix_size.argtypes = [P]
        lib.or_dmatrix_size.restype = ctypes.c_size_t
        lib.or_init_dmatrix.argtypes = [P, P]
        lib.or_matvec_fft.argtypes = [P, P, P, P, P, P, ctypes.c_int]
        lib.or_matvec_direct.argtypes = [P, P, P, P, P, ctypes.c_int]
        lib.or_couple_constant.argtypes = [OrCplx, ctypes.c_int, ctypes.c_double, ctypes.c_double, P, P]
        lib.or_couple_constant.restype = OrCplx
        lib.or_couple_constant_deps.argtypes = [OrCplx, ctypes.c_int, ctypes.c_double, ctypes.c_double, P, P]
        lib.or_couple_constant_deps.restype = OrCplx
        lib.or_adjoint_grad.argtypes = [P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, P, P]
        lib.or_generate_b.argtypes = [P, P, P, P]
        lib.or_iterative_solver.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                            P, P, P, ctypes.c_int, P, P]
        lib.or_iterative_solver.restype = ctypes.c_int
        lib.or_calc_field.argtypes = [P, P, P, P]
        lib.or_ext_cross.argtypes = [P, P, P, P]
        lib.or_ext_cross.restype = ctypes.c_double
        lib.or_abs_cross.argtypes = [P, P, P]
        lib.or_abs_cross.restype = ctypes.c_double
        lib.or_eplane_yz.argtypes = [P, P, P, P, P, ctypes.c_int, ctypes.c_int, P]
        lib.or_mueller_yz.argtypes = [P, P, ctypes.c_int, P]
        lib.or_refl_img.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, OrCplx, ctypes.c_double, ctypes.c_double, P]
        lib.or_rmatrix_size.argtypes = [P]
        lib.or_rmatrix_size.restype = ctypes.c_size_t
        lib.or_init_rmatrix.argtypes = [P, P]
        lib.or_generate_b_surf.argtypes = [P, OrCplx, ctypes.c_int, ctypes.c_double, P, P, P, ctypes.c_int, P]
        lib.or_generate_b_surf.restype = ctypes.c_double
        lib.or_ext_cross_gen.argtypes = [P, P, P]
        lib.or_ext_cross_gen.restype = ctypes.c_double
        _lib = lib
    return _lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Case:
    """One DDA problem: geometry, grid, wavelength, materials; matches ADDA's variables for the same command line."""

    def __init__(self, pos, box, m, dpl=None, lam=TWO_PI, mat=None, rule=1, reduced_fft=True, grid=None,
                 surface=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.uint16)
        self.box = tuple(int(b) for b in box)
        self.nd = len(self.pos)
        self.m = np.atleast_1d(np.asarray(m, dtype=np.complex128))
        self.mat = np.zeros(self.nd, np.int32) if mat is None else np.ascontiguousarray(mat, dtype=np.int32)
        if dpl is None:  # dpl_def = 10*|m| (src/make_particle.c:1353)
            dpl = 10 * float(np.max(np.abs(self.m)))
        self.dpl = dpl
        self.lam = lam
        self.wavenum = TWO_PI / lam
        self.gridspace = lam / dpl
        lib = oracle()
        if grid is None:  # ParSetup with one process (src/comm.c:489-494)
            grid = tuple(lib.or_fftfit(2 * b, 1, rule) for b in self.box)
        self.grid = tuple(grid)
        self.reduced_fft = reduced_fft
        self._g = OrGrid(*self.box, *self.grid, self.nd, self.pos.ctypes.data, self.mat.ctypes.data,
                         self.wavenum, self.gridspace, int(reduced_fft))
        # surface = (hsub, msub) with msub None for a perfect reflector ('-surf h inf'): ZsumShift
        # (src/make_particle.c:2312) and surfRCn (src/interaction.c:1411)
        self.surface = surface
        self._rmat = None
        if surface is not None:
            hsub, msub = surface
            rcn = -1.0 if msub is None else (1 - msub * msub) / (1 + msub * msub)
            self.zsum_shift = 2 * (hsub / self.gridspace - (self.box[2] - 1) / 2.0)
            self._surf = OrSurface(self.zsum_shift, OrCplx(complex(rcn).real, complex(rcn).imag), None)
            self._g.surf = ctypes.pointer(self._surf)
        self.gp = ctypes.byref(self._g)

    def cc(self, prop, incpol, polrel=0):
        """cc[nmat][3] (isotropic => 3 equal components) and cc_sqrt = csqrt(cc) (InitCC, src/calculator.c:266-289)."""
        lib = oracle()
        prop = np.ascontiguousarray(prop, np.float64)
        incpol = np.ascontiguousarray(incpol, np.float64)
        out = np.empty((len(self.m), 3), np.complex128)
        for i, mi in enumerate(self.m):
            r = lib.or_couple_constant(OrCplx(mi.real, mi.imag), polrel, self.wavenum, self.gridspace, ptr(prop),
                                       ptr(incpol))
            out[i, :] = complex(r.re, r.im)
        return out, np.sqrt(out)

    def cc_deps(self, prop, incpol, polrel=0):
        """d(cc)/d(eps) per material [nmat][3] at eps = m^2."""
        lib = oracle()
        prop = np.ascontiguousarray(prop, np.float64)
        incpol = np.ascontiguousarray(incpol, np.float64)
        out = np.empty((len(self.m), 3), np.complex128)
        for i, mi in enumerate(self.m):
            r = lib.or_couple_constant_deps(OrCplx(mi.real, mi.imag), polrel, self.wavenum, self.gridspace, ptr(prop),
                                            ptr(incpol))
            out[i, :] = complex(r.re, r.im)
        return out

    def with_m(self, m):
        """The same geometry and grid with another material table."""
        return Case(self.pos, self.box, m, dpl=self.dpl, lam=self.lam, mat=self.mat, reduced_fft=self.reduced_fft,
                    grid=self.grid, surface=self.surface)

    def adjoint_grad(self, dmat, prop, incpol, polrel=0, method=0, eps=1e-12):
        """Oracle adjoint gradient of (Cext, Cabs) w.r.t. each dipole's permittivity: forward solve, adjoint solves
        A.y = S.conj(Einc) and A.y' = S.conj(W.P), assembly (or_adjoint_grad). Returns gext, gabs [nd][2]
        (d/dRe eps, d/dIm eps), Cext, Cabs."""
        cc, ccs = self.cc(prop, incpol, polrel)
        e = self.einc(prop, incpol)
        P = self.solve(dmat, ccs, e, method=method, eps=eps)["p"]
        k3 = 2 * self.wavenum ** 3 / 3
        wabs = -np.imag(1 / cc) - k3
        lam = self.solve(dmat, ccs, np.conj(e), method=method, eps=eps)["p"]
        W = wabs[self.mat].reshape(-1)
        mu = self.solve(dmat, ccs, np.conj(W * P), method=method, eps=eps)["p"]
        q = np.ascontiguousarray(-self.cc_deps(prop, incpol, polrel) / cc ** 2)
        ge, ga = np.empty(2 * self.nd), np.empty(2 * self.nd)
        ce = ca = 4 * np.pi * self.wavenum
        oracle().or_adjoint_grad(self.gp, ptr(P), ptr(lam), ptr(mu), ptr(q), ptr(np.ascontiguousarray(wabs)), ce, ca,
                                 ptr(ge), ptr(ga))
        return ge.reshape(-1, 2), ga.reshape(-1, 2), self.cext(P, prop, incpol), self.cabs(P, cc)

    def fd_grad(self, dmat, prop, incpol, j, h=1e-6, polrel=0, method=0, eps=1e-13):
        """Central finite differences of (Cext, Cabs) in Re and Im of material j's permittivity (every dipole of
        material j), from forward solves only. C(eps +- h) - C(eps) is evaluated as an exact difference: with
        M' = M + dM (dM non-zero only on material j), P' - P solves M'(P' - P) = -dM.P -- a forward solve of the
        perturbed system with a right-hand side local to material j -- so no cancellation of O(1) sums occurs.
        Returns [[dCext/dRe, dCext/dIm], [dCabs/dRe, dCabs/dIm]]."""
        cc, ccs = self.cc(prop, incpol, polrel)
        e = self.einc(prop, incpol)
        P = self.solve(dmat, ccs, e, method=method, eps=eps)["p"]
        k3 = 2 * self.wavenum ** 3 / 3
        w = (-np.imag(1 / cc) - k3)[self.mat].reshape(-1)
        sel = np.repeat(self.mat == j, 3)
        out = np.zeros((2, 2))
        for col, dz in ((0, h), (1, 1j * h)):
            dC = []
            for sgn in (1, -1):
                m2 = self.m.copy()
                m2[j] = np.sqrt(self.m[j] ** 2 + sgn * dz)
                c2 = self.with_m(m2)
                cc2, ccs2 = c2.cc(prop, incpol, polrel)
                dMdiag = (1 / cc2 - 1 / cc)[self.mat].reshape(-1)  # non-zero on material j only
                rhs = np.where(sel, -dMdiag * P, 0)
                dP = c2.solve(dmat, ccs2, rhs, method=method, eps=eps)["p"]
                w2 = (-np.imag(1 / cc2) - k3)[self.mat].reshape(-1)
                dext = 4 * np.pi * self.wavenum * np.imag(np.sum(np.conj(e) * dP))
                dabs = 4 * np.pi * self.wavenum * (np.sum(w * (2 * np.real(np.conj(P) * dP) + np.abs(dP) ** 2))
                                                    + np.sum((w2 - w) * np.abs(P + dP) ** 2))
                dC.append((dext, dabs))
            out[0, col] = (dC[0][0] - dC[1][0]) / (2 * h)
            out[1, col] = (dC[0][1] - dC[1][1]) / (2 * h)
        return out

    def dmatrix(self):
        """D-matrix; near a substrate the R-matrix is built alongside and attached to the grid (the FFT MatVec and
        the solver then add the reflected term)."""
        lib = oracle()
        d = np.empty(lib.or_dmatrix_size(self.gp), np.complex128)
        lib.or_init_dmatrix(self.gp, ptr(d))
        if self.surface is not None and self._rmat is None:
            self._rmat = self.rmatrix()
            self._surf.rmat = self._rmat.ctypes.data
        return d

    def rmatrix(self):
        lib = oracle()
        r = np.empty(lib.or_rmatrix_size(self.gp), np.complex128)
        lib.or_init_rmatrix(self.gp, ptr(r))
        return r

    def einc_surf(self, prop, incpol_x, incpol_y, which_x):
        """Incident field near the substrate and inc_scale (or_generate_b_surf)."""
        hsub, msub = self.surface
        e = np.empty(3 * self.nd, np.complex128)
        ms = 1.0 if msub is None else complex(msub)
        sc = oracle().or_generate_b_surf(self.gp, OrCplx(ms.real, ms.imag), int(msub is None), hsub,
                                         ptr(np.ascontiguousarray(prop, np.float64)),
                                         ptr(np.ascontiguousarray(incpol_x, np.float64)),
                                         ptr(np.ascontiguousarray(incpol_y, np.float64)), int(which_x), ptr(e))
        return e, sc

    def cext_gen(self, p, einc):
        return oracle().or_ext_cross_gen(self.gp, ptr(p), ptr(np.ascontiguousarray(einc)))

    def matvec_fft(self, dmat, cc_sqrt, x, her=False):
        lib = oracle()
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty_like(x)
        ip = ctypes.c_double()
        lib.or_matvec_fft(self.gp, ptr(dmat), ptr(np.ascontiguousarray(cc_sqrt)), ptr(x), ptr(y), ctypes.byref(ip),
                          int(her))
        return y, ip.value

    def matvec_direct(self, cc_sqrt, x, her=False):
        lib = oracle()
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty_like(x)
        ip = ctypes.c_double()
        lib.or_matvec_direct(self.gp, ptr(np.ascontiguousarray(cc_sqrt)), ptr(x), ptr(y), ctypes.byref(ip), int(her))
        return y, ip.value

    def einc(self, prop, incpol):
        lib = oracle()
        e = np.empty(3 * self.nd, np.complex128)
        lib.or_generate_b(self.gp, ptr(np.ascontiguousarray(prop, np.float64)),
                          ptr(np.ascontiguousarray(incpol, np.float64)), ptr(e))
        return e

    def solve(self, dmat, cc_sqrt, einc, method=0, eps=1e-5, maxiter=None, init_field=0, hist=1000):
        lib = oracle()
        if maxiter is None:  # 3*Ndip (src/make_particle.c:1939)
            maxiter = 3 * self.box[0] * self.box[1] * self.box[2]
        x = np.empty(3 * self.nd, np.complex128)
        p = np.empty_like(x)
        h = np.zeros(hist)
        ni, nm = ctypes.c_int(), ctypes.c_int()
        rc = lib.or_iterative_solver(self.gp, None if dmat is None else ptr(dmat), ptr(np.ascontiguousarray(cc_sqrt)),
                                     ptr(np.ascontiguousarray(einc)), method, eps, maxiter, init_field, ptr(x), ptr(p),
                                     ptr(h), hist, ctypes.byref(ni), ctypes.byref(nm))
        return dict(rc=rc, x=x, p=p, hist=h[:ni.value + 1].copy(), niter=ni.value, nmatvec=nm.value)

    def cext(self, p, prop, incpol):
        return oracle().or_ext_cross(self.gp, ptr(p), ptr(np.ascontiguousarray(prop, np.float64)),
                                     ptr(np.ascontiguousarray(incpol, np.float64)))

    def cabs(self, p, cc):
        return oracle().or_abs_cross(self.gp, ptr(p), ptr(np.ascontiguousarray(cc)))

    def eplane(self, p, prop, par, per, ntheta, npoints=None):
        npoints = ntheta if npoints is None else npoints
        e = np.empty(2 * npoints, np.complex128)
        oracle().or_eplane_yz(self.gp, ptr(p), ptr(np.ascontiguousarray(prop, np.float64)),
                              ptr(np.ascontiguousarray(par, np.float64)), ptr(np.ascontiguousarray(per, np.float64)),
                              ntheta, npoints, ptr(e))
        return e

    def mueller(self, eX, eY, npoints):
        out = np.empty((npoints, 16))
        oracle().or_mueller_yz(ptr(eX), ptr(eY), npoints, ptr(out))
        return out


PROP_Z = np.array([0.0, 0.0, 1.0])
POL_Y = np.array([0.0, 1.0, 0.0])
POL_X = np.array([1.0, 0.0, 0.0])


def adj8_materials(nd):
    """The at-most-15-material variant of the per-voxel gradient workload (box 8^3): 14 refractive indices drawn
    like the per-voxel ones (Re ~ U(1.2,1.8), Im ~ U(0,0.05), seed 20261204, rounded to 4 decimals), voxels assigned at
    random; four finite-difference voxels (seed 20261205), each moved alone to material 15 = its own m +- ADJ_DELTA
    (real or imaginary part) in the reference runs of tests/golden/adj8/."""
    rng = np.random.default_rng(20261204)
    m = np.round(rng.uniform(1.2, 1.8, 14), 4) + 1j * np.round(rng.uniform(0, 0.05, 14), 4)
    mat = rng.integers(0, 14, nd)
    vox = np.random.default_rng(20261205).choice(nd, 4, replace=False)
    return m, mat, vox


ADJ_DELTA = 1e-3


def per_voxel_m(nd, seed=20261204):
    """Per-voxel refractive indices of the gradient workload: Re ~ U(1.2,1.8), Im ~ U(0,0.05)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(1.2, 1.8, nd) + 1j * rng.uniform(0, 0.05, nd)


def orientations(n, seed=20261203):
    """n random particle orientations as zyz Euler angles in degrees (ADDA's -orient alpha beta gamma,
    src/crosssec.c:82-110), uniform over the rotation group: alpha, gamma ~ U[0,360), beta = arccos(U[-1,1]);
    rounded to 6 decimals, the precision they are passed to the reference with."""
    u = np.random.default_rng(seed).random((n, 3))
    a = np.c_[360 * u[:, 0], np.degrees(np.arccos(2 * u[:, 1] - 1)), 360 * u[:, 2]]
    return np.round(a, 6)


def sample_case():
    """Config 1: sphere 16^3 (2176 dipoles), m=1.5, dpl=15 (ADDA defaults), grid 32^3."""
    pos, box = make_shape("sphere", 16)
    return Case(pos, box, 1.5, dpl=15)
