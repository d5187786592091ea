"""GPU parity of the HIP MatVec (src/matvec.c:128-425) against the oracle's FFT matvec and direct sum.

FP64 bar: relative error <= 1e-12 (max-norm over the vector, relative to max |y|). FP32 bar: <= 2e-5.
Shapes cover the configs' grid factorizations (32, 64, 96=2^5*3, 108=2^2*3^3, 128) and ragged boxes.
"""
import numpy as np
import pytest

from cases import Case, make_shape, PROP_Z, POL_Y

pytestmark = pytest.mark.gpu


def _case(shape, box, m, ay=1.0, az=1.0, **kw):
    pos, b = make_shape(shape, box, ay, az)
    return Case(pos, b, m, **kw)


def _op(adda, c, fp32=False):
    return adda.Operator(c.pos, c.box, c.grid, c.mat, len(c.m), c.wavenum, c.gridspace, c.reduced_fft, fp32=fp32)


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


CASES = [
    ("sphere", 16, 1.5, 1.0, 1.0, {}),                     # config 1 (grid 32^3)
    ("ellipsoid", 16, 1.33 + 0.001j, 0.8, 0.5, {}),        # ragged box 16x14x8
    ("box", 10, 1.2 + 0.05j, 0.7, 0.4, {}),                # box 10x8x4 -> grid 20x16x8
    ("sphere", 12, 1.1 + 0.1j, 1.0, 1.0, {"reduced_fft": False}),
    ("ellipsoid", 24, 1.5 + 0.01j, 1.0, 0.5, {}),          # grid 48x48x24
]


@pytest.mark.parametrize("shape,box,m,ay,az,kw", CASES)
def test_dmatrix_matches_oracle(hipdev, shape, box, m, ay, az, kw):
    c = _case(shape, box, m, ay, az, **kw)
    op = _op(hipdev, c)
    d_gpu = op.dmatrix()
    d_or = c.dmatrix()
    assert d_gpu.shape == d_or.shape
    assert _rel(d_gpu, d_or) < 1e-12


@pytest.mark.parametrize("shape,box,m,ay,az,kw", CASES)
@pytest.mark.parametrize("her", [False, True])
def test_matvec_fp64_vs_oracle(hipdev, shape, box, m, ay, az, kw, her):
    c = _case(shape, box, m, ay, az, **kw)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c)
    op.set_cc_sqrt(ccs)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y, ip = op.matvec(x, her=her)
    y_ref, ip_ref = c.matvec_fft(c.dmatrix(), ccs, x, her=her)
    assert _rel(y, y_ref) < 1e-12
    assert abs(ip - ip_ref) / ip_ref < 1e-12


def test_matvec_fp64_vs_direct_sum(hipdev):
    """FFT convolution == direct sum (what ADDA's SPARSE build computes), tiny ragged particle."""
    c = _case("ellipsoid", 8, 1.3 + 0.2j, 0.7, 0.5)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c)
    op.set_cc_sqrt(ccs)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y, _ = op.matvec(x)
    y_ref, _ = c.matvec_direct(ccs, x)
    assert _rel(y, y_ref) < 1e-12


@pytest.mark.parametrize("shape,box,m,ay,az,kw", CASES[:3])
def test_matvec_fp32(hipdev, shape, box, m, ay, az, kw):
    c = _case(shape, box, m, ay, az, **kw)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    op.set_cc_sqrt(ccs)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y, _ = op.matvec(x, prec=hipdev.F32)
    y_ref, _ = c.matvec_fft(c.dmatrix(), ccs, x)
    assert _rel(y, y_ref) < 2e-5


@pytest.mark.parametrize("her", [False, True])
@pytest.mark.parametrize("shape,box,m,ay,az,kw", CASES[:3])
def test_matvec_fp64_vectors_fp32_transform(hipdev, shape, box, m, ay, az, kw, her):
    """Matvec mode 1: FP64 vectors in and out, FFTs and D-matrix product in FP32. Within the FP32 bar of the oracle's
    FP64 matvec; the transform part equals the FP32 matvec's to FP32 rounding; mode 0 gives back the FP64 matvec bit
    for bit; a batch of K = 3 equals the single calls bit for bit."""
    c = _case(shape, box, m, ay, az, **kw)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    try:
        op.set_cc_sqrt(ccs)
        rng = np.random.default_rng(5)
        x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
        y64, _ = op.matvec(x, her=her)
        op.set_matvec_mode(1)
        y, n2 = op.matvec(x, her=her)
        ref = c.matvec_fft(c.dmatrix(), ccs, x, her=her)[0]
        assert _rel(y, ref) < 2e-5
        assert abs(n2 - np.vdot(y, y).real) / n2 < 1e-12
        y32, _ = op.matvec(x, her=her, prec=hipdev.F32)
        assert _rel(y - x, y32 - x.astype(np.complex64)) < 2e-5
        K = 3
        op.reserve(K)
        op.set_cc_sqrt_batch([ccs] * K)  # the same coupling constants for every RHS
        xs = [x * (k + 1) + 0.5j * k for k in range(K)]
        a, r = op.alloc_batch(K), op.alloc_batch(K)
        for k in range(K):
            op.upload(op.rhs(a, k), xs[k])
        op.matvec_batch(a, r, K, her=her)
        for k in range(K):
            assert np.array_equal(op.download(op.rhs(r, k)), op.matvec(xs[k], her=her)[0]), k
        op.free(a)
        op.free(r)
        op.set_matvec_mode(0)
        assert np.array_equal(op.matvec(x, her=her)[0], y64)
    finally:
        op.close()


@pytest.mark.parametrize("her", [False, True])
def test_matvec_any_dipole_order(hipdev, her):
    """The gather of P1/P5 takes rows whose dipoles are consecutive at consecutive cells (ADDA's z, y, x order)
    arithmetically and every other row through the cell map. A spherical shell (rows through the hollow core have a
    gap) with three materials, its dipoles in ADDA's order and in a random order: both match the oracle (FP64 1e-12,
    FP32 2e-5), and the random order gives the permuted product of the ordered one bit for bit."""
    pos, box = make_shape("sphere", 20)
    ctr = (np.array(box) - 1) / 2
    pos = np.ascontiguousarray(pos[((pos - ctr) ** 2).sum(1) >= 5.0 ** 2])
    nd = len(pos)
    mat = (pos[:, 0] + pos[:, 1] // 3 + pos[:, 2] // 5) % 3
    m = [1.5 + 0.01j, 1.2 + 0.1j, 1.33]
    perm = np.random.default_rng(11).permutation(nd)
    x = np.random.default_rng(12).standard_normal(3 * nd) * (1 - 0.4j)
    xp = x.reshape(nd, 3)[perm].reshape(-1)
    ys = {}
    for name, p, mt, xv in (("ordered", pos, mat, x), ("shuffled", pos[perm], mat[perm], xp)):
        c = Case(np.ascontiguousarray(p), box, m, mat=np.ascontiguousarray(mt))
        cc, ccs = c.cc(PROP_Z, POL_Y)
        op = _op(hipdev, c, fp32=True)
        try:
            op.set_cc_sqrt(ccs)
            y64, _ = op.matvec(xv, her=her)
            y32, _ = op.matvec(xv, her=her, prec=hipdev.F32)
            ref = c.matvec_fft(c.dmatrix(), ccs, xv, her=her)[0]
            assert _rel(y64, ref) < 1e-12, name
            assert _rel(y32, ref) < 2e-5, name
            ys[name] = (y64, y32)
        finally:
            op.close()
    for i in range(2):
        assert np.array_equal(ys["shuffled"][i].reshape(nd, 3), ys["ordered"][i].reshape(nd, 3)[perm])


def test_matvec_host_vectors(hipdev):
    """adda_hip_matvec_host (ADDA's MatVec signature on host vectors; the operator keeps its device copies between
    calls) equals the device-resident matvec bit for bit, ||y||^2 included, over repeated calls and both her."""
    c = _case("ellipsoid", 16, 1.33 + 0.001j, 0.8, 0.5)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c)
    try:
        op.set_cc_sqrt(ccs)
        rng = np.random.default_rng(13)
        for her in (False, True, False):
            x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
            yh, ih = op.matvec_host(x, her=her)
            yd, idv = op.matvec(x, her=her)
            assert np.array_equal(yh, yd) and ih == idv
    finally:
        op.close()


def test_vector_kernels(hipdev):
    c = _case("sphere", 8, 1.5)
    op = _op(hipdev, c)
    n = 3 * c.nd
    rng = np.random.default_rng(4)
    a, b, cv = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3)]
    da, db, dc, do = op.alloc(), op.alloc(), op.alloc(), op.alloc()
    op.upload(da, a); op.upload(db, b); op.upload(dc, cv)
    ca, cb, cc = 0.3 - 1.1j, -2.0 + 0.5j, 0.7j
    n2 = op.lincomb(do, da, ca, db, cb, dc, cc, flags=hipdev.CONJ_B, norm=True)
    ref = ca * a + cb * np.conj(b) + cc * cv
    assert _rel(op.download(do), ref) < 1e-15
    assert abs(n2 - np.vdot(ref, ref).real) / n2 < 1e-13
    r, _ = op.dot(da, db, hipdev.DOT_CONJ)
    assert abs(r - np.sum(a * np.conj(b))) / abs(r) < 1e-13
    r, _ = op.dot(da, db, hipdev.DOT_BILIN)
    assert abs(r - np.sum(a * b)) / abs(r) < 1e-13
    r, nn = op.dot(da, None, hipdev.DOT_SELF_BILIN)
    assert abs(r - np.sum(a * a)) / abs(r) < 1e-13
    assert abs(nn - np.sum(np.abs(a) ** 2)) / nn < 1e-13


@pytest.mark.parametrize("cfg", ["x=8.4 y=8.4 z=8.4 LY=1 TC=4 LP=1", "x=4.8 y=4.4.2 z=2.4.4 LY=8 TC=32 LP=2",
                                 "x=2.2.2.4 y=16.2 z=8.2.2 LY=2 TC=8 LP=4", "x=8.4 y=8.4 z=16.2 LY=3 TC=8 LP=1",
                                 "x=16.2 y=8.4 z=8.4 LY=5 TC=16 LP=1", "x=8.4 y=16.2 z=8.4 LY=2 TC=8 LP=1 XP=0",
                                 "x=8.4 y=8.4 z=8.4 LY=4 TC=16 LP=1 XP=0 PT=1"])
def test_matvec_configs_agree(hipdev, cfg):
    """Every radix schedule / tile configuration the autotuner may pick gives the same product (grid 32^3)."""
    c = _case("sphere", 16, 1.5)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    op.set_cc_sqrt(ccs)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y0, _ = op.matvec(x)
    y032, _ = op.matvec(x, prec=hipdev.F32)
    for prec in (hipdev.F64, hipdev.F32):
        op.set_config(cfg, prec)
        assert op.get_config(prec) == cfg
    y1, _ = op.matvec(x)
    y132, _ = op.matvec(x, prec=hipdev.F32)
    assert _rel(y1, y0) < 1e-13
    assert _rel(y132, y032) < 1e-5


def test_autotune_keeps_results(hipdev, tmp_path):
    c = _case("ellipsoid", 24, 1.5 + 0.01j, 1.0, 0.5)
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    op.set_cc_sqrt(ccs)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
    y0, _ = op.matvec(x)
    cache = str(tmp_path / "tune.jsonl")
    rep = op.autotune(hipdev.F64, reps=2, cache=cache)
    assert "x=" in rep and "candidates" in rep
    rep2 = op.autotune(hipdev.F64, reps=2, cache=cache)
    assert "from cache" in rep2
    op.autotune(hipdev.F32, reps=2, cache=cache)
    y1, _ = op.matvec(x)
    assert _rel(y1, y0) < 1e-13


COMPILED = {  # length -> compile-time schedules (mv_*.hip) + one run-time schedule of the same length
    256: ["16.16", "8.8.4", "4.8.8", "16.4.4", "2.16.8"],
    512: ["16.8.4", "8.8.8", "4.8.16"],
    128: ["16.8", "8.16", "8.4.4", "4.4.8", "2.8.8"],
    64: ["16.4", "8.8", "4.16", "2.4.8"],
    32: ["8.4", "4.8", "2.16"],
    96: ["8.4.3", "16.2.3", "3.8.4", "4.8.3"],
    108: ["4.3.3.3", "3.3.3.4", "2.2.3.3.3"],
    640: ["16.8.5", "8.16.5", "5.16.8", "16.5.8", "8.5.16", "4.4.8.5"],
}


def _p3_pairs_fit(n, rmax, es):
    """Whether P3 with column pairs (12 lines of length n at LP=1) fits 512 threads and 96 KB of LDS."""
    r = n + n // 16
    ls = r + (48 - r % 32) % 32  # padded line stride, a multiple of 32 plus 16
    return 12 * (n // rmax) <= 512 and (n + 12 * ls) * es <= 96 * 1024


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("n", sorted(COMPILED))
def test_matvec_compiled_plans(hipdev, axis, n):
    """Each compile-time FFT plan of length n along each axis (P1/P5 for x, P3 for y, P2/P4 for z), and a run-time
    plan of the same length, against the oracle: FP64 <= 1e-12, FP32 <= 2e-5 (box n/2 along that axis, 3 across)."""
    from cases import read_shape
    dims = [3, 3, 3]
    dims[axis] = n // 2
    g = np.indices(dims).reshape(3, -1).T
    pos, box = read_shape(g[(g[:, 0] + 2 * g[:, 1] + 3 * g[:, 2]) % 5 != 0])  # a ragged fill of the box
    c = Case(pos, box, 1.3 + 0.02j, dpl=12)
    assert c.grid[axis] == n
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    op.set_cc_sqrt(ccs)
    x = np.random.default_rng(9).standard_normal(3 * c.nd) * (1 + 0.3j)
    y_or, _ = c.matvec_fft(c.dmatrix(), ccs, x)
    for prec in (hipdev.F64, hipdev.F32):
        # tiles that fit every schedule; P3 on single columns (XP=0) unless its column pairs fit too
        base = op.get_config(prec).split()[:3] + ["LY=2", "TC=4", "LP=1", "XP=0"]
        for sched in COMPILED[n]:
            cfg = list(base)
            cfg[axis] = "xyz"[axis] + "=" + sched
            variants = [cfg]
            radices = [int(r) for r in sched.split(".")]
            if axis == 1 and _p3_pairs_fit(n, max(radices), 8 if prec == hipdev.F32 else 16):
                variants.append(cfg[:-1])  # column pairs: x and gridX - x share the D rows
            if axis == 1 and n & (n - 1) == 0 and sched != COMPILED[n][-1] \
                    and max(radices) <= (16 if prec == hipdev.F32 else 8):
                variants.append(cfg[:-1] + ["R3=1"])  # P3 with the D-matrix product in registers
            for v in variants:
                op.set_config(" ".join(v), prec)
                assert op.get_config(prec) == " ".join(v)
                y, _ = op.matvec(x, prec=prec)
                assert _rel(y, y_or) < (1e-12 if prec == hipdev.F64 else 2e-5), (prec, v)
                if axis == 0 and sched != COMPILED[n][-1]:
                    # P1 stores / P5 loads of Xa with the nontemporal hint (compile-time plans): a cache policy, the
                    # same arithmetic -- but kernels of their own, where the compiler may pair a few multiplies and
                    # adds into fused multiply-adds differently (measured: x=16.4, one rounding step in 51 of 690
                    # elements). So: the oracle's bar, and the plain variant's values to a few rounding steps.
                    for xn in ("XN=1", "XN=2", "XN=3"):
                        vn = v + [xn]
                        op.set_config(" ".join(vn), prec)
                        assert op.get_config(prec) == " ".join(vn)
                        yn, _ = op.matvec(x, prec=prec)
                        assert _rel(yn, y_or) < (1e-12 if prec == hipdev.F64 else 2e-5), (prec, vn)
                        assert _rel(yn, y) < (1e-15 if prec == hipdev.F64 else 1e-6), (prec, vn)
                if axis == 1 and sched != COMPILED[n][-1] and "R3=1" not in v:
                    # P3 reading the staged twiddles from global memory: the same table values, so the same bits
                    vg = v + ["TG=1"]
                    op.set_config(" ".join(vg), prec)
                    assert op.get_config(prec) == " ".join(vg)
                    yg, _ = op.matvec(x, prec=prec)
                    assert np.array_equal(yg, y), (prec, vg)
                    # one thread per D-matrix element (up to 512 per block): the same products on other threads, so
                    # the same bits
                    if n > 512:
                        continue
                    vp = v + ["PT=1"]
                    op.set_config(" ".join(vp), prec)
                    assert op.get_config(prec) == " ".join(vp)
                    yp, _ = op.matvec(x, prec=prec)
                    assert np.array_equal(yp, y), (prec, vp)
                    # the lines in unpadded swizzled LDS rows: the same arithmetic at other LDS addresses, so the same
                    # bits (one column unit per block, lengths divisible by 16)
                    if n % 16 == 0:
                        vs = vp + ["SW=1"]
                        op.set_config(" ".join(vs), prec)
                        assert op.get_config(prec) == " ".join(vs)
                        ys, _ = op.matvec(x, prec=prec)
                        assert np.array_equal(ys, y), (prec, vs)


@pytest.mark.parametrize("box,grid", [((14, 22, 26), (28, 44, 52)), ((52, 28, 44), (104, 56, 88))])
def test_matvec_fftw_grids(hipdev, box, grid):
    """Grids of ADDA's FFTW fftFit rule (src/fft.c:528-562: factors of 7 and one 11 or 13 per dimension), which an
    FFTW-built ADDA hands to MatVec (104 = 8*13 is a grid of the batched workload's class): radix-7/11/13 stages in
    the run-time plans and the D-matrix build, against the oracle's FFT matvec (its transform recurses over any
    prime factor), FP64 1e-12 and FP32 2e-5, both her; the small grid also against the direct sum."""
    from cases import read_shape
    g = np.indices(box).reshape(3, -1).T
    ctr = (np.array(box) - 1) / 2
    keep = (((g - ctr) / (np.array(box) / 2)) ** 2).sum(1) <= 1.0  # an ellipsoid filling the box
    pos, b = read_shape(g[keep])
    assert tuple(b) == tuple(box)
    c = Case(pos, b, 1.33 + 0.05j, dpl=15, grid=grid)
    assert tuple(hipdev.hip().adda_hip_fftfit(2 * n, 1) for n in box) != grid  # not grids of the 2/3/5 rule
    cc, ccs = c.cc(PROP_Z, POL_Y)
    op = _op(hipdev, c, fp32=True)
    try:
        op.set_cc_sqrt(ccs)
        rng = np.random.default_rng(17)
        x = rng.standard_normal(3 * c.nd) + 1j * rng.standard_normal(3 * c.nd)
        d = c.dmatrix()
        assert _rel(op.dmatrix(), d) < 1e-12
        for her in (False, True):
            y_or, ip_or = c.matvec_fft(d, ccs, x, her=her)
            y64, ip64 = op.matvec(x, her=her)
            assert _rel(y64, y_or) < 1e-12, her
            assert abs(ip64 - ip_or) / ip_or < 1e-12
            y32, _ = op.matvec(x, her=her, prec=hipdev.F32)
            assert _rel(y32, y_or) < 2e-5, her
        if c.nd < 5000:
            y_d, _ = c.matvec_direct(ccs, x)
            assert _rel(op.matvec(x)[0], y_d) < 1e-12
    finally:
        op.close()
