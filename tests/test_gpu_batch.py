"""GPU parity of the batched multi-RHS form: K right-hand sides (orientations x polarizations) sharing one operator.

Bars: every batched kernel gives, for each RHS, bit-for-bit the result of the single-RHS call (same kernels, same
per-RHS data, same reduction order); batched FP64 solves equal the separate single-RHS solves bit for bit, with and
without HIP-graph replay; cross sections of the batched FP64 solve match the reference's SPARSE outputs for each
orientation and polarization (abs 1e-14 / rel 1e-8); the FP32-mixed batched solve reaches true residual <= 1e-10 and
Cext within 1e-6 of the FP64 values.
"""
import numpy as np
import pytest

import golden_io as G
from adda_rot import rotate_zyz
from cases import Case, orientations, read_shape, PROP_Z, POL_X, POL_Y

pytestmark = pytest.mark.gpu


def _rhs(n_orient):
    props, pols, names = [], [], []
    for io, (a, b, g) in enumerate(orientations(n_orient)):
        prop, pY, pX = (rotate_zyz(a, b, g, v) for v in (PROP_Z, POL_Y, POL_X))
        for pol, vec in (("Y", pY), ("X", pX)):
            props.append(prop)
            pols.append(vec)
            names.append(("ell12_c3o%d" % io, pol))
    return np.array(props), np.array(pols), names


def _batch(adda, n_orient, fp32=False):
    pos, _ = G.geometry("ell12")
    p, box = read_shape(pos)
    props, pols, names = _rhs(n_orient)
    return adda.BatchProblem(p, box, 1.33 + 0.001j, props, pols, dpl=13.3, fp32=fp32), names


@pytest.mark.parametrize("prec", [0, 1])
def test_batched_kernels_equal_single(hipdev, prec):
    """matvec (both her forms), lincomb and dot over K = 5 RHS with distinct cc tables vs one call per RHS."""
    bp, _ = _batch(hipdev, 3, fp32=True)
    op, K = bp.op, 5
    try:
        rng = np.random.default_rng(7)
        n = 3 * bp.nd
        xs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(K)]
        a, r = op.alloc_batch(K, prec), op.alloc_batch(K, prec)
        b2 = op.alloc_batch(K, prec)
        for k in range(K):
            op.upload(op.rhs(a, k, prec), xs[k], prec)
            op.upload(op.rhs(b2, k, prec), xs[(k + 1) % K], prec)
        base = op.get_config(prec)
        # P3 with 1, 2 or 4 consecutive RHS per block (KL; K = 5 leaves a ragged last group)
        for cfg in (base, base + " KL=2", base + " KL=4"):
            op.set_config(cfg, prec)
            for her in (False, True):
                ip = op.matvec_batch(a, r, K, prec, her=her, inprod=True)
                yb = [op.download(op.rhs(r, k, prec), prec) for k in range(K)]
                for k in range(K):
                    op.set_cc_sqrt(bp.cc_sqrt[k])
                    y1, ip1 = op.matvec(xs[k], her=her, prec=prec)
                    assert np.array_equal(y1, yb[k]), (cfg, her, k, np.abs(y1 - yb[k]).max())
                    assert ip1 == ip[k]
                op.set_cc_sqrt_batch(bp.cc_sqrt)
        op.set_config(base, prec)
        ca = rng.standard_normal(K) + 1j * rng.standard_normal(K)
        n2 = op.lincomb_batch(K, r, a, ca, b2, 0.5 - 0.25j, None, 0, prec=prec, norm=True)
        d, dn = op.dot_batch(K, a, b2, hipdev.DOT_BILIN, prec=prec)
        for k in range(K):
            o1 = op.alloc(prec)
            e1 = op.lincomb(o1, op.rhs(a, k, prec), ca[k], op.rhs(b2, k, prec), 0.5 - 0.25j, prec=prec, norm=True)
            assert np.array_equal(op.download(o1, prec), op.download(op.rhs(r, k, prec), prec)) and e1 == n2[k]
            op.free(o1)
            assert op.dot(op.rhs(a, k, prec), op.rhs(b2, k, prec), hipdev.DOT_BILIN, prec=prec)[0] == d[k]
        for v in (a, r, b2):
            op.free(v)
    finally:
        bp.close()


@pytest.mark.parametrize("method", ["qmr", "bicgstab", "cgnr", "bicg", "qmr2", "csym", "bcgs2"])
def test_batched_fp64_equals_single_and_reference(hipdev, method):
    """8 RHS (4 orientations x 2 polarizations): x of every RHS equals the single-RHS solve bit for bit, graph replay
    equals eager enqueueing, and Cext/Cabs match the reference's outputs."""
    bp, names = _batch(hipdev, 4)
    try:
        res = bp.solve(method, eps=1e-10, graph=False)
        x_eager = [bp.field_x(k) for k in range(bp.K)]
        res_g = bp.solve(method, eps=1e-10, graph=True)
        for k, (gold, pol) in enumerate(names):
            assert res[k]["rc"] == 0 and res_g[k]["niter"] == res[k]["niter"]
            assert np.array_equal(bp.field_x(k), x_eager[k]), k
            g = G.crosssec(gold, pol)
            assert G.numeric_close([bp.cext(k), bp.cabs(k)], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0], (k, bp.cext(k), g)
        # single-RHS solves of two of them on a fresh operator
        pos, _ = G.geometry("ell12")
        p, box = read_shape(pos)
        pr = hipdev.Problem(p, box, 1.33 + 0.001j, dpl=13.3)
        try:
            for k in (0, 5):
                pr.set_polarization(bp.props[k], bp.pols[k])
                r1 = pr.solve(method, eps=1e-10)
                assert r1["niter"] == res[k]["niter"] and r1["nmatvec"] == res[k]["nmatvec"]
                assert np.array_equal(pr.field_x(), x_eager[k]), k
        finally:
            pr.close()
    finally:
        bp.close()


def test_batched_history_and_oracle_field(hipdev):
    """Per-RHS residual histories equal the single solves'; one RHS's field vs the oracle's FP64 solve (1e-8)."""
    bp, _ = _batch(hipdev, 2)
    pos, _ = G.geometry("ell12")
    p, box = read_shape(pos)
    try:
        res = bp.solve("qmr", eps=1e-10, hist_len=400)
        oc = Case(p, box, 1.33 + 0.001j, dpl=13.3, rule=1)
        k = 3
        cc, ccs = oc.cc(bp.props[k], bp.pols[k])
        ro = oc.solve(oc.dmatrix(), ccs, oc.einc(bp.props[k], bp.pols[k]), method=0, eps=1e-10)
        x = bp.field_x(k)
        assert np.abs(x - ro["x"]).max() / np.abs(ro["x"]).max() < 1e-8
        pr = hipdev.Problem(p, box, 1.33 + 0.001j, dpl=13.3)
        try:
            for k in range(bp.K):
                pr.set_polarization(bp.props[k], bp.pols[k])
                r1 = pr.solve("qmr", eps=1e-10, hist_len=400)
                assert np.array_equal(r1["hist"], res[k]["hist"]), k
        finally:
            pr.close()
    finally:
        bp.close()


@pytest.mark.parametrize("v64", [False, True])
@pytest.mark.parametrize("method", ["bicgstab", "qmr", "cgnr", "bcgs2", "csym"])
def test_batched_mixed_precision(hipdev, method, v64):
    """Inner solves of all RHS in shared launches (FP32, or FP64 around FP32 transforms) + per-RHS FP64 defect
    correction."""
    bp, names = _batch(hipdev, 4, fp32=True)
    try:
        ref = bp.solve(method, eps=1e-10)
        c64 = [bp.cext(k) for k in range(bp.K)]
        res = bp.solve(method, eps=1e-10, precision=hipdev.PREC_MIXED_V64 if v64 else hipdev.PREC_MIXED)
        for k in range(bp.K):
            assert res[k]["true_resid"] <= 1e-10, (k, res[k])
            assert res[k]["nouter"] >= 2
            assert abs(bp.cext(k) - c64[k]) / abs(c64[k]) < 1e-6
            g = G.crosssec(*names[k])
            assert abs(bp.cext(k) - g["Cext"]) / g["Cext"] < 1e-6
        assert all(r["rc"] == 0 for r in ref)
    finally:
        bp.close()


def test_batch_of_one_and_gating(hipdev):
    """A batched matvec gated by a solver state leaves the RHS whose status is non-zero untouched and computes the
    others exactly as the single-RHS matvec does."""
    import ctypes
    bp, _ = _batch(hipdev, 1)
    op = bp.op
    try:
        K = 2
        st = hipdev.hip().adda_hip_dk_create_batch(op.h, K, 2)
        assert st
        vals = np.array([1.0])
        slot = 1 * 96 + 6  # RHS 1, status slot (ADDA_DK_NSLOTS = 96, ADDA_DK_STATUS = 6)
        hipdev.check(hipdev.hip().adda_hip_dk_set(op.h, st, slot, vals.ctypes.data_as(ctypes.c_void_p), 1))
        a, r = op.alloc_batch(K), op.alloc_batch(K)
        x = np.arange(3 * bp.nd) * (1 + 0.5j) / bp.nd
        for k in range(K):
            op.upload(op.rhs(a, k), x)
        op.lincomb_batch(K, r, None, 0, None, 0, None, 0)  # zero
        sentinel = np.full(3 * bp.nd, 7 - 3j)
        op.upload(op.rhs(r, 1), sentinel)
        y1, _ = op.matvec(x)
        base = op.get_config(hipdev.F64)
        # with KL=2 the gated RHS shares its P3 blocks with the running one
        for cfg in (base, base + " KL=2"):
            op.set_config(cfg, hipdev.F64)
            op.upload(op.rhs(r, 1), sentinel)
            op.matvec_batch(a, r, K, gate=st)
            assert np.array_equal(op.download(op.rhs(r, 1)), sentinel), cfg
            assert np.array_equal(op.download(op.rhs(r, 0)), y1), cfg
        hipdev.check(hipdev.hip().adda_hip_dk_free(op.h, st))
        op.free(a)
        op.free(r)
    finally:
        bp.close()


def test_launch_list(hipdev):
    """Batched matvecs under a launch list compute the listed RHS exactly as the full launch does (every P3 form the
    tuner picks between: one RHS per block and KL=2), leave the others untouched, and a gate still applies to the
    listed ones; clearing the list restores the full launch, and a list a caller left set does not reach the batched
    solver."""
    import ctypes
    bp, _ = _batch(hipdev, 3)
    op = bp.op
    K = bp.K
    try:
        a, r = op.alloc_batch(K), op.alloc_batch(K)
        rng = np.random.default_rng(5)
        for k in range(K):
            op.upload(op.rhs(a, k), rng.standard_normal(3 * bp.nd) + 1j * rng.standard_normal(3 * bp.nd))
        op.matvec_batch(a, r, K)
        full = [op.download(op.rhs(r, k)) for k in range(K)]
        sentinel = np.full(3 * bp.nd, 7 - 3j)
        run = np.zeros(K, dtype=np.int32)
        run[[0, 3, 4]] = 1
        st = hipdev.hip().adda_hip_dk_create_batch(op.h, K, 2)
        assert st
        vals = np.array([1.0])
        hipdev.check(hipdev.hip().adda_hip_dk_set(op.h, st, 3 * 96 + 6, vals.ctypes.data_as(ctypes.c_void_p), 1))
        base = op.get_config(hipdev.F64)
        try:
            for cfg in (base, base + " KL=2"):
                op.set_config(cfg, hipdev.F64)
                for gate in (None, st):
                    for k in range(K):
                        op.upload(op.rhs(r, k), sentinel)
                    op.set_launch_list(K, run)
                    op.matvec_batch(a, r, K, gate=gate)
                    op.set_launch_list(K, None)
                    for k in range(K):
                        got = op.download(op.rhs(r, k))
                        listed = run[k] and not (gate is not None and k == 3)
                        assert np.array_equal(got, full[k] if listed else sentinel), (cfg, gate is not None, k)
            op.set_config(base, hipdev.F64)
            op.matvec_batch(a, r, K)
            for k in range(K):
                assert np.array_equal(op.download(op.rhs(r, k)), full[k]), k
            # a list left set by a caller does not reach the solver (it sets its own and clears it)
            ref = bp.solve("bicgstab", eps=1e-10)
            x_ref = [bp.field_x(k) for k in range(K)]
            op.set_launch_list(K, run)
            res = bp.solve("bicgstab", eps=1e-10)
            for k in range(K):
                assert res[k]["rc"] == 0 and res[k]["niter"] == ref[k]["niter"], k
                assert np.array_equal(bp.field_x(k), x_ref[k]), k
            op.matvec_batch(a, r, K)
            for k in range(K):
                assert np.array_equal(op.download(op.rhs(r, k)), full[k]), k
        finally:
            hipdev.check(hipdev.hip().adda_hip_dk_free(op.h, st))
            op.free(a)
            op.free(r)
    finally:
        bp.close()


@pytest.mark.parametrize("method", ["bicg", "qmr"])
def test_batched_breakdown_is_per_rhs(hipdev, method):
    """One RHS of a batch breaks down (incident field (1, i, 0) at every dipole: r^T r = 0 at iteration 1 of the
    complex-symmetric methods); it alone reports the error, and the other RHS converge to the same bits as in a batch
    without the breakdown."""
    bp, _ = _batch(hipdev, 1)
    try:
        assert bp.K == 2
        ref = bp.solve(method, eps=1e-10, init_field=hipdev.IF_ZERO)
        x1 = bp.field_x(1)
        e = np.zeros((bp.nd, 3), dtype=np.complex128)
        e[:, 0] = 1
        e[:, 1] = 1j
        bp.op.upload(bp.op.rhs(bp.einc, 0), e.ravel())
        res = bp.solve(method, eps=1e-10, init_field=hipdev.IF_ZERO, raise_on_error=False)
        assert res[0]["rc"] < 0
        assert res[1]["rc"] == 0 and res[1]["niter"] == ref[1]["niter"]
        assert np.array_equal(bp.field_x(1), x1)
    finally:
        bp.close()


def test_batched_postprocessing(hipdev):
    """Cext, Cabs and the scattering-plane fields of every RHS from one batched evaluation equal the single-RHS
    functions bit for bit; for the orientation the reference ran (30, 40, 50; m = 1.33+0.05i, dpl 12) the batched
    values reproduce its CrossSec and mueller outputs."""
    pos, _ = G.geometry("ell12")
    p, box = read_shape(pos)
    props, pols, pars, pers = [], [], [], []
    for a, b, g in [(30, 40, 50)] + [tuple(o) for o in orientations(2)]:
        prop, pY, pX = (rotate_zyz(a, b, g, v) for v in (PROP_Z, POL_Y, POL_X))
        for vec in (pY, pX):
            props.append(prop)
            pols.append(vec)
            pars.append(pY)
            pers.append(pX)
    bp = hipdev.BatchProblem(p, box, 1.33 + 0.05j, np.array(props), np.array(pols), dpl=12)
    try:
        res = bp.solve("qmr", eps=1e-10)
        assert all(r["rc"] == 0 for r in res)
        ce, ca = bp.cext_all(), bp.cabs_all()
        ep = bp.eplane_all(pars, pers, 20, 38)
        for k in range(bp.K):
            assert ce[k] == bp.cext(k) and ca[k] == bp.cabs(k), k
            assert np.array_equal(ep[k], bp.eplane(k, pars[k], pers[k], 20, 38)), k
        for j, pol in ((0, "Y"), (1, "X")):
            g = G.crosssec("ell12_orient", pol)
            assert G.numeric_close([ce[j], ca[j]], [g["Cext"], g["Cabs"]], 1e-14, 1e-8)[0], (pol, ce[j], ca[j], g)
        mu = hipdev.mueller_yz(ep[1], ep[0], 38)
        gm = G.mueller("ell12_orient")
        assert G.mueller_close(mu, gm[:, 1:], gm[:, 1])[0]
    finally:
        bp.close()
