"""GPU parity of the adjoint gradient d(Cext, Cabs)/d(eps_j) (two adjoint solves of the same operator, batched).

Bars: against the oracle's adjoint gradient (itself pinned by central differences of its forward solves and of the
reference's outputs) to 1e-8 of the largest component; against central differences of the reference's own Cext/Cabs
on the 15-material variant to 1e-4 plus the resolution of its 10-digit output; against batched GPU central
differences (step 1e-6) on 64 seeded voxels of the 48^3 per-voxel cube to 1e-4, for the FP64 and the FP32-mixed
solves; the torch.autograd view returns the same gradient in torch's complex convention.
"""
import numpy as np
import pytest

import golden_io as G
from cases import Case, adj8_materials, make_shape, per_voxel_m, read_shape, PROP_Z, POL_X, POL_Y
from gpu_fd import batched_fd

pytestmark = pytest.mark.gpu


def _per_voxel(A, box, fp32=False, dpl=18):
    pos, _ = make_shape("box", box)
    p, b = read_shape(pos)
    nd = len(p)
    m = per_voxel_m(nd)
    pr = A.Problem(p, b, m, mat=np.arange(nd, dtype=np.int32), dpl=dpl, fp32=fp32)
    return pr, p, b, m


@pytest.mark.parametrize("pol", ["Y", "X"])
def test_adjoint_vs_oracle(hipdev, pol):
    pr, p, b, m = _per_voxel(hipdev, 6)
    vec = POL_Y if pol == "Y" else POL_X
    try:
        pr.set_polarization(PROP_Z, vec)
        pr.solve("qmr", eps=1e-12)
        ge, ga, st = pr.adjoint_gradient("qmr", eps=1e-12)
        assert all(s["rc"] == 0 for s in st)
        oc = Case(p, b, m, dpl=18, mat=np.arange(len(p), dtype=np.int32), rule=1)
        oe, oa, cext, cabs = oc.adjoint_grad(oc.dmatrix(), PROP_Z, vec)
        assert abs(pr.cext() - cext) < 1e-9 * cext and abs(pr.cabs() - cabs) < 1e-9 * cabs
        assert np.abs(ge - oe).max() < 1e-8 * np.abs(oe).max()
        assert np.abs(ga - oa).max() < 1e-8 * np.abs(oa).max()
    finally:
        pr.close()


@pytest.mark.parametrize("pol", ["Y", "X"])
def test_adjoint_vs_reference_differences(hipdev, pol):
    """15-material box 8^3: per-voxel gradient converted to d/dRe(m), d/dIm(m) vs the reference's central
    differences (voxel moved alone to material 15)."""
    pos, _ = G.geometry("adj8")
    p, b = read_shape(pos)
    m, mat, vox = adj8_materials(len(p))
    pr = hipdev.Problem(p, b, m[mat], mat=np.arange(len(p), dtype=np.int32), dpl=18)
    vec = POL_Y if pol == "Y" else POL_X
    try:
        pr.set_polarization(PROP_Z, vec)
        pr.solve("qmr", eps=1e-12)
        base = G.crosssec("adj8/adj8_base", pol)
        assert abs(pr.cext() - base["Cext"]) < 1e-8 * base["Cext"]
        ge, ga, _ = pr.adjoint_gradient("qmr", eps=1e-12)
        fd = G.adj8_fd(pol)
        for j in vox:
            for row, g in ((0, ge[j]), (1, ga[j])):
                dre, dim = G.deps_to_dm(g, m[mat[j]])
                ref = fd[int(j)][row]
                tol = G.adj8_fd_resolution(base["Cext" if row == 0 else "Cabs"])
                assert abs(dre - ref[0]) < 1e-4 * abs(ref[0]) + tol and abs(dim - ref[1]) < 1e-4 * abs(ref[1]) + tol, \
                    (j, row, dre, dim, ref)
    finally:
        pr.close()


def test_torch_autograd_view(hipdev):
    import torch
    from adda_torch import ExtAbs
    pr, p, b, m = _per_voxel(hipdev, 5)
    try:
        eps = torch.tensor(m ** 2, dtype=torch.complex128, requires_grad=True)
        out = ExtAbs.apply(eps, pr, PROP_Z, POL_Y, {"method": "bicgstab", "eps": 1e-12})
        (g,) = torch.autograd.grad(out[0] + 2 * out[1], eps)
        ge, ga, _ = pr.adjoint_gradient("bicgstab", eps=1e-12)
        ref = (ge[:, 0] + 2 * ga[:, 0]) + 1j * (ge[:, 1] + 2 * ga[:, 1])
        assert np.abs(g.numpy() - ref).max() < 1e-12 * np.abs(ref).max()
        # torch's convention for a real output of a complex input: dL/dRe + i dL/dIm -- checked on central
        # differences of the oracle's forward solves
        oc = Case(p, b, m, dpl=18, mat=np.arange(len(p), dtype=np.int32), rule=1)
        d = oc.dmatrix()
        for j in (3, 77):
            fd = oc.fd_grad(d, PROP_Z, POL_Y, j)
            want = (fd[0, 0] + 2 * fd[1, 0]) + 1j * (fd[0, 1] + 2 * fd[1, 1])
            assert abs(g[j].item() - want) < 1e-6 * abs(want), (j, g[j].item(), want)
    finally:
        pr.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_cube48_vs_central_differences(hipdev, precision):
    """The per-voxel 48^3 cube (110,592 voxels, m_re ~ U(1.2,1.8), m_im ~ U(0,0.05), seed 20261204): gradient of the
    FP64 or FP32-mixed forward + adjoint solves vs batched FP64 central differences, step 1e-6, on 64 voxels (seed
    20261205), 1e-4 relative to the largest gradient component of those voxels."""
    pr, p, b, m = _per_voxel(hipdev, 48, fp32=True)
    try:
        pr.set_polarization(PROP_Z, POL_Y)
        r = pr.solve("qmr", eps=1e-10, precision=precision)
        assert r["rc"] == 0
        ge, ga, st = pr.adjoint_gradient("qmr", eps=1e-10, precision=precision)
        assert all(s["rc"] == 0 for s in st)
        vox = np.random.default_rng(20261205).choice(pr.nd, 64, replace=False)
        if precision == 1:  # differences are always taken around the FP64 solution
            pr.solve("qmr", eps=1e-12)
        fd = batched_fd(hipdev, pr, [int(j) for j in vox], h=1e-6, eps=1e-12)
        for i, j in enumerate(vox):
            for row, g in ((0, ge[j]), (1, ga[j])):
                scale = np.abs(fd[:, row, :]).max()
                assert np.abs(g - fd[i, row]).max() < 1e-4 * scale, (j, row, g, fd[i, row])
    finally:
        pr.close()
