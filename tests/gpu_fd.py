"""Central finite differences of Cext and Cabs in one dipole's permittivity from forward solves on the GPU, batched:
every (voxel, direction, sign) is one right-hand side with its own cc_sqrt table.

C(eps_j +- h) - C(eps_j) is evaluated as an exact difference: with M' = M + dM (dM non-zero on dipole j only),
P' - P solves M'(P' - P) = -dM.P, a forward solve of the perturbed system whose right-hand side is local to dipole j;
the differences of Cext and Cabs are then sums of O(h) terms only, so a step of 1e-6 loses no digits to cancellation.
Used by the GPU tests of the adjoint gradient; needs a model with one material per dipole.
"""
import ctypes

import numpy as np


def batched_fd(A, pr, voxels, h=1e-6, method="qmr", eps=1e-12):
    """pr: solved FP64 Problem (pr.pvec = P) with mat[j] == j. Returns fd[len(voxels)][2][2]:
    [[dCext/dRe eps, dCext/dIm eps], [dCabs/dRe eps, dCabs/dIm eps]] per voxel."""
    op, nd = pr.op, pr.nd
    assert pr.nmat == nd and np.array_equal(pr.mat, np.arange(nd))
    P = pr.polarization()
    e = op.download(pr.einc)
    k = pr.wavenum
    k3 = 2 * k ** 3 / 3
    cc = pr.cc[:, 0]
    w = -np.imag(1 / cc) - k3
    W = np.repeat(w, 3)
    runs = [(j, col, sgn) for j in voxels for col in (0, 1) for sgn in (1, -1)]
    K = len(runs)
    op.reserve(K)
    tables = np.repeat(pr.cc_sqrt[None], K, axis=0)
    E = op.alloc_batch(K)
    X = op.alloc_batch(K)
    DP = op.alloc_batch(K)
    meta = []
    try:
        rhs = np.zeros(3 * nd, np.complex128)
        for r, (j, col, sgn) in enumerate(runs):
            m2 = np.sqrt(pr.m[j] ** 2 + sgn * (h if col == 0 else 1j * h))
            c2 = A.couple_constant(m2, pr.polrel, pr.wavenum, pr.gridspace, pr.prop, pr.incpol)
            tables[r, j, :] = np.sqrt(c2)
            rhs[:] = 0
            rhs[3 * j:3 * j + 3] = -(1 / c2 - 1 / cc[j]) * P[3 * j:3 * j + 3]
            op.upload(op.rhs(E, r), rhs)
            meta.append((j, col, sgn, -np.imag(1 / c2) - k3))
        op.set_cc_sqrt_batch(tables)
        o = A.SolverOpts(A.ITER_NAMES[method], eps, int(3 * np.prod(pr.box)), A.IF_AUTO, 0, A.PREC_FP64, 0.0, 0, 0, 0, 0,
                         1)
        st = (A.SolveStats * K)()
        A.hcheck(A.host().adda_iterative_solver_batch(op.h, ctypes.byref(o), K, E, X, DP, None, 0, st))
        assert all(s.rc == 0 for s in st), [s.rc for s in st]
        d = {}
        for r, (j, col, sgn, w2) in enumerate(meta):
            dP = op.download(op.rhs(DP, r))
            dext = 4 * np.pi * k * np.imag(np.sum(np.conj(e) * dP))
            dabs = 4 * np.pi * k * (np.sum(W * (2 * np.real(np.conj(P) * dP) + np.abs(dP) ** 2))
                                     + (w2 - w[j]) * np.sum(np.abs(P[3 * j:3 * j + 3] + dP[3 * j:3 * j + 3]) ** 2))
            d[(j, col, sgn)] = (dext, dabs)
        out = np.zeros((len(voxels), 2, 2))
        for i, j in enumerate(voxels):
            for col in (0, 1):
                for row in (0, 1):
                    out[i, row, col] = (d[(j, col, 1)][row] - d[(j, col, -1)][row]) / (2 * h)
        return out
    finally:
        for v in (E, X, DP):
            op.free(v)
        op.set_cc_sqrt(pr.cc_sqrt)
