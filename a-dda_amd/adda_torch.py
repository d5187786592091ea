"""torch.autograd view of the adjoint gradient, for the test harness.

`ExtAbs.apply(eps, problem, prop, incpol, opts)` maps the permittivities eps (complex128 tensor, one per material of
`problem`; one material per dipole gives a per-voxel model) to the tensor [Cext, Cabs] of a converged solve on the
GPU; its backward uses the gradient of adda_adjoint_gradient (two adjoint solves of the same operator), returned in
torch's convention for complex inputs of a real output: dL/dRe(eps) + i dL/dIm(eps). Everything runs through the C
ABI (libadda.so / libadda_hip.so); there is no CPU path.
"""
import numpy as np
import torch

import adda_amd as A


class ExtAbs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, problem, prop, incpol, opts):
        e = eps.detach().cpu().numpy().astype(np.complex128)
        problem.set_materials(np.sqrt(e))
        problem.set_polarization(prop, incpol)
        o = dict(opts or {})
        method, tol = o.pop("method", "qmr"), o.pop("eps", 1e-10)
        r = problem.solve(method, eps=tol, **o)
        if r["rc"] != 0:
            raise RuntimeError(f"forward solve did not converge: {r}")
        cext, cabs = problem.cext(), problem.cabs()
        ge, ga, _ = problem.adjoint_gradient(method, eps=tol, precision=o.get("precision", A.PREC_FP64))
        # per-dipole -> per-material: a material's permittivity moves all its dipoles
        gm = [np.bincount(problem.mat, weights=g[:, 0], minlength=problem.nmat)
              + 1j * np.bincount(problem.mat, weights=g[:, 1], minlength=problem.nmat) for g in (ge, ga)]
        ctx.save_for_backward(torch.from_numpy(gm[0]), torch.from_numpy(gm[1]))
        return torch.tensor([cext, cabs], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        ge, ga = ctx.saved_tensors
        g = gout[0].to(torch.complex128) * ge + gout[1].to(torch.complex128) * ga
        return g, None, None, None, None
