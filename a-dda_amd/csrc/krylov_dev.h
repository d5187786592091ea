// krylov_dev.h -- device-resident scalar recurrences of ADDA's Krylov solvers (src/iterative.c).
//
// The FP64 scalar state of one solve lives in device memory (slots below). Each iteration is a short chain of kernels
// on the operator's stream: fused vector updates that also emit per-block partial reductions, and one-block "scalar"
// kernels that finish those reductions in a fixed order and advance the recurrence exactly as the reference does on
// the host (QMR_CS 898-1032, BiCGStab 640-724, CGNR 730-770, ProgressReport 282-318). Nothing returns to the host
// inside an iteration; the host polls the status slot every few iterations. Once the status is non-zero every kernel
// of the chain returns immediately, so iterations enqueued past convergence cost only their launch.
#pragma once
#include "fft_device.h"

namespace adda {

// state slots (doubles)
enum : int {
	DK_INPRODR = 0,  // best ||r||^2 so far (inprodR)
	DK_INPRODRP1,    // ||r_k||^2 of the last iteration
	DK_EPSB,         // eps^2 ||b||^2
	DK_RSCALE,       // 1/||b||^2
	DK_NITER,        // niter (next iteration number, starts at 1)
	DK_COUNTER,      // iterations without residual decrease
	DK_STATUS,       // 0 running, 1 converged, 2 maxiter, -1 breakdown, -2 stagnation
	DK_MAXITER,
	DK_MC,           // params[].mc
	DK_HISTLEN,
	DK_CODE,         // which breakdown test failed (diagnostics)
	DK_SC = 16,      // method scalars start here
	DK_NSLOTS = 96
};

struct dcx {
	double x, y;
};
__device__ __forceinline__ dcx dmk(double a, double b) { dcx r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ dcx dadd(dcx a, dcx b) { return dmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ dcx dsub(dcx a, dcx b) { return dmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ dcx dmul(dcx a, dcx b) { return dmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ dcx dscl(dcx a, double s) { return dmk(a.x * s, a.y * s); }
__device__ __forceinline__ dcx dconj(dcx a) { return dmk(a.x, -a.y); }
__device__ __forceinline__ dcx dneg(dcx a) { return dmk(-a.x, -a.y); }
__device__ __forceinline__ double dabs(dcx a) { return hypot(a.x, a.y); }
__device__ __forceinline__ double dabs2(dcx a) { return a.x * a.x + a.y * a.y; }
// a/b, scaled to avoid overflow (Smith's algorithm)
__device__ __forceinline__ dcx ddiv(dcx a, dcx b)
{
	if (fabs(b.x) >= fabs(b.y)) {
		const double r = b.y / b.x, d = b.x + b.y * r;
		return dmk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
	}
	const double r = b.x / b.y, d = b.x * r + b.y;
	return dmk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
// principal square root, branch cut along the negative real axis (C99 csqrt)
__device__ __forceinline__ dcx dsqrt(dcx z)
{
	const double r = hypot(z.x, z.y);
	if (r == 0) return dmk(0, z.y);
	if (z.x >= 0) {
		const double t = sqrt(0.5 * (r + z.x));
		return dmk(t, z.y / (2 * t));
	}
	const double t = sqrt(0.5 * (r - z.x));
	return dmk(fabs(z.y) / (2 * t), copysign(t, z.y));
}
__device__ __forceinline__ dcx ld_c(const double *s, int i) { return dmk(s[i], s[i + 1]); }
__device__ __forceinline__ void st_c(double *s, int i, dcx v) { s[i] = v.x; s[i + 1] = v.y; }

}  // namespace adda
