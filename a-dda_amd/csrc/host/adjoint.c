/* adjoint.c -- gradient of Cext and Cabs with respect to every dipole's permittivity by one adjoint solve.
 *
 * ADDA has no gradient; this differentiates the reference's own quantities. With S = diag(cc_sqrt) and
 * A = I + S.D.S (the matrix of MatVec, src/matvec.c:156-180), the polarization P = S.x of A.x = S.Einc solves
 * M.P = Einc with M = diag(1/cc) + D; M is complex symmetric (D = D^T, reduced-FFT symmetry of the interaction
 * matrix), so M^-T = M^-1 and every adjoint system is again a system of the same operator A:
 *   lam = M^-1 conj(Einc)   = S.y,  A.y  = S.conj(Einc)
 *   mu  = M^-1 conj(W.P)    = S.y', A.y' = S.conj(W.P)
 * (W = diag(wabs), wabs = -Im(1/cc) - 2k^3/3, the DRAINE weights of AbsCross, src/crosssec.c:873-884). Both are
 * solved together as a batch of two right-hand sides. With q = d(1/cc)/d(eps) per material (cc = CoupleConstant,
 * src/calculator.c:100-133, at eps = m^2):
 *   Cext = 4 pi k Im sum conj(Einc).P  (ExtCross for a plane wave, src/crosssec.c:811-823, by the optical theorem)
 *     dCext = 4 pi k Im(h d eps_j),  h = -sum_c q lam_jc P_jc
 *   Cabs = 4 pi k sum wabs |P|^2       (src/crosssec.c:852-884)
 *     dCabs = 4 pi k [-Im(q d eps_j) |P_j|^2 + 2 Re(z d eps_j)],  z = -sum_c q mu_jc P_jc
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/adda.h"
#include "host_internal.h"

typedef double complex dcx;

#define FOUR_PI 12.566370614359172953850573533118
/* LDR coefficients, src/const.h:234-236 */
#define LDR_B1 1.8915316
#define LDR_B2 -0.1648469
#define LDR_B3 1.7700004

adda_cplx adda_couple_constant_deps(adda_cplx mo, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3])
{
	dcx m = mo.re + I * mo.im, e = m * m, res;
	const double V = gridspace * gridspace * gridspace, kd = wavenum * gridspace;
	const dcx a = (3 / FOUR_PI) * V * (e - 1) / (e + 2), da = (3 / FOUR_PI) * V * 3 / ((e + 2) * (e + 2));
	if (polrel == ADDA_POL_CM) res = da;
	else {
		dcx M = I * 2 * kd * kd * kd / 3, dM = 0;
		if (polrel == ADDA_POL_LDR) {
			double S = 0;
			for (int i = 0; i < 3; i++) S += prop[i] * prop[i] * incpol[i] * incpol[i];
			M += (LDR_B1 + (LDR_B2 + LDR_B3 * S) * e) * kd * kd;
			dM = (LDR_B2 + LDR_B3 * S) * kd * kd;
		}
		const dcx den = 1 - a * M / V;
		res = (da + a * a * dM / V) / (den * den);
	}
	adda_cplx r = {creal(res), cimag(res)};
	return r;
}

int adda_adjoint_gradient(adda_hip_op *op, const adda_particle *p, const adda_solver_opts *opt, const adda_cplx *m,
	int polrel, double wavenum, double gridspace, const double prop[3], const double incpol[3], const void *d_einc,
	const void *d_pvec, double *grad_ext, double *grad_abs, adda_solve_stats *st)
{
	if (!op || !p || !opt || !m || !d_einc || !d_pvec || !grad_ext || !grad_abs || !st)
		return host_err(__FILE__, __LINE__, "null argument");
	const int nmat = p->nmat, K = 2;
	const size_t nq = (size_t)nmat * 3, stride = adda_hip_batch_stride(op);
	int rc = 0;
	adda_cplx *ccs = malloc(2 * nq * sizeof(adda_cplx)), *q = malloc(nq * sizeof(adda_cplx));
	double *wabs = malloc(nq * sizeof(double));
	void *e = NULL, *x = NULL, *pa = NULL;
	if (!ccs || !q || !wabs) { rc = host_err(__FILE__, __LINE__, "out of host memory"); goto out; }
	const double k3 = 2 * wavenum * wavenum * wavenum / 3;
	for (int i = 0; i < nmat; i++) {
		adda_cplx c = adda_couple_constant(m[i], polrel, wavenum, gridspace, prop, incpol);
		adda_cplx dc = adda_couple_constant_deps(m[i], polrel, wavenum, gridspace, prop, incpol);
		const dcx cc = c.re + I * c.im, dcc = dc.re + I * dc.im, s = csqrt(cc), qq = -dcc / (cc * cc);
		for (int comp = 0; comp < 3; comp++) {
			const size_t t = 3 * (size_t)i + comp;
			ccs[t].re = ccs[nq + t].re = creal(s);
			ccs[t].im = ccs[nq + t].im = cimag(s);
			q[t].re = creal(qq);
			q[t].im = cimag(qq);
			wabs[t] = -cimag(1 / cc) - k3;
		}
	}
	/* the two adjoint systems share the forward operator and its cc_sqrt: one batch of K = 2 right-hand sides */
	if (adda_hip_set_cc_sqrt_batch(op, K, ccs)) { rc = host_err_hip(__FILE__, __LINE__); goto out; }
	e = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
	x = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
	pa = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
	if (!e || !x || !pa) { rc = host_err_hip(__FILE__, __LINE__); goto out; }
	if (adda_hip_weighted_conj(op, e, d_einc, NULL) ||
		adda_hip_weighted_conj(op, (char *)e + stride * sizeof(adda_cplx), d_pvec, wabs)) {
		rc = host_err_hip(__FILE__, __LINE__);
		goto out;
	}
	if (!opt->host_scalars && (opt->method == ADDA_IT_QMR_CS || opt->method == ADDA_IT_BICGSTAB ||
		opt->method == ADDA_IT_CGNR))
		rc = adda_iterative_solver_batch(op, opt, K, e, x, pa, NULL, 0, st);
	else /* host-scalar recurrences: one right-hand side at a time (both use cc_sqrt table 0) */
		for (int k = 0; k < K && rc >= 0; k++) {
			const size_t off = k * stride * sizeof(adda_cplx);
			int r = adda_iterative_solver(op, opt, (char *)e + off, (char *)x + off, (char *)pa + off, NULL, 0, &st[k]);
			if (r < 0 || (r > 0 && rc == 0)) rc = r;
		}
	if (rc < 0) goto out;
	if (adda_hip_adjoint_grad(op, d_pvec, pa, (char *)pa + stride * sizeof(adda_cplx), q, FOUR_PI * wavenum,
		FOUR_PI * wavenum, grad_ext, grad_abs)) {
		rc = host_err_hip(__FILE__, __LINE__);
		goto out;
	}
out:
	if (e) adda_hip_vec_free(op, e);
	if (x) adda_hip_vec_free(op, x);
	if (pa) adda_hip_vec_free(op, pa);
	free(ccs);
	free(q);
	free(wabs);
	return rc;
}
