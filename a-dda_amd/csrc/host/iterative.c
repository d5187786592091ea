/* iterative.c -- ADDA's Krylov driver (src/iterative.c:1398-1530) over the MI355X backend.
 *
 * The scalar recurrences stay on the host in FP64 complex, exactly as in the reference; every vector operation and
 * every inner product runs on the device through include/adda_hip.h (one fused kernel per linalg call, FP64
 * accumulation in every reduction). Vectors of the iteration live in the working precision `prec`: FP64 for the
 * plain solver, FP32 for the inner solves of the mixed-precision form, whose outer loop recomputes the true residual
 * r = b - A.x in FP64 (ResidualNorm2, src/iterative.c:322-338) and accumulates x += e in FP64 (defect correction).
 */
#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/adda.h"
#include "host_internal.h"

typedef double complex dcx;

static adda_cplx AC(dcx z) { adda_cplx r = {creal(z), cimag(z)}; return r; }
static dcx CA(adda_cplx z) { return z.re + I * z.im; }

/* solver context: operator, working precision, matvec counter */
typedef struct {
	adda_hip_op *op;
	int prec;
	int nmv;
} kctx;

#define CK(x) do { if ((x) != 0) return host_err_hip(__FILE__, __LINE__); } while (0)

static int mv(kctx *k, const void *x, void *y, double *inprod, int her)
{
	k->nmv++;
	return adda_hip_matvec(k->op, k->prec, x, y, inprod, her);
}
/* a = ca*a + cb*b + cc*c style wrapper around the fused lincomb kernel */
static int lc(kctx *k, void *out, const void *a, dcx ca, const void *b, dcx cb, const void *c, dcx cc, int fl,
	double *n2)
{
	return adda_hip_lincomb(k->op, k->prec, out, a, AC(ca), b, AC(cb), c, AC(cc), fl, n2);
}
static int dotp(kctx *k, const void *a, const void *b, int mode, dcx *res, double *n2)
{
	adda_cplx r;
	int rc = adda_hip_dot(k->op, k->prec, a, b, mode, &r, n2);
	if (res) *res = CA(r);
	return rc;
}

static void *kvec(kctx *k) { return adda_hip_vec_alloc(k->op, k->prec); }

/* One Krylov solve of A.x = b in precision k->prec (IterativeSolver body, src/iterative.c:1398-1530).
 * b: right-hand side (already S.Einc); x: in/out (initial field per init_field); on return *niter_out iterations.
 * resid_scale/epsB follow src/iterative.c:1423-1425. The per-method bodies restate src/iterative.c:640-1032. */
static int krylov(kctx *k, int method, double eps, int maxiter, int init_field, int verbose, const void *b, void *x,
	double *hist, int hist_len, int *niter_out, double *final_rel)
{
	const int mc_tab[5] = {50000, 30000, 10, 50000, 10}; /* params[].mc, src/iterative.c:118-132 */
	int mc = mc_tab[method], rc = 0;
	void *r = kvec(k), *Av = kvec(k), *p = kvec(k), *v1 = kvec(k), *v2 = kvec(k), *v3 = kvec(k);
	if (!r || !Av || !p || !v1 || !v2 || !v3) { rc = host_err_hip(__FILE__, __LINE__); goto done; }
	double temp;
#define CKG(x) do { if ((x) != 0) { rc = host_err_hip(__FILE__, __LINE__); goto done; } } while (0)
	CKG(dotp(k, b, NULL, ADDA_DOT_NORM2, NULL, &temp));
	if (!(temp > 0)) { rc = host_err(__FILE__, __LINE__, "right-hand side is zero"); goto done; }
	double resid_scale = 1 / temp, epsB = eps * eps * temp, inprodR = 0, inprodRp1 = 0;
	int matvec_ready = 0;
	/* CalcInitField, src/iterative.c:1342-1394 */
	if (init_field == ADDA_IF_AUTO) {
		CKG(mv(k, b, Av, NULL, 0));
		CKG(lc(k, r, b, 1, Av, -1, NULL, 0, 0, &inprodR));
		if (temp < inprodR) {
			CKG(lc(k, x, NULL, 0, NULL, 0, NULL, 0, 0, NULL));
			CKG(lc(k, r, b, 1, NULL, 0, NULL, 0, 0, NULL));
			inprodR = temp;
			matvec_ready = 1;
			if (verbose) printf("x_0 = 0\n");
		}
		else {
			CKG(lc(k, x, b, 1, NULL, 0, NULL, 0, 0, NULL));
			if (verbose) printf("x_0 = E_inc\n");
		}
	}
	else if (init_field == ADDA_IF_ZERO) {
		CKG(lc(k, x, NULL, 0, NULL, 0, NULL, 0, 0, NULL));
		CKG(lc(k, r, b, 1, NULL, 0, NULL, 0, 0, NULL));
		inprodR = temp;
	}
	else {
		CKG(lc(k, x, b, 1, NULL, 0, NULL, 0, 0, NULL));
		CKG(mv(k, x, Av, NULL, 0));
		CKG(lc(k, r, b, 1, Av, -1, NULL, 0, 0, &inprodR));
	}
	double prev_err = sqrt(resid_scale * inprodR);
	if (hist && hist_len > 0) hist[0] = prev_err;
	if (verbose) printf("RE_000 = %.10E\n", prev_err);
	int niter = 1, counter = 0;

	double c_old = 0, c_new = 0, omega_old = 0, omega_new = 0, ro_old_d = 0, beta_d = 0;
	dcx beta = 0, tautilda = 0, s_new = 0, s_old = 0, ro_old = 0, omega = 0, alpha = 0, tau = 0;
	void *v = v1, *vtilda = v2, *p_new = p, *p_old = v3;    /* QMR_CS renames, src/iterative.c:916-921 */
	void *q_new = r, *q_old = v1, *pc_new = p, *pc_old = v2; /* CSYM renames, src/iterative.c:795-799 */
	/* PHASE_INIT */
	if (method == ADDA_IT_QMR_CS) {
		dcx t;
		CKG(dotp(k, r, NULL, ADDA_DOT_SELF_BILIN, &t, NULL));
		beta = csqrt(t);
		omega_old = 0.0;
		omega_new = sqrt(inprodR) / cabs(beta);
		CKG(lc(k, v, r, 1 / beta, NULL, 0, NULL, 0, 0, NULL));
		tautilda = omega_new * beta;
		c_new = c_old = 1.0;
		s_new = s_old = 0.0;
	}
	else if (method == ADDA_IT_BICGSTAB) CKG(lc(k, v3, r, 1, NULL, 0, NULL, 0, 0, NULL)); /* r~ = r_0 */
	else if (method == ADDA_IT_CSYM) {
		tau = sqrt(inprodR);
		CKG(lc(k, q_new, q_new, 1 / creal(tau), NULL, 0, NULL, 0, ADDA_CONJ_A, NULL));
		c_new = 1;
		c_old = 0;
		s_new = s_old = 0;
	}
	while (inprodR > epsB && niter <= maxiter && counter <= mc) {
		if (method == ADDA_IT_QMR_CS) { /* src/iterative.c:962-1030 */
			double dtmp1 = 1 / (omega_new * omega_new), dtmp2, zetaabs;
			dcx theta, eta, zetatilda, zeta, temp1, temp2, temp4;
			if (dtmp1 < 1e-10) { rc = host_err(__FILE__, __LINE__, "QMR_CS fails: |vT.v|/(v.v) is too small (%g)", dtmp1); break; }
			if (niter == 1 && matvec_ready) CKG(lc(k, Av, Av, 1 / beta, NULL, 0, NULL, 0, 0, NULL));
			else CKG(mv(k, v, Av, NULL, 0));
			CKG(dotp(k, v, Av, ADDA_DOT_BILIN, &alpha, NULL));
			temp2 = -alpha;
			if (niter == 1) CKG(lc(k, vtilda, v, temp2, Av, 1, NULL, 0, 0, NULL));
			else CKG(lc(k, vtilda, vtilda, -beta, v, temp2, Av, 1, 0, NULL));
			theta = conj(s_old) * omega_old * beta;
			eta = c_old * c_new * omega_old * beta;
			eta += alpha * conj(s_new) * omega_new;
			zetatilda = c_new * omega_new * alpha - s_new * c_old * omega_old * beta;
			omega_old = omega_new;
			CKG(dotp(k, vtilda, NULL, ADDA_DOT_SELF_BILIN, &temp1, &dtmp1));
			beta = csqrt(temp1);
			omega_new = sqrt(dtmp1) / cabs(beta);
			dtmp2 = creal(zetatilda) * creal(zetatilda) + cimag(zetatilda) * cimag(zetatilda);
			zetaabs = sqrt(dtmp2 + dtmp1);
			dtmp1 = sqrt(dtmp2);
			if (dtmp1 < 1e-40) zeta = zetaabs;
			else zeta = (zetaabs / dtmp1) * zetatilda;
			c_old = c_new;
			c_new = dtmp1 / zetaabs;
			s_old = s_new;
			s_new = omega_new * beta / zeta;
			temp4 = 1 / zeta;
			if (niter == 1) CKG(lc(k, p_new, v, temp4, NULL, 0, NULL, 0, 0, NULL));
			else {
				temp2 = -eta * temp4;
				if (niter == 2) CKG(lc(k, p_old, p_new, temp2, v, temp4, NULL, 0, 0, NULL));
				else CKG(lc(k, p_old, p_old, -theta * temp4, p_new, temp2, v, temp4, 0, NULL));
				void *t = p_old; p_old = p_new; p_new = t;
			}
			tau = c_new * tautilda;
			tautilda = -s_new * tautilda;
			CKG(lc(k, x, x, 1, p_new, tau, NULL, 0, 0, NULL));
			CKG(lc(k, vtilda, vtilda, 1 / beta, NULL, 0, NULL, 0, 0, NULL));
			{ void *t = v; v = vtilda; vtilda = t; }
			temp1 = (c_new / omega_new) * tautilda;
			double cs = creal(s_new) * creal(s_new) + cimag(s_new) * cimag(s_new);
			CKG(lc(k, r, r, cs, v, temp1, NULL, 0, 0, &inprodRp1));
		}
		else if (method == ADDA_IT_BICGSTAB) { /* src/iterative.c:675-718; v=vec1 s=vec2 rtilda=vec3 */
			void *vv = v1, *sv = v2, *rt = v3;
			dcx ro_new, temp1, temp2;
			CKG(dotp(k, r, rt, ADDA_DOT_CONJ, &ro_new, NULL));
			if (niter == 1) CKG(lc(k, p, r, 1, NULL, 0, NULL, 0, 0, NULL));
			else {
				temp1 = ro_new * alpha;
				temp2 = ro_old * omega;
				if (cabs(temp2) / cabs(temp1) < 1e-10) { rc = host_err(__FILE__, __LINE__, "BiCGStab fails: 1/|beta| is too small"); break; }
				beta = temp1 / temp2;
				CKG(lc(k, p, p, beta, vv, -beta * omega, r, 1, 0, NULL));
			}
			if (niter == 1 && matvec_ready) CKG(lc(k, vv, Av, 1, NULL, 0, NULL, 0, 0, NULL));
			else CKG(mv(k, p, vv, NULL, 0));
			CKG(dotp(k, vv, rt, ADDA_DOT_CONJ, &temp1, NULL));
			if (cabs(temp1) / cabs(ro_new) < 1e-10) { rc = host_err(__FILE__, __LINE__, "BiCGStab fails: |v.r~|/|r.r~| is too small"); break; }
			alpha = ro_new / temp1;
			CKG(lc(k, sv, vv, -alpha, r, 1, NULL, 0, 0, &inprodRp1));
			if (inprodRp1 < epsB) CKG(lc(k, x, x, 1, p, alpha, NULL, 0, 0, NULL));
			else {
				double denumOmega;
				CKG(mv(k, sv, Av, &denumOmega, 0));
				dcx st;
				CKG(dotp(k, sv, Av, ADDA_DOT_CONJ, &st, NULL));
				omega = st / denumOmega;
				CKG(lc(k, x, x, 1, p, alpha, sv, omega, 0, NULL));
				CKG(lc(k, r, Av, -omega, sv, 1, NULL, 0, 0, &inprodRp1));
				ro_old = ro_new;
			}
		}
		else if (method == ADDA_IT_CGNR) { /* src/iterative.c:745-766 */
			double ro_new, denum, a;
			if (niter == 1) CKG(mv(k, r, p, &ro_new, 1));
			else {
				CKG(mv(k, r, Av, &ro_new, 1));
				CKG(lc(k, p, p, ro_new / ro_old_d, Av, 1, NULL, 0, 0, NULL));
			}
			CKG(mv(k, p, Av, &denum, 0));
			a = ro_new / denum;
			CKG(lc(k, x, x, 1, p, a, NULL, 0, 0, NULL));
			CKG(lc(k, r, r, 1, Av, -a, NULL, 0, 0, &inprodRp1));
			ro_old_d = ro_new;
		}
		else if (method == ADDA_IT_BICG_CS) { /* src/iterative.c:549-630 */
			dcx ro_new, mu;
			CKG(dotp(k, r, NULL, ADDA_DOT_SELF_BILIN, &ro_new, NULL));
			double abs_ro_new = cabs(ro_new);
			if (abs_ro_new / inprodR < 1e-10) { rc = host_err(__FILE__, __LINE__, "BiCG_CS fails: |rT.r|/(r.r) is too small"); break; }
			if (niter == 1) CKG(lc(k, p, r, 1, NULL, 0, NULL, 0, 0, NULL));
			else CKG(lc(k, p, p, ro_new / ro_old, r, 1, NULL, 0, 0, NULL));
			if (!(niter == 1 && matvec_ready)) CKG(mv(k, p, Av, NULL, 0));
			CKG(dotp(k, p, Av, ADDA_DOT_BILIN, &mu, NULL));
			if (cabs(mu) / abs_ro_new < 1e-10) { rc = host_err(__FILE__, __LINE__, "BiCG_CS fails: |pT.A.p|/(rT.r) is too small"); break; }
			alpha = ro_new / mu;
			CKG(lc(k, x, x, 1, p, alpha, NULL, 0, 0, NULL));
			CKG(lc(k, r, r, 1, Av, -alpha, NULL, 0, 0, &inprodRp1));
			ro_old = ro_new;
		}
		else { /* CSYM, src/iterative.c:822-888 */
			dcx eta, gamma, theta, invksi, temp1;
			double dtmp, sum;
			CKG(mv(k, q_new, Av, NULL, 0));
			CKG(dotp(k, q_new, Av, ADDA_DOT_BILIN, &alpha, NULL));
			eta = c_old * c_new * beta_d + alpha * conj(s_new);
			gamma = c_new * alpha - c_old * s_new * beta_d;
			theta = beta_d * conj(s_old);
			temp1 = -alpha;
			if (niter == 1) CKG(lc(k, q_old, q_new, temp1, Av, 1, NULL, 0, ADDA_CONJ_A, &sum));
			else CKG(lc(k, q_old, q_old, -beta_d, q_new, temp1, Av, 1, ADDA_CONJ_A | ADDA_CONJ_B, &sum));
			beta_d = sqrt(sum);
			c_old = c_new;
			s_old = s_new;
			dtmp = cabs(gamma);
			if (dtmp == 0) {
				if (beta_d == 0) { rc = host_err(__FILE__, __LINE__, "CSYM: interaction matrix is singular"); break; }
				c_new = 0;
				s_new = 1;
				invksi = 1 / beta_d;
			}
			else {
				if (dtmp < beta_d) { dtmp = dtmp / beta_d; c_new = dtmp / sqrt(1 + dtmp * dtmp); }
				else { dtmp = beta_d / dtmp; c_new = 1 / sqrt(1 + dtmp * dtmp); }
				invksi = c_new / gamma;
				s_new = beta_d * invksi;
			}
			if (niter == 1) CKG(lc(k, pc_new, q_new, invksi, NULL, 0, NULL, 0, 0, NULL));
			else {
				temp1 = -eta * invksi;
				if (niter == 2) CKG(lc(k, pc_old, pc_new, temp1, q_new, invksi, NULL, 0, 0, NULL));
				else CKG(lc(k, pc_old, pc_old, -theta * invksi, pc_new, temp1, q_new, invksi, 0, NULL));
				void *t = pc_old; pc_old = pc_new; pc_new = t;
			}
			CKG(lc(k, x, x, 1, pc_new, c_new * tau, NULL, 0, 0, NULL));
			CKG(lc(k, q_old, q_old, 1 / beta_d, NULL, 0, NULL, 0, ADDA_CONJ_A, NULL));
			{ void *t = q_old; q_old = q_new; q_new = t; }
			tau *= -s_new;
			inprodRp1 = creal(tau) * creal(tau) + cimag(tau) * cimag(tau);
		}
		/* ProgressReport, src/iterative.c:282-318 */
		if (inprodRp1 <= inprodR) { inprodR = inprodRp1; counter = 0; }
		else counter++;
		double err = sqrt(resid_scale * inprodRp1);
		if (verbose) printf("RE_%03d = %.10E  %s\n", niter, err, counter == 0 ? "+ " : (1 - err / prev_err > 0 ? "-+" : "- "));
		prev_err = err;
		if (hist && niter < hist_len) hist[niter] = err;
		niter++;
	}
	if (rc == 0 && inprodR > epsB) {
		if (niter > maxiter) rc = 1; /* warning only, src/iterative.c:1506-1508 */
		else if (counter > mc) rc = host_err(__FILE__, __LINE__, "Residual norm haven't decreased for maximum allowed number of iterations (%d)", mc) - 1;
	}
	*niter_out = niter - 1;
	*final_rel = sqrt(resid_scale * inprodR);
done:
	{
		void *vs[] = {r, Av, p, v1, v2, v3};
		for (int i = 0; i < 6; i++)
			if (vs[i]) adda_hip_vec_free(k->op, vs[i]);
	}
	return rc;
#undef CKG
}

/* Device-resident form of krylov() for QMR_CS, BiCGStab and CGNR: CalcInitField and PHASE_INIT as above (host,
 * synchronising, once), then whole iterations enqueued with adda_hip_dk_*_iter -- whose FP64 scalar recurrences run
 * on the GPU -- in chunks of DK_CHUNK, polling the device status slot after each chunk. Iterations enqueued after
 * convergence are no-ops on the device. Results are identical in exact arithmetic to krylov() (same recurrences, same
 * order of operations per vector element; reductions differ only in summation order). */
#define DK_CHUNK 16
static int krylov_dev(kctx *k, int method, double eps, int maxiter, int init_field, int verbose, const void *b, void *x,
	double *hist, int hist_len, int *niter_out, double *final_rel)
{
	const int mc_tab[5] = {50000, 30000, 10, 50000, 10};
	int rc = 0;
	void *r = kvec(k), *Av = kvec(k), *p = kvec(k), *v1 = kvec(k), *v2 = kvec(k), *v3 = kvec(k);
	void *st = adda_hip_dk_create(k->op, hist_len > 1 ? hist_len : 2);
	if (!r || !Av || !p || !v1 || !v2 || !v3 || !st) { rc = host_err_hip(__FILE__, __LINE__); goto done; }
	double temp;
#define CKG(x) do { if ((x) != 0) { rc = host_err_hip(__FILE__, __LINE__); goto done; } } while (0)
	CKG(dotp(k, b, NULL, ADDA_DOT_NORM2, NULL, &temp));
	if (!(temp > 0)) { rc = host_err(__FILE__, __LINE__, "right-hand side is zero"); goto done; }
	double resid_scale = 1 / temp, epsB = eps * eps * temp, inprodR = 0;
	int matvec_ready = 0;
	if (init_field == ADDA_IF_AUTO) { /* src/iterative.c:1347-1368 */
		CKG(mv(k, b, Av, NULL, 0));
		CKG(lc(k, r, b, 1, Av, -1, NULL, 0, 0, &inprodR));
		if (temp < inprodR) {
			CKG(lc(k, x, NULL, 0, NULL, 0, NULL, 0, 0, NULL));
			CKG(lc(k, r, b, 1, NULL, 0, NULL, 0, 0, NULL));
			inprodR = temp;
			matvec_ready = 1;
		}
		else CKG(lc(k, x, b, 1, NULL, 0, NULL, 0, 0, NULL));
	}
	else if (init_field == ADDA_IF_ZERO) {
		CKG(lc(k, x, NULL, 0, NULL, 0, NULL, 0, 0, NULL));
		CKG(lc(k, r, b, 1, NULL, 0, NULL, 0, 0, NULL));
		inprodR = temp;
	}
	else {
		CKG(lc(k, x, b, 1, NULL, 0, NULL, 0, 0, NULL));
		CKG(mv(k, x, Av, NULL, 0));
		CKG(lc(k, r, b, 1, Av, -1, NULL, 0, 0, &inprodR));
	}
	if (hist && hist_len > 0) hist[0] = sqrt(resid_scale * inprodR);
	if (verbose) printf("RE_000 = %.10E\n", sqrt(resid_scale * inprodR));
	{
		double base[11] = {inprodR, 0, epsB, resid_scale, 1, 0, inprodR > epsB ? 0 : 1, maxiter, mc_tab[method],
			hist_len > 1 ? hist_len : 2, 0};
		if (maxiter < 1 && base[6] == 0) base[6] = 2;
		CKG(adda_hip_dk_set(k->op, st, 0, base, 11));
	}
	/* PHASE_INIT, written into the method slots (layout: krylov_dev.h / adda_hip.hip k_dk_* kernels) */
	if (method == ADDA_IT_QMR_CS) {
		dcx t;
		CKG(dotp(k, r, NULL, ADDA_DOT_SELF_BILIN, &t, NULL));
		dcx beta = csqrt(t);
		double omega_new = sqrt(inprodR) / cabs(beta);
		CKG(lc(k, v1, r, 1 / beta, NULL, 0, NULL, 0, 0, NULL));
		dcx tautilda = omega_new * beta, ib = 1 / beta;
		double sc[46] = {0};
		sc[0] = creal(beta); sc[1] = cimag(beta);
		sc[8] = creal(tautilda); sc[9] = cimag(tautilda);
		sc[10] = 1; sc[11] = 1; sc[12] = 0; sc[13] = omega_new;
		sc[44] = creal(ib); sc[45] = cimag(ib);
		CKG(adda_hip_dk_set(k->op, st, ADDA_DK_SC, sc, 46));
	}
	else if (method == ADDA_IT_BICGSTAB) {
		CKG(lc(k, v3, r, 1, NULL, 0, NULL, 0, 0, NULL)); /* r~ = r_0 */
		double sc[26] = {0};
		sc[24] = 1;
		CKG(adda_hip_dk_set(k->op, st, ADDA_DK_SC, sc, 26));
	}
	else {
		double sc[14] = {0};
		sc[6] = 1;
		CKG(adda_hip_dk_set(k->op, st, ADDA_DK_SC, sc, 14));
	}
	{
		void *v = v1, *vtilda = v2, *p_new = p, *p_old = v3;
		int it = 1;
		double status = 0;
		while (status == 0) {
			for (int c = 0; c < DK_CHUNK; c++, it++) {
				if (method == ADDA_IT_QMR_CS) {
					CKG(adda_hip_dk_qmr_iter(k->op, k->prec, st, it, matvec_ready, v, vtilda, Av, p_old, p_new, x, r));
					if (it >= 2) { void *t = p_old; p_old = p_new; p_new = t; }
					{ void *t = v; v = vtilda; vtilda = t; }
					k->nmv += !(it == 1 && matvec_ready);
				}
				else if (method == ADDA_IT_BICGSTAB) {
					/* v = vec1, s = vec2, rtilda = vec3, t = Avecbuffer (src/iterative.c:660-662) */
					CKG(adda_hip_dk_bicgstab_iter(k->op, k->prec, st, it, matvec_ready, x, r, p, v1, v2, v3, Av));
					k->nmv += 2 - (it == 1 && matvec_ready);
				}
				else {
					CKG(adda_hip_dk_cgnr_iter(k->op, k->prec, st, it, x, r, p, Av));
					k->nmv += 2;
				}
			}
			CKG(adda_hip_dk_get(k->op, st, ADDA_DK_STATUS, &status, 1));
		}
		double fin[11];
		CKG(adda_hip_dk_get(k->op, st, 0, fin, 11));
		int niter = (int)fin[4];
		/* matvecs of the enqueued iterations that the device skipped are not counted */
		int skipped = it - niter;
		if (method == ADDA_IT_QMR_CS) k->nmv -= skipped;
		else k->nmv -= 2 * skipped;
		inprodR = fin[0];
		if (hist && hist_len > 1 && niter > 1) {
			int n = niter < hist_len ? niter : hist_len;
			double *tmp = malloc(n * sizeof(double));
			if (!tmp) { rc = host_err(__FILE__, __LINE__, "out of host memory"); goto done; }
			CKG(adda_hip_dk_history(k->op, st, tmp, n));
			for (int i = 1; i < n; i++) hist[i] = tmp[i];
			free(tmp);
		}
		if (verbose)
			for (int i = 1; i < niter && hist && i < hist_len; i++) printf("RE_%03d = %.10E\n", i, hist[i]);
		*niter_out = niter - 1;
		*final_rel = sqrt(resid_scale * inprodR);
		if (status == 2) rc = 1;
		else if (status == -1) rc = host_err(__FILE__, __LINE__, "%s breakdown (test %d failed)",
			method == ADDA_IT_QMR_CS ? "QMR_CS" : method == ADDA_IT_BICGSTAB ? "BiCGStab" : "CGNR", (int)fin[10]);
		else if (status == -2) rc = host_err(__FILE__, __LINE__,
			"Residual norm haven't decreased for maximum allowed number of iterations (%d)", mc_tab[method]) - 1;
	}
done:
	{
		void *vs[] = {r, Av, p, v1, v2, v3};
		for (int i = 0; i < 6; i++)
			if (vs[i]) adda_hip_vec_free(k->op, vs[i]);
		if (st) adda_hip_dk_free(k->op, st);
	}
	return rc;
#undef CKG
}

/* dispatch: device-resident recurrences where they exist, unless the caller asks for the host form */
static int krylov_any(kctx *k, int method, int host_scalars, double eps, int maxiter, int init_field, int verbose,
	const void *b, void *x, double *hist, int hist_len, int *niter_out, double *final_rel)
{
	if (!host_scalars && (method == ADDA_IT_QMR_CS || method == ADDA_IT_BICGSTAB || method == ADDA_IT_CGNR))
		return krylov_dev(k, method, eps, maxiter, init_field, verbose, b, x, hist, hist_len, niter_out, final_rel);
	return krylov(k, method, eps, maxiter, init_field, verbose, b, x, hist, hist_len, niter_out, final_rel);
}

int adda_iterative_solver(adda_hip_op *op, const adda_solver_opts *opt, const void *d_einc, void *d_x, void *d_pvec,
	double *resid_hist, int hist_len, adda_solve_stats *st)
{
	if (!op || !opt || !d_einc || !d_x || !d_pvec || !st) return host_err(__FILE__, __LINE__, "null argument");
	memset(st, 0, sizeof *st);
	st->true_resid = -1;
	if (opt->method < 0 || opt->method > 4) return host_err(__FILE__, __LINE__, "unknown iterative method %d", opt->method);
	int rc = 0;
	/* b = S.Einc into pvec (nMult_mat, src/iterative.c:1421-1422) */
	CK(adda_hip_mult_mat(op, ADDA_F64, d_pvec, d_einc));
	if (opt->precision == ADDA_PREC_FP64) {
		kctx k = {op, ADDA_F64, 0};
		double rel = 0;
		rc = krylov_any(&k, opt->method, opt->host_scalars, opt->eps, opt->maxiter, opt->init_field, opt->verbose,
			d_pvec, d_x, resid_hist, hist_len, &st->niter, &rel);
		st->nmatvec = k.nmv;
		st->resid = rel;
		if (rc < 0) { st->rc = rc; return rc; }
		st->rc = rc;
		if (opt->recalc_resid) { /* ResidualNorm2, src/iterative.c:1512-1520 */
			void *Ax = adda_hip_vec_alloc(op, ADDA_F64), *r = adda_hip_vec_alloc(op, ADDA_F64);
			double b2, r2;
			if (!Ax || !r) return host_err_hip(__FILE__, __LINE__);
			CK(adda_hip_matvec(op, ADDA_F64, d_x, Ax, NULL, 0));
			CK(adda_hip_lincomb(op, ADDA_F64, r, d_pvec, AC(1), Ax, AC(-1), NULL, AC(0), 0, &r2));
			CK(adda_hip_dot(op, ADDA_F64, d_pvec, NULL, ADDA_DOT_NORM2, NULL, &b2));
			st->true_resid = sqrt(r2 / b2);
			st->nmatvec_fp64 = 1;
			adda_hip_vec_free(op, Ax);
			adda_hip_vec_free(op, r);
		}
	}
	else {
		/* FP64 defect correction around FP32 Krylov solves:
		 *   x = 0; r = b; repeat { e ~ A^-1 r (FP32, relative tol inner_eps); x += e; r = b - A.x (FP64) }
		 * until ||r||/||b|| <= eps. Inner solves start from e = 0 (init_field zero) on the normalized residual. */
		const double inner = opt->inner_eps > 0 ? opt->inner_eps : 1e-5;
		const int max_outer = opt->max_outer > 0 ? opt->max_outer : 30;
		void *r64 = adda_hip_vec_alloc(op, ADDA_F64), *Ax = adda_hip_vec_alloc(op, ADDA_F64);
		void *r32 = adda_hip_vec_alloc(op, ADDA_F32), *e32 = adda_hip_vec_alloc(op, ADDA_F32);
		void *e64 = adda_hip_vec_alloc(op, ADDA_F64);
		if (!r64 || !Ax || !r32 || !e32 || !e64) return host_err_hip(__FILE__, __LINE__);
		double b2, r2;
		CK(adda_hip_dot(op, ADDA_F64, d_pvec, NULL, ADDA_DOT_NORM2, NULL, &b2));
		CK(adda_hip_lincomb(op, ADDA_F64, d_x, NULL, AC(0), NULL, AC(0), NULL, AC(0), 0, NULL));
		CK(adda_hip_lincomb(op, ADDA_F64, r64, d_pvec, AC(1), NULL, AC(0), NULL, AC(0), 0, &r2));
		int hist_pos = 0;
		if (resid_hist && hist_len > 0) resid_hist[hist_pos++] = sqrt(r2 / b2);
		kctx k = {op, ADDA_F32, 0};
		while (sqrt(r2 / b2) > opt->eps && st->nouter < max_outer) {
			const double rn = sqrt(r2);
			/* normalized residual in FP32 */
			CK(adda_hip_lincomb(op, ADDA_F64, Ax, r64, AC(1 / rn), NULL, AC(0), NULL, AC(0), 0, NULL));
			CK(adda_hip_vec_convert(op, r32, ADDA_F32, Ax, ADDA_F64));
			/* relative target of this step: reach eps overall but no deeper than FP32 can resolve */
			double tgt = opt->eps / sqrt(r2 / b2);
			if (tgt < inner) tgt = inner;
			int ni = 0;
			double rel;
			/* the inner history starts one slot early: its entry 0 (= 1, the normalized start) overwrites the last
			 * entry of the previous step, which holds the same residual */
			int hp = hist_pos > 0 ? hist_pos - 1 : 0;
			int cap = opt->max_inner > 0 && opt->max_inner < opt->maxiter ? opt->max_inner : opt->maxiter;
			int irc = krylov_any(&k, opt->method, opt->host_scalars, tgt, cap, ADDA_IF_ZERO, 0, r32, e32,
				resid_hist ? resid_hist + hp : NULL, resid_hist ? hist_len - hp : 0, &ni, &rel);
			if (resid_hist) hist_pos = hp + 1;
			if (irc < 0) {
				/* an FP32 breakdown or stagnation ends this step; the FP64 residual below decides what happens next */
				if (ni == 0) { st->rc = irc; goto mixed_done; }
			}
			if (resid_hist) { /* inner residuals relative to this step's rhs -> relative to b */
				const double scale = sqrt(r2 / b2);
				int upto = hist_pos + ni < hist_len ? hist_pos + ni : hist_len;
				for (int h = hist_pos; h < upto; h++) resid_hist[h] *= scale;
				hist_pos = upto;
			}
			st->niter += ni;
			/* x += rn * e (FP64 accumulation) */
			CK(adda_hip_vec_convert(op, e64, ADDA_F64, e32, ADDA_F32));
			CK(adda_hip_lincomb(op, ADDA_F64, d_x, d_x, AC(1), e64, AC(rn), NULL, AC(0), 0, NULL));
			/* true residual in FP64 */
			CK(adda_hip_matvec(op, ADDA_F64, d_x, Ax, NULL, 0));
			st->nmatvec_fp64++;
			double r2new;
			CK(adda_hip_lincomb(op, ADDA_F64, r64, d_pvec, AC(1), Ax, AC(-1), NULL, AC(0), 0, &r2new));
			st->nouter++;
			if (opt->verbose) printf("outer %d: inner iterations %d, true residual %.6E\n", st->nouter, ni, sqrt(r2new / b2));
			if (!(r2new < r2)) { /* no progress in FP64: give up with the best x so far */
				r2 = r2new;
				break;
			}
			r2 = r2new;
		}
		st->nmatvec = k.nmv;
		st->resid = st->true_resid = sqrt(r2 / b2);
		if (st->resid > opt->eps && st->rc == 0) st->rc = 1;
	mixed_done:
		adda_hip_vec_free(op, r64);
		adda_hip_vec_free(op, Ax);
		adda_hip_vec_free(op, r32);
		adda_hip_vec_free(op, e32);
		adda_hip_vec_free(op, e64);
		if (st->rc < 0) return st->rc;
		rc = st->rc;
	}
	/* P = S.x (src/iterative.c:1527) */
	CK(adda_hip_mult_mat(op, ADDA_F64, d_pvec, d_x));
	CK(adda_hip_sync(op));
	return rc;
}
