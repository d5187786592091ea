/* iterative.c -- ADDA's Krylov driver (src/iterative.c:1398-1530) over the MI355X backend.
 *
 * The scalar recurrences stay on the host in FP64 complex, exactly as in the reference; every vector operation and
 * every inner product runs on the device through include/adda_hip.h (one fused kernel per linalg call, FP64
 * accumulation in every reduction). Vectors of the iteration live in the working precision `prec`: FP64 for the
 * plain solver, FP32 for the inner solves of the mixed-precision form, whose outer loop recomputes the true residual
 * r = b - A.x in FP64 (ResidualNorm2, src/iterative.c:322-338) and accumulates x += e in FP64 (defect correction).
 */
#define _POSIX_C_SOURCE 200809L /* clock_gettime */
#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

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
	int chp_saved; /* an iteration checkpoint was written */
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

/* Moves a checkpoint's vectors (xvec, rvec, pvec and the method's own, dv[0..]) between the device and the file
 * (adda_chp_write / adda_chp_read, SaveIterChpoint / LoadIterChpoint of src/iterative.c:165-278); FP64 only. */
static int chp_move(kctx *k, const adda_solver_opts *o, int method, int save, adda_chp_head *h, double *sc,
	void *const *dv)
{
	int nv, rc = 0;
	const size_t nrows = adda_hip_nrows(k->op);
	const char *dir = o->chp_dir ? o->chp_dir : "chpoint";
	if (adda_chp_layout(method, NULL, NULL, &nv)) return -1;
	adda_cplx *hv[7] = {0};
	for (int i = 0; i < 3 + nv && rc == 0; i++)
		if (!(hv[i] = malloc(nrows * sizeof(adda_cplx)))) rc = host_err(__FILE__, __LINE__, "out of host memory");
		else if (save && adda_hip_vec_download(k->op, ADDA_F64, hv[i], dv[i])) rc = host_err_hip(__FILE__, __LINE__);
	if (rc == 0) {
		if (save) rc = adda_chp_write(dir, o->run_dir, method, nrows, h, sc, (const adda_cplx *const *)hv);
		else if ((rc = adda_chp_read(dir, method, nrows, h, sc, hv)) == 0)
			for (int i = 0; i < 3 + nv && rc == 0; i++)
				if (adda_hip_vec_upload(k->op, ADDA_F64, dv[i], hv[i])) rc = host_err_hip(__FILE__, __LINE__);
	}
	for (int i = 0; i < 7; i++) free(hv[i]);
	return rc;
}

static double wall_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Start of the checkpoint clock: the time the library was loaded, i.e. program start for a program linked against
 * it. The reference starts last_chp_wt at program start (src/timing.c:102) and never resets it per solve, so set-up
 * and earlier solves count towards -chpoint. */
static double g_chp_clock0;
__attribute__((constructor)) static void chp_clock_init(void) { g_chp_clock0 = wall_s(); }

/* One Krylov solve of A.x = b in precision k->prec (IterativeSolver body, src/iterative.c:1398-1530).
 * b: right-hand side (already S.Einc); x: in/out (initial field per init_field); on return *niter_out iterations.
 * resid_scale/epsB follow src/iterative.c:1423-1425. The per-method bodies restate src/iterative.c:342-1143
 * (BCGS2, BiCG_CS, BiCGStab, CGNR, CSYM, QMR_CS, QMR_CS_2). */
static int krylov(kctx *k, int method, double eps, int maxiter, int init_field, int verbose, const void *b, void *x,
	double *hist, int hist_len, int *niter_out, double *final_rel, const adda_solver_opts *chp)
{
	const int mc_tab[7] = {50000, 30000, 10, 50000, 10, 15000, 50000}; /* params[].mc, src/iterative.c:118-132 */
	int mc = mc_tab[method], rc = 0;
	void *r = kvec(k), *Av = kvec(k), *p = kvec(k), *v1 = kvec(k), *v2 = kvec(k), *v3 = kvec(k);
	void *v4 = method == ADDA_IT_BCGS2 ? kvec(k) : NULL;
	if (!r || !Av || !p || !v1 || !v2 || !v3 || (method == ADDA_IT_BCGS2 && !v4)) {
		rc = host_err_hip(__FILE__, __LINE__);
		goto done;
	}
	double temp;
#define CKG(x) do { if ((x) != 0) { rc = host_err_hip(__FILE__, __LINE__); goto done; } } while (0)
	CKG(dotp(k, b, NULL, ADDA_DOT_NORM2, NULL, &temp));
	if (!(temp > 0)) { rc = host_err(__FILE__, __LINE__, "right-hand side is zero"); goto done; }
	double resid_scale = 1 / temp, epsB = eps * eps * temp, inprodR = 0, inprodRp1 = 0;
	int matvec_ready = 0;
	const int load = chp && chp->chp_load;
	/* CalcInitField, src/iterative.c:1342-1394 (skipped when resuming from a checkpoint, src/iterative.c:1421) */
	if (load) {}
	else if (init_field == ADDA_IF_AUTO) {
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
	if (hist && hist_len > 0 && !load) hist[0] = prev_err;
	if (verbose && !load) printf("RE_000 = %.10E\n", prev_err);
	int niter = 1, counter = 0, chp_exit = 0;

	double c_old = 0, c_new = 0, omega_old = 0, omega_new = 0, ro_old_d = 0, beta_d = 0;
	dcx beta = 0, tautilda = 0, s_new = 0, s_old = 0, ro_old = 0, omega = 0, alpha = 0, tau = 0;
	void *v = v1, *vtilda = v2, *p_new = p, *p_old = v3;    /* QMR_CS renames, src/iterative.c:916-921 */
	void *q_new = r, *q_old = v1, *pc_new = p, *pc_old = v2; /* CSYM renames, src/iterative.c:795-799 */
	/* BCGS2 (l = 2): R[0..2] residuals, U[0..2] directions, p = r~_0 (src/iterative.c:365-377); U[1] is Av so that
	 * the initial field's A.b serves the first iteration */
	void *R[3] = {r, v1, v3}, *U[3] = {v2, Av, v4};
	dcx rho0 = 0;
	/* QMR_CS_2 state (src/iterative.c:1050-1053): v = v1, d = v2, p_k = p */
	double q2_ro = 0, q2_c = 0, q2_theta = 0;
	dcx q2_eps = 0, q2_eta = 0;
	/* The checkpoint state of each method: its scalars (PHASE_VARS order, a complex as two doubles) and the physical
	 * buffers xvec, rvec, pvec, vec1.. of the reference (src/iterative.c:365-1068) */
	double sc[20];
	void *dv[7] = {x, r, p, NULL, NULL, NULL, NULL};
	if (method == ADDA_IT_QMR_CS || method == ADDA_IT_BICGSTAB) { dv[3] = v1; dv[4] = v2; dv[5] = v3; }
	else if (method == ADDA_IT_CSYM || method == ADDA_IT_QMR_CS_2) { dv[3] = v1; dv[4] = v2; }
	else if (method == ADDA_IT_BCGS2) dv[3] = v2;
	double *sd[10] = {0};  /* real scalars, then complex ones */
	dcx *sz[4] = {0};
	int nsd = 0, nsz = 0;
	switch (method) {
		case ADDA_IT_QMR_CS:
			sd[0] = &omega_old; sd[1] = &omega_new; sd[2] = &c_old; sd[3] = &c_new; nsd = 4;
			sz[0] = &beta; sz[1] = &tautilda; sz[2] = &s_old; sz[3] = &s_new; nsz = 4;
			break;
		case ADDA_IT_BICGSTAB: sz[0] = &ro_old; sz[1] = &omega; sz[2] = &alpha; nsz = 3; break;
		case ADDA_IT_CGNR: sd[0] = &ro_old_d; nsd = 1; break;
		case ADDA_IT_BICG_CS: sz[0] = &ro_old; nsz = 1; break;
		case ADDA_IT_CSYM:
			sd[0] = &beta_d; sd[1] = &c_old; sd[2] = &c_new; nsd = 3;
			sz[0] = &tau; sz[1] = &s_old; sz[2] = &s_new; nsz = 3;
			break;
		case ADDA_IT_BCGS2: sz[0] = &rho0; sz[1] = &alpha; nsz = 2; break;
		default: sd[0] = &q2_c; sd[1] = &q2_theta; sd[2] = &q2_ro; nsd = 3; sz[0] = &q2_eps; sz[1] = &q2_eta; nsz = 2;
	}
	adda_chp_head ch;
	if (load) { /* LoadIterChpoint, src/iterative.c:226-278 */
		CKG(chp_move(k, chp, method, 0, &ch, sc, dv));
		niter = ch.niter;
		counter = ch.counter;
		inprodR = ch.inprodR;
		prev_err = ch.prev_err;
		resid_scale = ch.resid_scale;
		epsB = eps * eps / resid_scale;
		for (int i = 0; i < nsd; i++) *sd[i] = sc[i];
		for (int i = 0; i < nsz; i++) *sz[i] = sc[nsd + 2 * i] + I * sc[nsd + 2 * i + 1];
		if (verbose) printf("Checkpoint (iteration) loaded\n");
		/* the renamed buffers of QMR_CS and CSYM by the parity of the iteration count (src/iterative.c:941-944, 817-820) */
		if (method == ADDA_IT_QMR_CS) {
			void *t;
			if (niter % 2 == 0) { t = v; v = vtilda; vtilda = t; }
			else { t = p_old; p_old = p_new; p_new = t; }
		}
		else if (method == ADDA_IT_CSYM) {
			void *t;
			if (niter % 2 == 0) { t = q_old; q_old = q_new; q_new = t; }
			else { t = pc_old; pc_old = pc_new; pc_new = t; }
		}
	}
	double t_last = g_chp_clock0;
	/* SaveIterChpoint, src/iterative.c:165-222 */
#define CHP_SAVE() do { \
		for (int i_ = 0; i_ < nsd; i_++) sc[i_] = *sd[i_]; \
		for (int i_ = 0; i_ < nsz; i_++) { sc[nsd + 2 * i_] = creal(*sz[i_]); sc[nsd + 2 * i_ + 1] = cimag(*sz[i_]); } \
		ch.niter = niter; ch.counter = counter; ch.inprodR = inprodR; ch.prev_err = prev_err; \
		ch.resid_scale = resid_scale; \
		CKG(chp_move(k, chp, method, 1, &ch, sc, dv)); \
		k->chp_saved = 1; \
		if (verbose) printf("Checkpoint (iteration) saved\n"); \
	} while (0)
	/* PHASE_INIT */
	if (load) {}
	else if (method == ADDA_IT_QMR_CS) {
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
	else if (method == ADDA_IT_BCGS2) { /* src/iterative.c:384-387 */
		CKG(lc(k, p, r, 1, NULL, 0, NULL, 0, 0, NULL));
		rho0 = -1;
	}
	else if (method == ADDA_IT_QMR_CS_2) { /* src/iterative.c:1070-1080 */
		q2_ro = sqrt(inprodR);
		CKG(lc(k, v1, r, 1, NULL, 0, NULL, 0, 0, NULL));
		q2_c = 1;
		q2_eps = 1;
		q2_theta = 0;
		q2_eta = -1;
	}
	else if (method == ADDA_IT_CSYM) {
		tau = sqrt(inprodR);
		CKG(lc(k, q_new, q_new, 1 / creal(tau), NULL, 0, NULL, 0, ADDA_CONJ_A, NULL));
		c_new = 1;
		c_old = 0;
		s_new = s_old = 0;
	}
	const int chp_always = chp && chp->chp_type == ADDA_CHP_ALWAYS;
	while (inprodR > epsB && niter <= maxiter && counter <= mc && !chp_exit) {
		/* complete: the iteration ran to its end; an iteration stopped by a convergence test in its middle saves no
		 * timed checkpoint (src/iterative.c:81,309,703-707) */
		int complete = 1;
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
			/* convergence at the half step: skipped when every run ends with a checkpoint (type always), which must
			 * hold a consistent r and ro_old (src/iterative.c:702-707) */
			if (inprodRp1 < epsB && !chp_always) {
				CKG(lc(k, x, x, 1, p, alpha, NULL, 0, 0, NULL));
				complete = 0;
			}
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
		else if (method == ADDA_IT_BCGS2) { /* src/iterative.c:391-483 */
			dcx rho1, sigma;
			int brk = 0;
			for (int j = 0; j < 2 && !brk; j++) { /* BiCG part */
				CKG(dotp(k, R[j], p, ADDA_DOT_CONJ, &rho1, NULL));
				if (niter == 1 && j == 0) CKG(lc(k, U[0], R[0], 1, NULL, 0, NULL, 0, 0, NULL));
				else {
					if (cabs(rho0) / (cabs(rho1) * cabs(alpha)) < 1e-10) {
						rc = host_err(__FILE__, __LINE__, "BCGS2 fails: 1/|beta| is too small");
						brk = 1;
						break;
					}
					beta = alpha * rho1 / rho0;
					for (int i = 0; i <= j; i++) CKG(lc(k, U[i], U[i], -beta, R[i], 1, NULL, 0, 0, NULL));
				}
				rho0 = rho1;
				if (!(niter == 1 && j == 0 && matvec_ready)) CKG(mv(k, U[j], U[j + 1], NULL, 0));
				CKG(dotp(k, U[j + 1], p, ADDA_DOT_CONJ, &sigma, NULL));
				if (cabs(sigma) / cabs(rho1) < 1e-10) {
					rc = host_err(__FILE__, __LINE__, "BCGS2 fails: |u_%d.r~|/|r_%d.r~| is too small", j + 1, j);
					brk = 1;
					break;
				}
				alpha = rho1 / sigma;
				CKG(lc(k, x, x, 1, U[0], alpha, NULL, 0, 0, NULL));
				for (int i = 0; i <= j; i++) CKG(lc(k, R[i], R[i], 1, U[i + 1], -alpha, NULL, 0, 0, NULL));
				CKG(mv(k, R[j], R[j + 1], NULL, 0));
			}
			if (brk) break;
			/* minimal-residual polynomial of degree 2 over Z = R^H R, Z[i][j] = R_j . conj(R_i) */
			dcx Z[3][3], y0[3], yl[3], zy0[3], zyl[3], vr = 0;
			for (int i = 0; i < 3; i++)
				for (int j = 0; j <= i; j++) {
					CKG(dotp(k, R[j], R[i], ADDA_DOT_CONJ, &Z[i][j], NULL));
					Z[j][i] = conj(Z[i][j]);
				}
			y0[0] = -1; y0[1] = Z[1][0] / Z[1][1]; y0[2] = 0;
			yl[0] = 0; yl[1] = Z[1][2] / Z[1][1]; yl[2] = -1;
			double k0 = 0, kl = 0;
			for (int i = 0; i < 3; i++) {
				zy0[i] = Z[i][0] * y0[0] + Z[i][1] * y0[1] + Z[i][2] * y0[2];
				zyl[i] = Z[i][0] * yl[0] + Z[i][1] * yl[1] + Z[i][2] * yl[2];
			}
			for (int i = 0; i < 3; i++) {
				k0 += creal(zy0[i] * conj(y0[i]));
				kl += creal(zyl[i] * conj(yl[i]));
				vr += zy0[i] * conj(yl[i]);
			}
			k0 = sqrt(k0);
			kl = sqrt(kl);
			vr /= k0 * kl;
			double avr = cabs(vr);
			dcx hg = vr * (avr > 0.7 ? avr : 0.7) / avr;
			for (int i = 0; i < 3; i++) y0[i] += (-hg * k0 / kl) * yl[i];
			omega = y0[2];
			for (int i = 1; i <= 2; i++) {
				CKG(lc(k, U[0], U[0], 1, U[i], -y0[i], NULL, 0, 0, NULL));
				CKG(lc(k, x, x, 1, R[i - 1], y0[i], NULL, 0, 0, NULL));
				CKG(lc(k, R[0], R[0], 1, R[i], -y0[i], NULL, 0, 0, NULL));
			}
			inprodRp1 = 0;
			for (int i = 0; i < 3; i++)
				inprodRp1 += creal((Z[i][0] * y0[0] + Z[i][1] * y0[1] + Z[i][2] * y0[2]) * conj(y0[i]));
			rho0 *= -omega;
		}
		else if (method == ADDA_IT_QMR_CS_2) { /* src/iterative.c:1082-1136 */
			void *vq = v1, *d = v2;
			dcx delta, bt;
			double nv2;
			CKG(lc(k, vq, vq, 1 / q2_ro, NULL, 0, NULL, 0, 0, NULL));
			CKG(dotp(k, vq, NULL, ADDA_DOT_SELF_BILIN, &delta, NULL));
			if (cabs(delta) < 1e-10) { rc = host_err(__FILE__, __LINE__, "QMR_CS_2 fails: |vT.v| is too small"); break; }
			if (niter == 1) CKG(lc(k, p, vq, 1, NULL, 0, NULL, 0, 0, NULL));
			else CKG(lc(k, p, p, -q2_ro * delta / q2_eps, vq, 1, NULL, 0, 0, NULL));
			if (niter == 1 && matvec_ready) CKG(lc(k, Av, Av, 1 / q2_ro, NULL, 0, NULL, 0, 0, NULL));
			else CKG(mv(k, p, Av, NULL, 0));
			CKG(dotp(k, p, Av, ADDA_DOT_BILIN, &q2_eps, NULL));
			bt = q2_eps / delta;
			if (cabs(q2_eps) < 1e-10) { rc = host_err(__FILE__, __LINE__, "QMR_CS_2 fails: |pT.A.p| is too small"); break; }
			CKG(lc(k, vq, vq, -bt, Av, 1, NULL, 0, 0, &nv2));
			double ro_new = sqrt(nv2);
			double th = ro_new / (q2_c * cabs(bt));
			double c_nw = 1 / sqrt(1 + th * th), sabs2 = th * th / (1 + th * th), cr = c_nw / q2_c;
			q2_eta = -q2_ro * cr * cr * q2_eta / bt;
			if (niter == 1) CKG(lc(k, d, p, q2_eta, NULL, 0, NULL, 0, 0, NULL));
			else {
				double f = q2_theta * c_nw;
				CKG(lc(k, d, d, f * f, p, q2_eta, NULL, 0, 0, NULL));
			}
			CKG(lc(k, x, x, 1, d, 1, NULL, 0, 0, NULL));
			CKG(lc(k, r, r, sabs2, vq, -q2_eta, NULL, 0, 0, &inprodRp1));
			q2_ro = ro_new;
			q2_theta = th;
			q2_c = c_nw;
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
		/* checkpoint on time (src/iterative.c:307-317): normal and always stop after it, regular goes on */
		if (chp && chp->chp_type != ADDA_CHP_NONE && chp->chp_time > 0 && complete && wall_s() - t_last > chp->chp_time) {
			CHP_SAVE();
			t_last = wall_s();
			if (chp->chp_type != ADDA_CHP_REGULAR) chp_exit = 1;
		}
	}
	if (rc == 0 && chp && chp->chp_type == ADDA_CHP_ALWAYS && !chp_exit) CHP_SAVE(); /* src/iterative.c:1496 */
	if (rc == 0 && chp_exit) rc = 2;
	else if (rc == 0 && inprodR > epsB) {
		if (niter > maxiter) rc = 1; /* warning only, src/iterative.c:1506-1508 */
		else if (counter > mc) rc = host_err(__FILE__, __LINE__, "Residual norm haven't decreased for maximum allowed number of iterations (%d)", mc) - 1;
	}
	*niter_out = niter - 1;
	*final_rel = sqrt(resid_scale * inprodR);
#undef CHP_SAVE
done:
	{
		void *vs[] = {r, Av, p, v1, v2, v3, v4};
		for (int i = 0; i < 7; i++)
			if (vs[i]) adda_hip_vec_free(k->op, vs[i]);
	}
	return rc;
#undef CKG
}

/* ---- device-resident recurrences, K right-hand sides at once ---------------------------------------------------
 * CalcInitField and PHASE_INIT as in krylov() (host, synchronising, once), then whole iterations enqueued with
 * adda_hip_dk_*_iter -- whose FP64 scalar recurrences run on the GPU -- in chunks of up to DK_CHUNK, polling the
 * device status slots after each chunk. Iterations enqueued after convergence are no-ops on the device, but they
 * still cost their launches: near convergence the chunks shrink to the number of iterations the residuals' decay
 * over the last chunk predicts (the slowest running RHS; at least one). With K > 1 every
 * launch advances all K solves (vectors batched at adda_hip_batch_stride); each RHS has its own scalars, tolerance,
 * status and history, so the K results equal K separate solves. Results are identical in exact arithmetic to
 * krylov() (same recurrences, same order of operations per vector element; reductions differ only in summation
 * order). With use_graph, the chunks after the first are recorded once as a HIP graph and replayed. */
#define DK_CHUNK 16
#define DK_EXCLUDED 3 /* status of a right-hand side that takes no part in a batched solve */

typedef struct {
	int rc;      /* 0, 1 (maxiter), -1 breakdown/error, -2 stagnation */
	int niter;   /* iterations */
	int nmv;     /* matvecs */
	double rel;  /* final relative residual of the recurrence */
	int chp_saved;
} kres;

static const char *method_name(int m)
{
	static const char *n[7] = {"QMR_CS", "BiCGStab", "CGNR", "BiCG_CS", "CSYM", "BCGS2", "QMR_CS_2"};
	return m >= 0 && m < 7 ? n[m] : "?";
}

static int enqueue_chunk(adda_hip_op *op, int prec, int method, void *st, int n, int *it, int matvec_ready, void **v,
	void **vtilda, void *Av, void **p_old, void **p_new, void *x, void *r, void *p, void *v1, void *v2, void *v3,
	void *v4)
{
	for (int c = 0; c < n; c++, (*it)++) {
		if (method == ADDA_IT_QMR_CS) {
			if (adda_hip_dk_qmr_iter(op, prec, st, *it, matvec_ready, *v, *vtilda, Av, *p_old, *p_new, x, r)) return -1;
			if (*it >= 2) { void *t = *p_old; *p_old = *p_new; *p_new = t; }
			{ void *t = *v; *v = *vtilda; *vtilda = t; }
		}
		else if (method == ADDA_IT_BICGSTAB) {
			/* v = vec1, s = vec2, rtilda = vec3, t = Avecbuffer (src/iterative.c:660-662) */
			if (adda_hip_dk_bicgstab_iter(op, prec, st, *it, matvec_ready, x, r, p, v1, v2, v3, Av)) return -1;
		}
		else if (method == ADDA_IT_BICG_CS) {
			if (adda_hip_dk_bicgcs_iter(op, prec, st, *it, matvec_ready, x, r, p, Av)) return -1;
		}
		else if (method == ADDA_IT_BCGS2) {
			/* residuals r, vec1, vec3; directions vec2, Avecbuffer, vec4; r~_0 in pvec (as in krylov()) */
			if (adda_hip_dk_bcgs2_iter(op, prec, st, *it, matvec_ready, x, p, r, v1, v3, v2, Av, v4)) return -1;
		}
		else if (method == ADDA_IT_QMR_CS_2) {
			/* v = vec1, d = vec2, p_k = pvec (src/iterative.c:1050-1053) */
			if (adda_hip_dk_qmr2_iter(op, prec, st, *it, matvec_ready, v1, v2, p, Av, x, r)) return -1;
		}
		else if (method == ADDA_IT_CSYM) {
			/* q_new, q_old, p_new, p_old renamed every iteration as in src/iterative.c:880-882 */
			if (adda_hip_dk_csym_iter(op, prec, st, *it, *v, *vtilda, *p_new, *p_old, Av, x)) return -1;
			if (*it >= 2) { void *t = *p_old; *p_old = *p_new; *p_new = t; }
			{ void *t = *v; *v = *vtilda; *vtilda = t; }
		}
		else if (adda_hip_dk_cgnr_iter(op, prec, st, *it, x, r, p, Av)) return -1;
	}
	return 0;
}

/* active: [K] or NULL (all); eps: [K]; b, x: batched vectors in precision prec; hist: [K][hist_len] or NULL,
 * hist[k*hist_len + i] = relative residual after iteration i (i = 0: the initial one). res: [K]. Returns the most
 * severe per-RHS code (an error message is set for the first failing RHS). */
static int krylov_dev_batch(adda_hip_op *op, int prec, int K, int method, const double *eps, int maxiter,
	int init_field, int verbose, const int *active, const void *b, void *x, double *hist, int hist_len, int use_graph,
	kres *res)
{
	const int mc_tab[7] = {50000, 30000, 10, 50000, 10, 15000, 50000}; /* params[].mc, src/iterative.c:118-132 */
	const int NS = ADDA_DK_NSLOTS, hl = hist_len > 1 ? hist_len : 2;
	int rc = 0, matvec_ready = 0, init_mv = 0;
	void *graph = NULL;
	void *r = adda_hip_vec_alloc_batch(op, prec, K), *Av = adda_hip_vec_alloc_batch(op, prec, K);
	void *p = adda_hip_vec_alloc_batch(op, prec, K), *v1 = adda_hip_vec_alloc_batch(op, prec, K);
	void *v2 = adda_hip_vec_alloc_batch(op, prec, K), *v3 = adda_hip_vec_alloc_batch(op, prec, K);
	void *v4 = method == ADDA_IT_BCGS2 ? adda_hip_vec_alloc_batch(op, prec, K) : NULL;
	void *st = adda_hip_dk_create_batch(op, K, hl);
	double *S = calloc((size_t)K * NS, sizeof(double)), *temp = calloc(K, sizeof(double));
	double *inprodR = calloc(K, sizeof(double));
	int *ready = calloc(K, sizeof(int));
	adda_cplx *ca = calloc(K, sizeof(adda_cplx)), *cb = calloc(K, sizeof(adda_cplx)), *cc = calloc(K, sizeof(adda_cplx));
	adda_cplx *t = calloc(K, sizeof(adda_cplx));
	double *hbuf = NULL;
	/* launch list of the matvecs (K <= 64): the RHS still running at the last poll */
	const int use_list = K > 1 && K <= 64;
	int *run = calloc(K, sizeof(int)), list_set = 0;
	unsigned long long list_mask = 0;
#define CKG(x) do { if ((x) != 0) { rc = host_err_hip(__FILE__, __LINE__); goto done; } } while (0)
	if (!r || !Av || !p || !v1 || !v2 || !v3 || !st || (method == ADDA_IT_BCGS2 && !v4)) {
		rc = host_err_hip(__FILE__, __LINE__);
		goto done;
	}
	if (!S || !temp || !inprodR || !ready || !ca || !cb || !cc || !t || !run) {
		rc = host_err(__FILE__, __LINE__, "out of host memory");
		goto done;
	}
	for (int k = 0; k < K; k++) {
		res[k].rc = res[k].niter = res[k].nmv = 0;
		res[k].rel = 0;
		S[(size_t)k * NS + ADDA_DK_STATUS] = (active && !active[k]) ? DK_EXCLUDED : 0;
		S[(size_t)k * NS + 9] = hl;
	}
	/* the statuses gate every batched call below: excluded RHS are never touched */
	CKG(adda_hip_dk_set(op, st, 0, S, K * NS));
	CKG(adda_hip_dot_batch(op, prec, K, st, b, NULL, ADDA_DOT_NORM2, NULL, temp));
	for (int k = 0; k < K; k++)
		if (!(active && !active[k]) && !(temp[k] > 0)) { rc = host_err(__FILE__, __LINE__, "right-hand side %d is zero", k); goto done; }
	/* CalcInitField, src/iterative.c:1342-1394 */
#define SETC(arr, v) for (int k_ = 0; k_ < K; k_++) { arr[k_].re = creal(v); arr[k_].im = cimag(v); }
	if (init_field == ADDA_IF_AUTO) {
		CKG(adda_hip_matvec_batch(op, prec, K, st, b, Av, NULL, 0));
		init_mv = 1;
		SETC(ca, 1); SETC(cb, -1); SETC(cc, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, r, b, ca, Av, cb, NULL, cc, 0, inprodR));
		int any = 0, all = 1;
		for (int k = 0; k < K; k++) {
			ready[k] = temp[k] < inprodR[k];
			if (ready[k]) inprodR[k] = temp[k];
			if (!(active && !active[k])) { any |= ready[k]; all &= ready[k]; }
			ca[k].re = ready[k] ? 0 : 1;  /* x = b, or 0 */
			cb[k].re = ready[k] ? 0 : -1; /* r = b - A.b, or b */
		}
		CKG(adda_hip_lincomb_batch(op, prec, K, st, x, b, ca, NULL, cc, NULL, cc, 0, NULL));
		if (any) {
			SETC(ca, 1);
			CKG(adda_hip_lincomb_batch(op, prec, K, st, r, b, ca, Av, cb, NULL, cc, 0, NULL));
		}
		matvec_ready = all;
	}
	else if (init_field == ADDA_IF_ZERO) {
		SETC(ca, 0); SETC(cb, 1);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, x, NULL, ca, NULL, ca, NULL, ca, 0, NULL));
		CKG(adda_hip_lincomb_batch(op, prec, K, st, r, b, cb, NULL, ca, NULL, ca, 0, NULL));
		for (int k = 0; k < K; k++) inprodR[k] = temp[k];
	}
	else {
		SETC(ca, 1); SETC(cb, -1); SETC(cc, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, x, b, ca, NULL, cc, NULL, cc, 0, NULL));
		CKG(adda_hip_matvec_batch(op, prec, K, st, x, Av, NULL, 0));
		init_mv = 1;
		CKG(adda_hip_lincomb_batch(op, prec, K, st, r, b, ca, Av, cb, NULL, cc, 0, inprodR));
	}
	for (int k = 0; k < K; k++) {
		double *s = S + (size_t)k * NS;
		const int excl = active && !active[k];
		const double epsB = eps[k] * eps[k] * temp[k];
		s[0] = inprodR[k];
		s[2] = epsB;
		s[3] = excl ? 0 : 1 / temp[k];
		s[4] = 1;
		s[6] = excl ? DK_EXCLUDED : (inprodR[k] > epsB ? 0 : 1);
		if (maxiter < 1 && s[6] == 0) s[6] = 2;
		s[7] = maxiter;
		s[8] = mc_tab[method];
		s[9] = hl;
		if (hist && hist_len > 0) hist[(size_t)k * hist_len] = excl ? 0 : sqrt(inprodR[k] / temp[k]);
	}
	if ((verbose & 1) && K == 1) printf("RE_000 = %.10E\n", sqrt(inprodR[0] / temp[0]));
	/* PHASE_INIT, written into the method slots (layout: the k_dk_* kernels of adda_hip.hip) */
	if (method == ADDA_IT_QMR_CS) {
		CKG(adda_hip_dot_batch(op, prec, K, st, r, NULL, ADDA_DOT_SELF_BILIN, t, NULL));
		for (int k = 0; k < K; k++) {
			double *sc = S + (size_t)k * NS + ADDA_DK_SC;
			if (active && !active[k]) { ca[k].re = ca[k].im = 0; continue; }
			const dcx beta = csqrt(CA(t[k]));
			const double omega_new = sqrt(inprodR[k]) / cabs(beta);
			const dcx tautilda = omega_new * beta, ib = 1 / beta;
			sc[0] = creal(beta); sc[1] = cimag(beta);
			sc[8] = creal(tautilda); sc[9] = cimag(tautilda);
			sc[10] = 1; sc[11] = 1; sc[12] = 0; sc[13] = omega_new;
			sc[44] = creal(ib); sc[45] = cimag(ib);
			ca[k] = AC(ib);
		}
		SETC(cb, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, v1, r, ca, NULL, cb, NULL, cb, 0, NULL));
	}
	else if (method == ADDA_IT_BICGSTAB) {
		SETC(ca, 1); SETC(cb, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, v3, r, ca, NULL, cb, NULL, cb, 0, NULL)); /* r~ = r_0 */
		for (int k = 0; k < K; k++) S[(size_t)k * NS + ADDA_DK_SC + 24] = 1;
	}
	else if (method == ADDA_IT_BICG_CS)
		for (int k = 0; k < K; k++) S[(size_t)k * NS + ADDA_DK_SC + 22] = 1;
	else if (method == ADDA_IT_BCGS2) { /* src/iterative.c:384-387: r~_0 = r_0, rho_0 = -1 */
		SETC(ca, 1); SETC(cb, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, p, r, ca, NULL, cb, NULL, cb, 0, NULL));
		for (int k = 0; k < K; k++) {
			double *sc = S + (size_t)k * NS + ADDA_DK_SC;
			sc[0] = -1;
			sc[10] = 1;
		}
	}
	else if (method == ADDA_IT_QMR_CS_2) { /* src/iterative.c:1070-1080 */
		for (int k = 0; k < K; k++) {
			double *sc = S + (size_t)k * NS + ADDA_DK_SC;
			if (active && !active[k]) continue;
			sc[0] = sqrt(inprodR[k]); /* ro */
			sc[1] = 1;                /* c */
			sc[4] = 1;                /* eps */
			sc[6] = -1;               /* eta */
			sc[10] = 1 / sc[0];
			sc[14] = sc[36] = sc[38] = 1;
		}
		SETC(ca, 1); SETC(cb, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, v1, r, ca, NULL, cb, NULL, cb, 0, NULL));
	}
	else if (method == ADDA_IT_CSYM) { /* src/iterative.c:808-818: tau = ||r_0||, q_1 = r_0* / tau (in rvec) */
		for (int k = 0; k < K; k++) {
			double *sc = S + (size_t)k * NS + ADDA_DK_SC;
			ca[k].re = ca[k].im = 0;
			if (active && !active[k]) continue;
			sc[4] = sqrt(inprodR[k]); /* tau */
			sc[2] = 1;                /* c_new; c_old, s_old, s_new, beta = 0 */
			sc[24] = 1;
			ca[k].re = 1 / sc[4];
		}
		SETC(cb, 0);
		CKG(adda_hip_lincomb_batch(op, prec, K, st, r, r, ca, NULL, cb, NULL, cb, ADDA_CONJ_A, NULL));
	}
	else
		for (int k = 0; k < K; k++) S[(size_t)k * NS + ADDA_DK_SC + 6] = 1;
	CKG(adda_hip_dk_set(op, st, 0, S, K * NS));
#undef SETC
	{
		void *v = v1, *vtilda = v2, *p_new = p, *p_old = v3;
		if (method == ADDA_IT_CSYM) { v = r; vtilda = v1; p_old = v2; } /* q_new, q_old, p_new, p_old */
		int it = 1, running = 0, chunk = DK_CHUNK, tail = 0;
		for (int k = 0; k < K; k++) {
			running |= S[(size_t)k * NS + ADDA_DK_STATUS] == 0;
			inprodR[k] = S[(size_t)k * NS];  /* the residual before the chunk, for the decay rate */
		}
		struct timespec tp0, tp1;
		clock_gettime(CLOCK_MONOTONIC, &tp0);
		/* the list follows the statuses of the last poll (RHS that stop inside a chunk stay gated until its end); a
		 * recorded graph holds the list it was recorded with, so a new list drops it (it is recorded again) */
#define UPDATE_LIST() do { \
		if (use_list) { \
			unsigned long long m_ = 0; \
			for (int k_ = 0; k_ < K; k_++) { \
				run[k_] = S[(size_t)k_ * NS + ADDA_DK_STATUS] == 0; \
				if (run[k_]) m_ |= 1ull << k_; \
			} \
			if (m_ && m_ != list_mask) { \
				CKG(adda_hip_set_launch_list(op, K, run)); \
				list_set = 1; \
				list_mask = m_; \
				if (graph) { adda_hip_graph_free(op, graph); graph = NULL; } \
			} \
		} } while (0)
		UPDATE_LIST();
		while (running) {
			/* full chunks replay the recorded graph; the shorter chunks near convergence (tail) are enqueued directly,
			 * and the graph -- recorded for one pointer rotation of QMR_CS -- is not replayed after them */
			if (graph && !tail && chunk == DK_CHUNK) {
				CKG(adda_hip_graph_launch(op, graph));
				it += DK_CHUNK;
			}
			else CKG(enqueue_chunk(op, prec, method, st, chunk, &it, matvec_ready, &v, &vtilda, Av, &p_old, &p_new, x,
				r, p, v1, v2, v3, v4));
			if (use_graph && !graph && !tail && it > 3) {
				/* record the next full chunk while the GPU runs the one just enqueued (a capture runs nothing); the
				 * pointer rotation of a chunk is the identity (DK_CHUNK is even), so the recorded chunk is valid for
				 * every later one */
				int it0 = it;
				CKG(adda_hip_capture_begin(op));
				int erc = enqueue_chunk(op, prec, method, st, DK_CHUNK, &it, matvec_ready, &v, &vtilda, Av, &p_old,
					&p_new, x, r, p, v1, v2, v3, v4);
				graph = adda_hip_capture_end(op);
				if (erc || !graph) { rc = host_err_hip(__FILE__, __LINE__); goto done; }
				it = it0;
			}
			const int ran = chunk;
			CKG(adda_hip_dk_get(op, st, 0, S, K * NS));
			UPDATE_LIST();
			running = 0;
			/* iterations the slowest running RHS still needs at the decay rate of its best residual over this chunk */
			double need = 0;
			for (int k = 0; k < K; k++) {
				const double *sk = S + (size_t)k * NS;
				if (sk[ADDA_DK_STATUS] != 0) continue;
				running = 1;
				const double r0 = inprodR[k], r1 = sk[0], eb = sk[2];
				const double n = r1 < r0 && r1 > 0 && eb > 0 ? ran * log(r1 / eb) / log(r0 / r1) : DK_CHUNK;
				if (n > need) need = n;
				inprodR[k] = r1;
			}
			chunk = need >= DK_CHUNK ? DK_CHUNK : need < 1 ? 1 : (int)ceil(need);
			if (chunk < DK_CHUNK) tail = 1;
			if (verbose & 2) { /* progress of a long solve: the largest relative residual of the running RHS */
				clock_gettime(CLOCK_MONOTONIC, &tp1);
				if (tp1.tv_sec - tp0.tv_sec + 1e-9 * (tp1.tv_nsec - tp0.tv_nsec) >= 10) {
					double worst = 0;
					int nrun = 0;
					for (int k = 0; k < K; k++) {
						const double *sk = S + (size_t)k * NS;
						if (sk[ADDA_DK_STATUS] != 0) continue;
						nrun++;
						if (sqrt(sk[0] * sk[3]) > worst) worst = sqrt(sk[0] * sk[3]);
					}
					printf("%s %s: iteration %d, %d RHS running, residual %.3E\n", method_name(method),
						prec == ADDA_F32 ? "FP32" : "FP64", it - 1, nrun, worst);
					fflush(stdout);
					tp0 = tp1;
				}
			}
		}
#undef UPDATE_LIST
		if (hist && hist_len > 1) {
			hbuf = malloc((size_t)K * hl * sizeof(double));
			if (!hbuf) { rc = host_err(__FILE__, __LINE__, "out of host memory"); goto done; }
			CKG(adda_hip_dk_history(op, st, hbuf, K * hl));
		}
		for (int k = 0; k < K; k++) {
			const double *s = S + (size_t)k * NS;
			const int status = (int)s[ADDA_DK_STATUS];
			if (status == DK_EXCLUDED) continue;
			const int niter = (int)s[ADDA_DK_NITER]; /* next iteration number */
			const int done_it = niter - 1;
			res[k].niter = done_it;
			res[k].rel = sqrt(s[3] * s[0]);
			/* matvecs of the iterations the device ran */
			if (method == ADDA_IT_QMR_CS || method == ADDA_IT_BICG_CS || method == ADDA_IT_QMR_CS_2)
				res[k].nmv = done_it - (done_it > 0 && matvec_ready);
			else if (method == ADDA_IT_CSYM) res[k].nmv = done_it;
			else if (method == ADDA_IT_BCGS2) res[k].nmv = 4 * done_it - (done_it > 0 && matvec_ready);
			else if (method == ADDA_IT_BICGSTAB) res[k].nmv = 2 * done_it - (done_it > 0 && matvec_ready);
			else res[k].nmv = 2 * done_it;
			res[k].nmv += init_mv;
			if (hbuf)
				for (int i = 1; i < niter && i < hist_len; i++) hist[(size_t)k * hist_len + i] = hbuf[(size_t)k * hl + i];
			if ((verbose & 1) && K == 1 && hist)
				for (int i = 1; i < niter && i < hist_len; i++) printf("RE_%03d = %.10E\n", i, hist[i]);
			if (status == 2) res[k].rc = 1;
			else if (status == -1) {
				res[k].rc = -1;
				if (rc > -1) rc = host_err(__FILE__, __LINE__, "%s breakdown (test %d failed)%s", method_name(method),
					(int)s[10], K > 1 ? " in a batched solve" : "");
			}
			else if (status == -2) {
				res[k].rc = -2;
				if (rc > -1) rc = host_err(__FILE__, __LINE__,
					"Residual norm haven't decreased for maximum allowed number of iterations (%d)", mc_tab[method]) - 1;
			}
			if (res[k].rc == 1 && rc == 0) rc = 1;
		}
	}
done:
	{
		void *vs[] = {r, Av, p, v1, v2, v3, v4};
		for (int i = 0; i < 7; i++)
			if (vs[i]) adda_hip_vec_free(op, vs[i]);
		if (graph) adda_hip_graph_free(op, graph);
		if (list_set) adda_hip_set_launch_list(op, K, NULL);
		if (st) adda_hip_dk_free(op, st);
		free(S); free(temp); free(inprodR); free(ready); free(ca); free(cb); free(cc); free(t); free(hbuf); free(run);
	}
	return rc;
#undef CKG
}

static int device_method(const adda_solver_opts *opt)
{
	return !opt->host_scalars && opt->chp_type == ADDA_CHP_NONE && !opt->chp_load &&
		(opt->method == ADDA_IT_QMR_CS || opt->method == ADDA_IT_BICGSTAB || opt->method == ADDA_IT_CGNR ||
		opt->method == ADDA_IT_BICG_CS || opt->method == ADDA_IT_QMR_CS_2 || opt->method == ADDA_IT_CSYM ||
		opt->method == ADDA_IT_BCGS2);
}

/* one Krylov solve of K right-hand sides: device-resident recurrences, or (K = 1) the host form */
static int krylov_run(adda_hip_op *op, const adda_solver_opts *opt, int prec, int K, const double *eps, int maxiter,
	int init_field, int verbose, const int *active, const void *b, void *x, double *hist, int hist_len, kres *res)
{
	if (device_method(opt))
		return krylov_dev_batch(op, prec, K, opt->method, eps, maxiter, init_field, verbose, active, b, x, hist,
			hist_len, opt->graph, res);
	kctx k = {op, prec, 0, 0};
	int ni = 0;
	double rel = 0;
	const int chp = opt->chp_type != ADDA_CHP_NONE || opt->chp_load;
	int rc = krylov(&k, opt->method, eps[0], maxiter, init_field, verbose & 1, b, x, hist, hist_len, &ni, &rel,
		chp && prec == ADDA_F64 ? opt : NULL);
	res[0].rc = rc;
	res[0].niter = ni;
	res[0].nmv = k.nmv;
	res[0].rel = rel;
	res[0].chp_saved = k.chp_saved;
	return rc;
}

int adda_iterative_solver_batch(adda_hip_op *op, const adda_solver_opts *opt, int K, const void *d_einc, void *d_x,
	void *d_pvec, double *resid_hist, int hist_len, adda_solve_stats *st)
{
	if (!op || !opt || !d_einc || !d_x || !d_pvec || !st) return host_err(__FILE__, __LINE__, "null argument");
	if (K < 1) return host_err(__FILE__, __LINE__, "batch of %d right-hand sides", K);
	for (int k = 0; k < K; k++) {
		memset(&st[k], 0, sizeof st[k]);
		st[k].true_resid = -1;
	}
	if (opt->method < 0 || opt->method > 6) return host_err(__FILE__, __LINE__, "unknown iterative method %d", opt->method);
	if ((opt->chp_type != ADDA_CHP_NONE || opt->chp_load) && (K != 1 || opt->precision != ADDA_PREC_FP64))
		return host_err(__FILE__, __LINE__, "checkpoints are supported for FP64 solves of one right-hand side");
	if (K > 1 && !device_method(opt))
		return host_err(__FILE__, __LINE__, "batched solves need the device-resident recurrences (QMR_CS, BiCGStab, CGNR)");
	int rc = 0;
	const int NS = ADDA_DK_NSLOTS;
	if (K > 1 && K <= 64 && adda_hip_set_launch_list(op, K, NULL)) return host_err_hip(__FILE__, __LINE__);
	double *eps = malloc(K * sizeof(double)), *b2 = malloc(K * sizeof(double)), *r2 = malloc(K * sizeof(double));
	double *r2new = malloc(K * sizeof(double)), *ih = NULL, *og_h = NULL;
	kres *kr = malloc(K * sizeof(kres));
	adda_cplx *ca = malloc(K * sizeof(adda_cplx)), *cb = malloc(K * sizeof(adda_cplx)), *cz = calloc(K, sizeof(adda_cplx));
	int *active = malloc(K * sizeof(int)), *hist_pos = calloc(K, sizeof(int)), *sel_run = calloc(K, sizeof(int));
	void *r64 = NULL, *Ax = NULL, *r32 = NULL, *e32 = NULL, *e64 = NULL, *xt = NULL, *og = NULL;
	int *fin = NULL, *acc = NULL;
#define CKB(x) do { if ((x) != 0) { rc = host_err_hip(__FILE__, __LINE__); goto out; } } while (0)
	if (!eps || !b2 || !r2 || !r2new || !kr || !ca || !cb || !cz || !active || !hist_pos || !sel_run) {
		rc = host_err(__FILE__, __LINE__, "out of host memory");
		goto out;
	}
	for (int k = 0; k < K; k++) eps[k] = opt->eps;
	/* b = S.Einc into pvec (nMult_mat, src/iterative.c:1421-1422) */
	CKB(adda_hip_mult_mat_batch(op, ADDA_F64, K, d_pvec, d_einc));
	if (opt->precision == ADDA_PREC_FP64) {
		rc = krylov_run(op, opt, ADDA_F64, K, eps, opt->maxiter, opt->init_field, opt->verbose, NULL, d_pvec, d_x,
			resid_hist, hist_len, kr);
		for (int k = 0; k < K; k++) {
			st[k].rc = kr[k].rc;
			st[k].niter = kr[k].niter;
			st[k].nmatvec = kr[k].nmv;
			st[k].resid = kr[k].rel;
			st[k].chp_saved = kr[k].chp_saved;
		}
		if (rc < 0) goto out;
		if (opt->recalc_resid) { /* ResidualNorm2, src/iterative.c:1512-1520 */
			Ax = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
			r64 = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
			if (!Ax || !r64) { rc = host_err_hip(__FILE__, __LINE__); goto out; }
			for (int k = 0; k < K; k++) { ca[k].re = 1; ca[k].im = 0; cb[k].re = -1; cb[k].im = 0; }
			CKB(adda_hip_matvec_batch(op, ADDA_F64, K, NULL, d_x, Ax, NULL, 0));
			CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, NULL, r64, d_pvec, ca, Ax, cb, NULL, cz, 0, r2));
			CKB(adda_hip_dot_batch(op, ADDA_F64, K, NULL, d_pvec, NULL, ADDA_DOT_NORM2, NULL, b2));
			for (int k = 0; k < K; k++) {
				st[k].true_resid = sqrt(r2[k] / b2[k]);
				st[k].nmatvec_fp64 = 1;
			}
		}
	}
	else {
		/* FP64 defect correction around inner Krylov solves (FP32 ones, or with ADDA_PREC_MIXED_V64 FP64 ones whose
		 * matvecs run their transforms in FP32), per RHS:
		 *   x = 0; r = b; repeat { e ~ A^-1 (r/|r|) (FP32, relative tol inner_eps); x' = x + |r| e;
		 *                          r' = b - A.x' (FP64); accept x', r' if |r'| < |r| }
		 * until ||r||/||b|| <= eps. Inner solves start from e = 0 on the normalized residual. All RHS still
		 * correcting run their inner solves together; the others are gated off (status non-zero in og). A correction
		 * that does not reduce the FP64 residual is rejected, and that RHS (like one still above eps after max_outer
		 * steps) finishes with an FP64 Krylov solve of the same correction equation -- the FP64 residual decides
		 * every outcome, so no RHS ends on an unchecked FP32 result. */
		const int v64 = opt->precision == ADDA_PREC_MIXED_V64;
		/* Default step 1e-4: each inner solve stays two decades above the FP32 operator's accuracy (about 1e-6 true
		 * residual on the 1.1 M-dipole sphere), where its recurrence residual still tracks the true one. Over four
		 * test problems this took 2.5 % more iterations than FP64 on average, 1e-5 took 5 %. */
		const double inner = opt->inner_eps > 0 ? opt->inner_eps : 1e-4;
		/* inner Krylov method (adda_solver_opts.inner_method): the complex-symmetric Lanczos methods assume A^T = A,
		 * which the FP32 operator holds to about 1e-7; their FP32 inner solves lose the symmetric Lanczos basis
		 * (measured on the 1.1 M-dipole sphere: the true residual stalls near 1e-3 of the step's right-hand side while
		 * the recurrence residual keeps falling), so by default they run BiCGStab, which assumes no symmetry. */
		adda_solver_opts iopt = *opt;
		if (opt->inner_method > 0 && opt->inner_method <= 7) iopt.method = opt->inner_method - 1;
		else if (opt->method != ADDA_IT_BICGSTAB && opt->method != ADDA_IT_CGNR && opt->method != ADDA_IT_BCGS2)
			iopt.method = ADDA_IT_BICGSTAB;
		if (K > 1 && !device_method(&iopt))
			{ rc = host_err(__FILE__, __LINE__, "batched mixed solves need a device-resident inner method"); goto out; }
		const int max_outer = opt->max_outer > 0 ? opt->max_outer : 30;
		const int cap = opt->max_inner > 0 && opt->max_inner < opt->maxiter ? opt->max_inner : opt->maxiter;
		const int ihl = (hist_len < cap ? hist_len : cap) + 2; /* inner entries beyond hist_len are not kept */
		r64 = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
		Ax = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
		e64 = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
		xt = adda_hip_vec_alloc_batch(op, ADDA_F64, K);
		if (!v64) {
			r32 = adda_hip_vec_alloc_batch(op, ADDA_F32, K);
			e32 = adda_hip_vec_alloc_batch(op, ADDA_F32, K);
		}
		og = adda_hip_dk_create_batch(op, K, 2);
		og_h = calloc((size_t)K * NS, sizeof(double));
		fin = calloc(K, sizeof(int));
		acc = calloc(K, sizeof(int));
		if (resid_hist && hist_len > 0) ih = malloc((size_t)K * ihl * sizeof(double));
		if (!r64 || !Ax || !e64 || !xt || (!v64 && (!r32 || !e32)) || !og) { rc = host_err_hip(__FILE__, __LINE__); goto out; }
		if (!og_h || !fin || !acc || (resid_hist && hist_len > 0 && !ih)) {
			rc = host_err(__FILE__, __LINE__, "out of host memory");
			goto out;
		}
#define GATE(cond) do { for (int k_ = 0; k_ < K; k_++) og_h[(size_t)k_ * NS + ADDA_DK_STATUS] = (cond) ? 0 : 1; \
		CKB(adda_hip_dk_set(op, og, 0, og_h, K * NS)); } while (0)
		for (int k = 0; k < K; k++) { ca[k].re = 1; ca[k].im = 0; cb[k].re = -1; cb[k].im = 0; }
		CKB(adda_hip_dot_batch(op, ADDA_F64, K, NULL, d_pvec, NULL, ADDA_DOT_NORM2, NULL, b2));
		for (int k = 0; k < K; k++)
			if (!(b2[k] > 0)) { rc = host_err(__FILE__, __LINE__, "right-hand side %d is zero", k); goto out; }
		CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, NULL, d_x, NULL, cz, NULL, cz, NULL, cz, 0, NULL));
		CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, NULL, r64, d_pvec, ca, NULL, cz, NULL, cz, 0, r2));
		int nact = 0;
		for (int k = 0; k < K; k++) {
			if (resid_hist && hist_len > 0) resid_hist[(size_t)k * hist_len] = sqrt(r2[k] / b2[k]);
			hist_pos[k] = 1;
			active[k] = sqrt(r2[k] / b2[k]) > opt->eps;
			nact += active[k];
		}
		/* x' = x + |r| e, r' = b - A.x' for the RHS where `sel`; r2new receives |r'|^2; e64 ends up holding r' */
#define CANDIDATE(sel) do { \
		GATE(sel); \
		for (int k_ = 0; k_ < K; k_++) { ca[k_].re = 1; cb[k_].re = sqrt(r2[k_]); } \
		CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, og, xt, d_x, ca, e64, cb, NULL, cz, 0, NULL)); \
		if (K > 1 && K <= 64) { /* the FP64 matvec launches only the selected RHS */ \
			for (int k_ = 0; k_ < K; k_++) sel_run[k_] = (sel) ? 1 : 0; \
			CKB(adda_hip_set_launch_list(op, K, sel_run)); \
		} \
		CKB(adda_hip_matvec_batch(op, ADDA_F64, K, og, xt, Ax, NULL, 0)); \
		if (K > 1 && K <= 64) CKB(adda_hip_set_launch_list(op, K, NULL)); \
		for (int k_ = 0; k_ < K; k_++) cb[k_].re = -1; \
		CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, og, e64, d_pvec, ca, Ax, cb, NULL, cz, 0, r2new)); } while (0)
		/* x = x', r = r' for the RHS where acc */
#define ACCEPT() do { \
		GATE(acc[k_]); \
		for (int k_ = 0; k_ < K; k_++) ca[k_].re = 1; \
		CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, og, d_x, xt, ca, NULL, cz, NULL, cz, 0, NULL)); \
		CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, og, r64, e64, ca, NULL, cz, NULL, cz, 0, NULL)); } while (0)
		int nouter = 0;
		while (nact > 0 && nouter < max_outer) {
			GATE(active[k_]);
			/* normalized residual in FP32; the relative target of each step: reach eps overall but no deeper than
			 * FP32 can resolve */
			for (int k = 0; k < K; k++) {
				const double rn = sqrt(r2[k]);
				ca[k].re = active[k] ? 1 / rn : 0;
				eps[k] = opt->eps / sqrt(r2[k] / b2[k]);
				if (eps[k] < inner) eps[k] = inner;
			}
			CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, og, Ax, r64, ca, NULL, cz, NULL, cz, 0, NULL));
			/* per-RHS codes in kr: an inner breakdown or stagnation ends that RHS's step; the FP64 residual below
			 * decides what happens next */
			if (v64) { /* FP64 vectors and recurrences, FP32 transforms */
				CKB(adda_hip_set_matvec_mode(op, 1));
				krylov_run(op, &iopt, ADDA_F64, K, eps, cap, ADDA_IF_ZERO, opt->verbose & 2, active, Ax, e64, ih,
					ih ? ihl : 0, kr);
				CKB(adda_hip_set_matvec_mode(op, 0));
			}
			else {
				CKB(adda_hip_vec_convert_batch(op, K, r32, ADDA_F32, Ax, ADDA_F64));
				krylov_run(op, &iopt, ADDA_F32, K, eps, cap, ADDA_IF_ZERO, opt->verbose & 2, active, r32, e32, ih,
					ih ? ihl : 0, kr);
			}
			for (int k = 0; k < K; k++) {
				if (!active[k]) continue;
				if (kr[k].rc < 0 && kr[k].niter == 0) { /* the inner solver could not even start */
					active[k] = 0;
					fin[k] = 1;
					continue;
				}
				if (ih) { /* inner residuals relative to this step's rhs -> relative to b */
					const double scale = sqrt(r2[k] / b2[k]);
					for (int i = 1; i <= kr[k].niter && hist_pos[k] < hist_len; i++)
						resid_hist[(size_t)k * hist_len + hist_pos[k]++] = ih[(size_t)k * ihl + i] * scale;
				}
				st[k].niter += kr[k].niter;
				st[k].nmatvec += kr[k].nmv;
			}
			if (!v64) CKB(adda_hip_vec_convert_batch(op, K, e64, ADDA_F64, e32, ADDA_F32));
			CANDIDATE(active[k_]);
			nouter++;
			nact = 0;
			for (int k = 0; k < K; k++) {
				acc[k] = 0;
				if (!active[k]) continue;
				st[k].nmatvec_fp64++;
				st[k].nouter++;
				if (opt->verbose && K == 1) {
					printf("outer %d: inner iterations %d, true residual %.6E\n", st[k].nouter, kr[k].niter,
						sqrt(r2new[k] / b2[k]));
					fflush(stdout);
				}
				if (r2new[k] < r2[k]) {
					acc[k] = 1;
					r2[k] = r2new[k];
					active[k] = sqrt(r2[k] / b2[k]) > opt->eps;
				}
				else { /* the FP32 correction made things worse: reject it, finish in FP64 */
					active[k] = 0;
					fin[k] = 1;
				}
				nact += active[k];
			}
			ACCEPT();
		}
		int nfin = 0;
		for (int k = 0; k < K; k++) {
			if (active[k]) fin[k] = 1; /* still above eps after max_outer steps */
			nfin += fin[k];
		}
		if (nfin) { /* FP64 Krylov solve of A.e = r/|r| to the remaining relative target, then the same acceptance.
		             * The target is halved: the solver stops on its recurrence residual, which can sit a little
		             * below the true residual that decides. */
			GATE(fin[k_]);
			for (int k = 0; k < K; k++) {
				ca[k].re = fin[k] ? 1 / sqrt(r2[k]) : 0;
				eps[k] = 0.5 * opt->eps / sqrt(r2[k] / b2[k]);
			}
			CKB(adda_hip_lincomb_batch(op, ADDA_F64, K, og, Ax, r64, ca, NULL, cz, NULL, cz, 0, NULL));
			krylov_run(op, opt, ADDA_F64, K, eps, opt->maxiter, ADDA_IF_ZERO, opt->verbose & 2, fin, Ax, e64, NULL, 0,
				kr);
			for (int k = 0; k < K; k++)
				if (fin[k]) {
					st[k].niter += kr[k].niter;
					st[k].nmatvec_fp64 += kr[k].nmv;
				}
			CANDIDATE(fin[k_]);
			for (int k = 0; k < K; k++) {
				acc[k] = 0;
				if (!fin[k]) continue;
				st[k].nmatvec_fp64++;
				st[k].nouter++;
				if (r2new[k] < r2[k]) {
					acc[k] = 1;
					r2[k] = r2new[k];
				}
				else if (kr[k].rc < 0) st[k].rc = kr[k].rc;
			}
			ACCEPT();
		}
#undef GATE
#undef CANDIDATE
#undef ACCEPT
		for (int k = 0; k < K; k++) {
			st[k].resid = st[k].true_resid = sqrt(r2[k] / b2[k]);
			if (st[k].resid > opt->eps && st[k].rc == 0) st[k].rc = 1;
		}
		rc = 0;
		for (int k = 0; k < K; k++)
			if (st[k].rc < rc || (rc == 0 && st[k].rc > 0)) rc = st[k].rc;
		if (rc < 0) goto out;
	}
	/* P = S.x (src/iterative.c:1527) */
	CKB(adda_hip_mult_mat_batch(op, ADDA_F64, K, d_pvec, d_x));
	CKB(adda_hip_sync(op));
out:
	if (opt->precision == ADDA_PREC_MIXED_V64) adda_hip_set_matvec_mode(op, 0); /* also after a failed inner solve */
	if (K > 1 && K <= 64) adda_hip_set_launch_list(op, K, NULL);                  /* also after a failed step */
	{
		void *vs[] = {r64, Ax, r32, e32, e64, xt};
		for (int i = 0; i < 6; i++)
			if (vs[i]) adda_hip_vec_free(op, vs[i]);
		if (og) adda_hip_dk_free(op, og);
	}
	free(eps); free(b2); free(r2); free(r2new); free(ih); free(og_h); free(kr); free(ca); free(cb); free(cz);
	free(active); free(hist_pos); free(fin); free(acc); free(sel_run);
	return rc;
#undef CKB
}

int adda_iterative_solver(adda_hip_op *op, const adda_solver_opts *opt, const void *d_einc, void *d_x, void *d_pvec,
	double *resid_hist, int hist_len, adda_solve_stats *st)
{
	return adda_iterative_solver_batch(op, opt, 1, d_einc, d_x, d_pvec, resid_hist, hist_len, st);
}
