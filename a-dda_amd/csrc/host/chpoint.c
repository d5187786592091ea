/* chpoint.c -- ADDA's iteration checkpoint files (SaveIterChpoint / LoadIterChpoint, src/iterative.c:165-278).
 *
 * The binary file <dir>/chp.0 holds, in this order and in the machine's native layout: int ind_m (the solver's index
 * in the reference's params[] table, src/iterative.c:118-132), size_t local_nRows, int niter, int counter, double
 * inprodR, double prev_err, double resid_scale, the solver's own scalars (PHASE_VARS of each ITER_FUNC), the vectors
 * xvec, rvec, pvec and the solver's own vectors, each local_nRows doublecomplex values. <dir>/chp.log holds one line
 * naming the run's output directory. The files are interchangeable with the reference's: a state saved here resumes
 * there and the reverse. Host memory only; the solver moves its vectors between host and device.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "../../../include/adda.h"
#include "host_internal.h"

/* per method of enum adda_iter: index in params[] (alphabetical there), scalars in doubles (a doublecomplex counts
 * two), extra vectors -- src/iterative.c:118-132 and the PHASE_VARS blocks of each solver */
static const int chp_tab[7][3] = {
	{5, 12, 3}, /* QMR_CS: omega_old omega_new c_old c_new (double), beta tautilda s_old s_new (complex); vec1-3 */
	{2, 6, 3},  /* BiCGStab: ro_old omega alpha (complex); vec1 (v) vec2 (s) vec3 (r~) */
	{3, 1, 0},  /* CGNR: ro_old (double) */
	{1, 2, 0},  /* BiCG_CS: ro_old (complex) */
	{4, 9, 2},  /* CSYM: beta c_old c_new (double), tau s_old s_new (complex); vec1 vec2 */
	{0, 4, 1},  /* BCGS2: rho0 alpha (complex); vec2 (u_0) */
	{6, 7, 2},  /* QMR_CS_2: c_old theta_old ro_old (double), eps eta (complex); vec1 (v) vec2 (d) */
};

int adda_chp_layout(int method, int *ind_m, int *nscal, int *nvec)
{
	if (method < 0 || method > 6) return host_err(__FILE__, __LINE__, "unknown iterative method %d", method);
	if (ind_m) *ind_m = chp_tab[method][0];
	if (nscal) *nscal = chp_tab[method][1];
	if (nvec) *nvec = chp_tab[method][2];
	return 0;
}

static void chp_name(char *b, size_t n, const char *dir, const char *f) { snprintf(b, n, "%s/%s", dir, f); }

int adda_chp_write(const char *dir, const char *run_dir, int method, size_t nrows, const adda_chp_head *h,
	const double *scal, const adda_cplx *const *vecs)
{
	int ind_m, ns, nv;
	if (adda_chp_layout(method, &ind_m, &ns, &nv)) return -1;
	if (!dir || !h || (!scal && ns) || !vecs) return host_err(__FILE__, __LINE__, "null argument");
	char fn[4096];
	chp_name(fn, sizeof fn, dir, "chp.log");
	FILE *f = fopen(fn, "w");
	if (!f) {  /* create the directory, as MkDirErr in the reference */
		if (mkdir(dir, 0755) && errno != EEXIST) return host_err(__FILE__, __LINE__, "cannot create directory '%s'", dir);
		if (!(f = fopen(fn, "w"))) return host_err(__FILE__, __LINE__, "cannot open '%s' for writing", fn);
	}
	fprintf(f, "Info about the run, which produced the checkpoint, can be found in ../%s", run_dir ? run_dir : ".");
	if (fclose(f)) return host_err(__FILE__, __LINE__, "failed writing '%s'", fn);
	chp_name(fn, sizeof fn, dir, "chp.0");
	if (!(f = fopen(fn, "wb"))) return host_err(__FILE__, __LINE__, "cannot open '%s' for writing", fn);
	int ok = fwrite(&ind_m, sizeof(int), 1, f) == 1 && fwrite(&nrows, sizeof(size_t), 1, f) == 1 &&
		fwrite(&h->niter, sizeof(int), 1, f) == 1 && fwrite(&h->counter, sizeof(int), 1, f) == 1 &&
		fwrite(&h->inprodR, sizeof(double), 1, f) == 1 && fwrite(&h->prev_err, sizeof(double), 1, f) == 1 &&
		fwrite(&h->resid_scale, sizeof(double), 1, f) == 1 && (ns == 0 || fwrite(scal, sizeof(double), ns, f) == (size_t)ns);
	for (int i = 0; ok && i < 3 + nv; i++) ok = fwrite(vecs[i], sizeof(adda_cplx), nrows, f) == nrows;
	if (fclose(f) || !ok) return host_err(__FILE__, __LINE__, "Failed writing to file '%s'", fn);
	return 0;
}

int adda_chp_read(const char *dir, int method, size_t nrows, adda_chp_head *h, double *scal, adda_cplx *const *vecs)
{
	int ind_m, ns, nv, ind_f;
	size_t nrows_f;
	if (adda_chp_layout(method, &ind_m, &ns, &nv)) return -1;
	if (!dir || !h || (!scal && ns) || !vecs) return host_err(__FILE__, __LINE__, "null argument");
	char fn[4096];
	chp_name(fn, sizeof fn, dir, "chp.0");
	FILE *f = fopen(fn, "rb");
	if (!f) return host_err(__FILE__, __LINE__, "cannot open '%s' for reading", fn);
	int rc = 0;
	if (fread(&ind_f, sizeof(int), 1, f) != 1 || fread(&nrows_f, sizeof(size_t), 1, f) != 1)
		rc = host_err(__FILE__, __LINE__, "Failed reading from file '%s'", fn);
	else if (ind_f != ind_m) rc = host_err(__FILE__, __LINE__, "File '%s' is for different iterative method", fn);
	else if (nrows_f != nrows) rc = host_err(__FILE__, __LINE__, "File '%s' is for different vector size", fn);
	else if (fread(&h->niter, sizeof(int), 1, f) != 1 || fread(&h->counter, sizeof(int), 1, f) != 1 ||
		fread(&h->inprodR, sizeof(double), 1, f) != 1 || fread(&h->prev_err, sizeof(double), 1, f) != 1 ||
		fread(&h->resid_scale, sizeof(double), 1, f) != 1 || (ns && fread(scal, sizeof(double), ns, f) != (size_t)ns))
		rc = host_err(__FILE__, __LINE__, "Failed reading from file '%s'", fn);
	for (int i = 0; rc == 0 && i < 3 + nv; i++)
		if (fread(vecs[i], sizeof(adda_cplx), nrows, f) != nrows)
			rc = host_err(__FILE__, __LINE__, "Failed reading from file '%s'", fn);
	char ch;
	if (rc == 0 && fread(&ch, 1, 1, f) != 0) rc = host_err(__FILE__, __LINE__, "File '%s' is too long", fn);
	fclose(f);
	return rc;
}
