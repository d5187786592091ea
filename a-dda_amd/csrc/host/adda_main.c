/* adda_main.c -- command-line front end with ADDA's option names (src/param.c:401-650) for the hot path:
 *
 *   adda_hip [-shape sphere | ellipsoid <y/x> <z/x> | box <y/x> <z/x> | read <file>] [-grid <n>] [-size <d>]
 *            [-dpl <n>] [-lambda <l>] [-m <re> <im> [<re> <im> ...]] [-iter qmr|qmr2|bicgstab|bcgs2|cgnr|bicg|csym]
 *            [-eps <k>] [-maxiter <n>] [-init_field auto|zero|inc] [-recalc_resid] [-pol ldr|cm|rrc]
 *            [-ntheta <n>] [-orient <alpha> <beta> <gamma>] [-no_reduced_fft] [-gpu <i>] [-dir <d>]
 *            [-prec fp64|mixed|mixed64] [-eps_inner <k>] [-max_inner <n>] [-host_scalars] [-graph]
 *            [-store_dip_pol] [-store_int_field] [-chpoint <time>] [-chp_type normal|regular|always]
 *            [-chp_dir <d>] [-chp_load] [-sym auto|no|enf] [-iter_inner <method>] [-prop <x> <y> <z>]
 *            [-surf <h> <mre> <mim> | -surf <h> inf] [-int_surf img]
 *
 * With -surf the particle sits above a substrate (image-dipole reflection). Cext and Cabs are those of ExtCross's
 * general form and AbsCross with inc_scale (src/crosssec.c:824-846,915); the scattered field near the substrate
 * (CalcFieldSurf) is not computed, so no mueller file is written.
 *
 * Checkpoints (ADDA's chp.0 format, interchangeable with the reference's) cover the first polarization's solve, and,
 * as in the reference, need a run whose symmetry computes the internal field once (rotational symmetry, or -sym enf).
 *
 * It runs the Calculator flow for one orientation (src/calculator.c:293-344): InitCC, GenerateB, IterativeSolver,
 * Cext/Cabs, the yz-plane amplitudes, for the Y then the X incident polarization, and writes CrossSec-Y, CrossSec-X,
 * mueller and log in ADDA's formats (GFORM/EFORM, src/const.h:142-146; src/CalculateE.c:236-246,758-759), and on
 * request DipPol-Y/X and IntField-Y/X (StoreFields, src/CalculateE.c:617-681,833-877).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../../include/adda.h"

#define PI 3.14159265358979323846264338327950288

static double now(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void die(const char *what)
{
	fprintf(stderr, "ERROR: %s: %s\n", what, adda_last_error()[0] ? adda_last_error() : adda_hip_last_error());
	exit(EXIT_FAILURE);
}

/* seconds from "%d[d,D[%d]][h,H[%d]][m,M[%d]][s,S]" (ScanTime, src/param.c:905-924) */
static double scan_time(const char *str)
{
	int t[4], n;
	char c[4] = {0, 0, 0, 0};
	n = sscanf(str, "%d%c%d%c%d%c%d%c", t, c, t + 1, c + 1, t + 2, c + 2, t + 3, c + 3);
	if (n < 1) { fprintf(stderr, "ERROR: Wrong time format '%s'\n", str); exit(EXIT_FAILURE); }
	double tim = 0;
	for (int i = 0; n > 0; n -= 2, i++) {
		const char u = c[i];
		const double f = u == 'd' || u == 'D' ? 86400 : u == 'h' || u == 'H' ? 3600 : u == 'm' || u == 'M' ? 60 :
			u == 's' || u == 'S' || u == 0 ? 1 : -1;
		if (f < 0) { fprintf(stderr, "ERROR: Illegal time format specifier (%c)\n", u); exit(EXIT_FAILURE); }
		tim += t[i] * f;
	}
	return tim;
}

static void usage(void)
{
	fprintf(stderr, "usage: adda_hip [-shape ...] [-grid n] [-size d] [-dpl n] [-m re im] [-iter name] [-eps k] ...\n");
	exit(EXIT_FAILURE);
}

/* InitRotation, zyz Euler angles in degrees: src/crosssec.c:82-138 */
/* StoreFields (src/CalculateE.c:617-681) for a complex field on the dipoles: a header of column labels, then per dipole
 * its coordinates (DipoleCoord: (position - (box-1)/2)*gridspace, src/make_particle.c:1979-1981,2266-2275), |F|^2
 * and the real and imaginary parts of the three components, all GFORM. scale: per-material factor applied to the field
 * (chi_inv for the internal field E = chi_inv*P, src/CalculateE.c:833-840), or NULL. */
static void store_field(const char *dir, const char *prefix, const char *pol, const char *name, const adda_particle *p,
	double gridspace, const adda_cplx *v, const adda_cplx *scale)
{
	char fn[1024];
	snprintf(fn, sizeof fn, "%s/%s-%s", dir, prefix, pol);
	FILE *f = fopen(fn, "w");
	if (!f) die("writing a field file");
	fprintf(f, "x y z |%s|^2 %sx.r %sx.i %sy.r %sy.i %sz.r %sz.i\n", name, name, name, name, name, name, name);
	const double c[3] = {(p->boxX - 1) / 2.0, (p->boxY - 1) / 2.0, (p->boxZ - 1) / 2.0};
	for (size_t i = 0; i < p->nd; i++) {
		double re[3], im[3], n2 = 0;
		for (int k = 0; k < 3; k++) {
			adda_cplx a = v[3 * i + k];
			if (scale) {
				const adda_cplx s = scale[p->material[i]];
				const adda_cplx b = {s.re * a.re - s.im * a.im, s.re * a.im + s.im * a.re};
				a = b;
			}
			re[k] = a.re;
			im[k] = a.im;
			n2 += a.re * a.re + a.im * a.im;
		}
		fprintf(f, "%.10g %.10g %.10g %.10g %.10g %.10g %.10g %.10g %.10g %.10g\n",
			(p->position[3 * i] - c[0]) * gridspace, (p->position[3 * i + 1] - c[1]) * gridspace,
			(p->position[3 * i + 2] - c[2]) * gridspace, n2, re[0], im[0], re[1], im[1], re[2], im[2]);
	}
	fclose(f);
}

static void rotate(double a_deg, double b_deg, double g_deg, const double in[3], double out[3])
{
	double a = a_deg * PI / 180, b = b_deg * PI / 180, g = g_deg * PI / 180;
	double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
	double m[3][3] = {{ca * cb * cg - sa * sg, sa * cb * cg + ca * sg, -sb * cg},
		{-ca * cb * sg - sa * cg, -sa * cb * sg + ca * cg, sb * sg},
		{ca * sb, sa * sb, cb}};
	for (int i = 0; i < 3; i++) out[i] = m[i][0] * in[0] + m[i][1] * in[1] + m[i][2] * in[2];
}

int main(int argc, char **argv)
{
	const char *shape = "sphere", *geom = NULL, *dir = "run_adda_hip";
	double asp_y = 1, asp_z = 1, size = -1, dpl = -1, lambda = 2 * PI, eps_k = 5, eps_inner_k = 0;
	int grid = 16, maxiter = -1, ntheta = -1, gpu = 0, reduced = 1, have_orient = 0;
	int sym = 0; /* -sym: 0 auto, 1 no, 2 enf (src/param.c:1529-1540, FinalizeSymmetry 2160-2164) */
	int store_dip_pol = 0, store_int_field = 0;
	double orient[3] = {0, 0, 0};
	adda_cplx m[16];
	int nm = 0;
	adda_solver_opts opt = {ADDA_IT_QMR_CS, 1e-5, 0, ADDA_IF_AUTO, 0, ADDA_PREC_FP64, 0, 0, 1, 0, 0, 0,
		ADDA_CHP_NONE, 0, "chpoint", 0, NULL, 0};
	int chp_time_set = 0;
	int polrel = ADDA_POL_LDR;
	double prop_in[3] = {0, 0, 1};
	int have_prop = 0, have_surf = 0;
	adda_surface surf = {0, {0, 0}, 0};
	for (int i = 1; i < argc; i++) {
		const char *a = argv[i];
#define NEED(n) do { if (i + (n) >= argc) usage(); } while (0)
		if (!strcmp(a, "-shape")) {
			NEED(1);
			shape = argv[++i];
			if (!strcmp(shape, "ellipsoid") || !strcmp(shape, "box")) {
				if (i + 2 < argc && argv[i + 1][0] != '-') { asp_y = atof(argv[++i]); asp_z = atof(argv[++i]); }
			}
			else if (!strcmp(shape, "read")) { NEED(1); geom = argv[++i]; }
		}
		else if (!strcmp(a, "-grid")) { NEED(1); grid = atoi(argv[++i]); }
		else if (!strcmp(a, "-size")) { NEED(1); size = atof(argv[++i]); }
		else if (!strcmp(a, "-dpl")) { NEED(1); dpl = atof(argv[++i]); }
		else if (!strcmp(a, "-lambda")) { NEED(1); lambda = atof(argv[++i]); }
		else if (!strcmp(a, "-m")) {
			while (i + 2 < argc && argv[i + 1][0] != '-' && nm < 16) {
				m[nm].re = atof(argv[++i]);
				m[nm].im = atof(argv[++i]);
				nm++;
			}
		}
		else if (!strcmp(a, "-iter") || !strcmp(a, "-iter_inner")) {
			NEED(1);
			const char *s = argv[++i];
			static const char *names[7] = {"qmr", "bicgstab", "cgnr", "bicg", "csym", "bcgs2", "qmr2"}; /* enum order */
			int m = -1;
			for (int j = 0; j < 7; j++)
				if (!strcmp(s, names[j])) m = j;
			if (m < 0) { fprintf(stderr, "ERROR: unsupported %s %s\n", a, s); return EXIT_FAILURE; }
			if (!strcmp(a, "-iter")) opt.method = m;
			else opt.inner_method = m + 1; /* the inner solver of the mixed forms (default: automatic) */
		}
		else if (!strcmp(a, "-eps")) { NEED(1); eps_k = atof(argv[++i]); }
		else if (!strcmp(a, "-eps_inner")) { NEED(1); eps_inner_k = atof(argv[++i]); }
		else if (!strcmp(a, "-maxiter")) { NEED(1); maxiter = atoi(argv[++i]); }
		else if (!strcmp(a, "-init_field")) {
			NEED(1);
			const char *s = argv[++i];
			opt.init_field = !strcmp(s, "zero") ? ADDA_IF_ZERO : !strcmp(s, "inc") ? ADDA_IF_INC : ADDA_IF_AUTO;
		}
		else if (!strcmp(a, "-recalc_resid")) opt.recalc_resid = 1;
		else if (!strcmp(a, "-pol")) {
			NEED(1);
			const char *s = argv[++i];
			polrel = !strcmp(s, "cm") ? ADDA_POL_CM : !strcmp(s, "rrc") ? ADDA_POL_RRC : ADDA_POL_LDR;
		}
		else if (!strcmp(a, "-ntheta")) { NEED(1); ntheta = atoi(argv[++i]) + 1; } /* intervals, src/param.c:1285-1290 */
		else if (!strcmp(a, "-orient")) {
			NEED(3);
			for (int c = 0; c < 3; c++) orient[c] = atof(argv[++i]);
			have_orient = 1;
		}
		else if (!strcmp(a, "-prop")) { /* src/param.c:1372-1382 */
			NEED(3);
			for (int c = 0; c < 3; c++) prop_in[c] = atof(argv[++i]);
			const double n2 = prop_in[0] * prop_in[0] + prop_in[1] * prop_in[1] + prop_in[2] * prop_in[2];
			if (n2 == 0) { fprintf(stderr, "ERROR: Given propagation vector is null\n"); return EXIT_FAILURE; }
			for (int c = 0; c < 3; c++) prop_in[c] /= sqrt(n2);
			have_prop = 1;
		}
		else if (!strcmp(a, "-surf")) { /* src/param.c:1510-1528 */
			NEED(2);
			surf.hsub = atof(argv[++i]);
			if (!strcmp(argv[i + 1], "inf")) { surf.inf = 1; i++; }
			else {
				NEED(2);
				surf.msub.re = atof(argv[++i]);
				surf.msub.im = atof(argv[++i]);
			}
			have_surf = 1;
		}
		else if (!strcmp(a, "-int_surf")) {
			NEED(1);
			if (strcmp(argv[++i], "img")) {
				fprintf(stderr, "ERROR: -int_surf %s is not implemented (image-dipole reflection only: img)\n", argv[i]);
				return EXIT_FAILURE;
			}
		}
		else if (!strcmp(a, "-no_reduced_fft")) reduced = 0;
		else if (!strcmp(a, "-sym")) {
			NEED(1);
			const char *v = argv[++i];
			if (!strcmp(v, "auto")) sym = 0;
			else if (!strcmp(v, "no")) sym = 1;
			else if (!strcmp(v, "enf")) sym = 2;
			else { fprintf(stderr, "ERROR: Unknown argument '%s' to '-sym' option\n", v); return EXIT_FAILURE; }
		}
		else if (!strcmp(a, "-gpu")) { NEED(1); gpu = atoi(argv[++i]); }
		else if (!strcmp(a, "-dir")) { NEED(1); dir = argv[++i]; }
		else if (!strcmp(a, "-host_scalars")) opt.host_scalars = 1;
		else if (!strcmp(a, "-max_inner")) { NEED(1); opt.max_inner = atoi(argv[++i]); }
		else if (!strcmp(a, "-graph")) opt.graph = 1;
		else if (!strcmp(a, "-chpoint")) { /* time "%d[d[%d]][h[%d]][m[%d]][s]", src/param.c:905-924,1009-1017 */
			NEED(1);
			opt.chp_time = scan_time(argv[++i]);
			chp_time_set = opt.chp_time > 0;
		}
		else if (!strcmp(a, "-chp_type")) {
			NEED(1);
			const char *s = argv[++i];
			if (!strcmp(s, "normal")) opt.chp_type = ADDA_CHP_NORMAL;
			else if (!strcmp(s, "regular")) opt.chp_type = ADDA_CHP_REGULAR;
			else if (!strcmp(s, "always")) opt.chp_type = ADDA_CHP_ALWAYS;
			else usage();
		}
		else if (!strcmp(a, "-chp_dir")) { NEED(1); opt.chp_dir = argv[++i]; }
		else if (!strcmp(a, "-chp_load")) opt.chp_load = 1;
		else if (!strcmp(a, "-store_dip_pol")) store_dip_pol = 1;
		else if (!strcmp(a, "-store_int_field")) store_int_field = 1;
		else if (!strcmp(a, "-prec")) {
			NEED(1);
			++i;
			opt.precision = !strcmp(argv[i], "mixed") ? ADDA_PREC_MIXED : !strcmp(argv[i], "mixed64") ?
				ADDA_PREC_MIXED_V64 : ADDA_PREC_FP64;
		}
		else { fprintf(stderr, "ERROR: unknown option %s\n", a); usage(); }
	}
	if (nm == 0) { m[0].re = 1.5; m[0].im = 0; nm = 1; } /* default refractive index, src/param.c:1852 */
	if (have_surf) { /* src/param.c:2070-2073 */
		if (have_orient) { fprintf(stderr, "ERROR: Currently '-orient' and '-surf' can not be used together\n"); return 1; }
		if (!reduced) {
			fprintf(stderr, "ERROR: '-surf' needs the reduced D-matrix here (drop -no_reduced_fft)\n");
			return EXIT_FAILURE;
		}
	}
	/* -chpoint without -chp_type: normal for a positive time, always for zero (src/param.c:1011-1016) */
	for (int i = 1; i < argc; i++)
		if (!strcmp(argv[i], "-chpoint") && opt.chp_type == ADDA_CHP_NONE)
			opt.chp_type = chp_time_set ? ADDA_CHP_NORMAL : ADDA_CHP_ALWAYS;
	if (opt.chp_type != ADDA_CHP_NONE && opt.chp_type != ADDA_CHP_ALWAYS && !chp_time_set) {
		fprintf(stderr, "ERROR: You must specify time for this checkpoint type\n"); /* src/param.c:2058-2059 */
		return EXIT_FAILURE;
	}
	opt.run_dir = dir;
	opt.eps = pow(10, -eps_k);
	if (eps_inner_k > 0) opt.inner_eps = pow(10, -eps_inner_k);
	double t0 = now();
	adda_particle p;
	if (geom) { if (adda_read_particle(geom, &p)) die("reading the geometry"); }
	else if (adda_make_particle(shape, grid, asp_y, asp_z, &p)) die("making the particle");
	if (p.nmat > nm) { fprintf(stderr, "ERROR: geometry has %d materials, -m gives %d\n", p.nmat, nm); return 1; }
	/* dpl, gridspace: src/make_particle.c:2185-2205 (volume correction when -size is given) */
	if (size > 0) {
		double vr = geom ? -1 : adda_volume_ratio(shape, asp_y, asp_z);
		dpl = vr > 0 ? lambda * cbrt(p.nd / vr) / size : lambda * p.boxX / size;
	}
	else if (dpl <= 0) {
		double mx = 0;
		for (int i = 0; i < nm; i++) { double a = hypot(m[i].re, m[i].im); if (a > mx) mx = a; }
		dpl = 10 * mx; /* dpl_def, src/make_particle.c:1353 */
	}
	const double gridspace = lambda / dpl, wavenum = 2 * PI / lambda;
	const double a_eq = cbrt(3 / (4 * PI) * p.nd) * gridspace, inv_G = 1 / (PI * a_eq * a_eq);
	if (maxiter <= 0) maxiter = 3 * p.boxX * p.boxY * p.boxZ; /* src/make_particle.c:1939 */
	opt.maxiter = maxiter;
	if (ntheta <= 0) { /* src/make_particle.c:1941-1946 */
		size_t nbox = (size_t)p.boxX * p.boxY * p.boxZ;
		ntheta = nbox < 1000 ? 91 : nbox < 10000 ? 181 : nbox < 100000 ? 361 : 721;
	}
	const int gx = adda_hip_fftfit(2 * p.boxX, 1), gy = adda_hip_fftfit(2 * p.boxY, 1), gz = adda_hip_fftfit(2 * p.boxZ, 1);
	/* incident frame: polarizations Y (par) and X (per) of prop (src/param.c:2116-2131); rotated by -orient */
	double prop0[3] = {prop_in[0], prop_in[1], prop_in[2]}, py0[3] = {0, 1, 0}, px0[3] = {1, 0, 0}, prop[3], incY[3],
		incX[3];
	if (fabs(prop0[0]) < 1e-300 && fabs(prop0[1]) < 1e-300) px0[0] = prop0[2] > 0 ? 1 : -1;
	else {
		const double t = sqrt(1 - prop0[2] * prop0[2]);
		px0[0] = prop0[0] * prop0[2] / t;
		px0[1] = prop0[1] * prop0[2] / t;
		px0[2] = -t;
		py0[0] = -prop0[1] / t;
		py0[1] = prop0[0] / t;
		py0[2] = 0;
	}
	if (have_orient) {
		rotate(orient[0], orient[1], orient[2], prop0, prop);
		rotate(orient[0], orient[1], orient[2], py0, incY);
		rotate(orient[0], orient[1], orient[2], px0, incX);
	}
	else { memcpy(prop, prop0, sizeof prop); memcpy(incY, py0, sizeof incY); memcpy(incX, px0, sizeof incX); }
	mkdir(dir, 0755);
	char fn[1024];
	snprintf(fn, sizeof fn, "%s/log", dir);
	FILE *log = fopen(fn, "w");
	if (!log) { fprintf(stderr, "ERROR: cannot write %s\n", fn); return 1; }
	fprintf(log, "Generated by adda_hip (MI355X backend of ADDA's FFT MatVec + Krylov path)\n");
	fprintf(log, "command: '");
	for (int i = 0; i < argc; i++) fprintf(log, "%s%s", argv[i], i + 1 < argc ? " " : "");
	fprintf(log, "'\nlambda: %.10g\nshape: %s\nbox dimensions: %dx%dx%d\n", lambda, geom ? "read" : shape, p.boxX,
		p.boxY, p.boxZ);
	fprintf(log, "refractive index: %.10g%+.10gi\nDipoles/lambda: %.10g\nRequired relative residual norm: %g\n",
		m[0].re, m[0].im, dpl, opt.eps);
	fprintf(log, "Total number of occupied dipoles: %zu\nVolume-equivalent size parameter: %.10g\n", p.nd,
		wavenum * a_eq);
	fprintf(log, "The FFT grid is: %dx%dx%d\nPrecision: %s\n", gx, gy, gz,
		opt.precision == ADDA_PREC_MIXED ? "FP32 inner iterations, FP64 defect correction" :
		opt.precision == ADDA_PREC_MIXED_V64 ? "FP64 inner iterations with FP32 transforms, FP64 defect correction" :
		"FP64");
	double tD = now();
	adda_hip_op *op = adda_hip_op_create(gpu, p.boxX, p.boxY, p.boxZ, gx, gy, gz, p.nd, p.position, p.material, p.nmat,
		wavenum, gridspace, reduced, opt.precision != ADDA_PREC_FP64);
	if (!op) die("creating the operator");
	if (have_surf && adda_set_surface(op, &p, gridspace, &surf)) die("setting the substrate");
	tD = now() - tD;
	if (have_surf) {
		if (surf.inf) fprintf(log, "Particle is placed near the perfectly reflecting substrate\n");
		else fprintf(log, "Particle is placed near the substrate with refractive index %.10g%+.10gi\n", surf.msub.re,
			surf.msub.im);
		fprintf(log, "  height of the particle center: %.10g\n", surf.hsub);
	}
	void *einc = adda_hip_vec_alloc(op, ADDA_F64), *x = adda_hip_vec_alloc(op, ADDA_F64),
		*pv = adda_hip_vec_alloc(op, ADDA_F64);
	if (!einc || !x || !pv) die("allocating vectors");
	/* symmetries of the predefined shapes (src/make_particle.c:1478,1633; param.c:1903,2138) and the 360-degree
	 * extension of the yz-plane range when incPolY and -incPolY are not equivalent (TestExtendThetaRange,
	 * src/CalculateE.c:396-418, calculator.c:674) */
	int symXY = !geom && !have_orient && !have_prop, symR = symXY && (!strcmp(shape, "sphere") || asp_y == 1),
		symZ = !geom && !have_orient && !have_prop && !have_surf;
	if (sym) symXY = symR = symZ = sym == 2; /* -sym no / enf */
	int sym_incY = ((symXY || fabs(incY[0]) < 1e-10) && (symXY || fabs(incY[1]) < 1e-10)) || symR;
	sym_incY = sym_incY && (symZ || fabs(incY[2]) < 1e-10);
	const int npoints = sym_incY ? ntheta : 2 * (ntheta - 1);
	/* the reference takes checkpoints only on runs that compute the internal field once, for a single polarization
	 * (FinalizeSymmetry, src/param.c:2171-2174); this front end solves the second polarization of such a run afresh
	 * (equal to the reference's rotated first one), without checkpoints */
	if ((opt.chp_type != ADDA_CHP_NONE || opt.chp_load) && !symR)
		die("Currently checkpoints can be used when internal fields are calculated only once, i.e. for a single "
			"incident polarization");
	adda_cplx *eplX = malloc(2 * npoints * sizeof(adda_cplx)), *eplY = malloc(2 * npoints * sizeof(adda_cplx));
	adda_cplx *cc = malloc((size_t)p.nmat * 3 * sizeof(adda_cplx)), *ccs = malloc((size_t)p.nmat * 3 * sizeof(adda_cplx));
	int total_mv = 0, total_iter = 0;
	double t_solve = 0;
	for (int pol = 0; pol < 2; pol++) { /* Y then X, src/calculator.c:312-330 */
		const double *inc = pol == 0 ? incY : incX;
		const char *nm_ = pol == 0 ? "Y" : "X";
		for (int i = 0; i < p.nmat; i++) { /* InitCC, src/calculator.c:266-289 */
			adda_cplx c = adda_couple_constant(m[i], polrel, wavenum, gridspace, prop, inc);
			double r = hypot(c.re, c.im), ph = atan2(c.im, c.re);
			for (int j = 0; j < 3; j++) {
				cc[3 * i + j] = c;
				ccs[3 * i + j].re = sqrt(r) * cos(ph / 2);
				ccs[3 * i + j].im = sqrt(r) * sin(ph / 2);
			}
		}
		fprintf(log, "\nhere we go, calc %s\n\nCoupleConstant:%.10g%+.10gi\n", nm_, cc[0].re, cc[0].im);
		if (adda_hip_set_cc_sqrt(op, ccs)) die("uploading cc_sqrt");
		double inc_scale = 1;
		if (have_surf) {
			if (adda_generate_b_surf(op, &p, &surf, wavenum, gridspace, prop, incX, incY, pol, einc, &inc_scale))
				die("GenerateB");
		}
		else if (adda_generate_b(op, &p, wavenum, gridspace, prop, inc, einc)) die("GenerateB");
		double hist[4096];
		adda_solve_stats st;
		double ts = now();
		/* checkpoints cover the solve of the first polarization (the reference allows them only for runs that solve
		 * one, src/param.c:2173-2174) */
		adda_solver_opts o = opt;
		if (pol == 1) { o.chp_type = ADDA_CHP_NONE; o.chp_load = 0; }
		int rc = adda_iterative_solver(op, &o, einc, x, pv, hist, 4096, &st);
		ts = now() - ts;
		t_solve += ts;
		if (rc < 0) die("IterativeSolver");
		if (rc == 1) fprintf(stderr, "WARNING: iterations haven't converged in %d iterations\n", maxiter);
		if (st.chp_saved) fprintf(log, "Checkpoint (iteration) saved\n");
		if (rc == 2) { /* stopped after saving a checkpoint (CHP_EXIT, src/ADDAmain.c) */
			printf("Checkpoint (iteration) saved after %d iterations; the run stops here\n", st.niter);
			fclose(log);
			return 0;
		}
		for (int i = 0; i <= st.niter && i < 4096; i++) fprintf(log, "RE_%03d = %.10E\n", i, hist[i]);
		total_mv += st.nmatvec + st.nmatvec_fp64;
		total_iter += st.niter;
		double cext, cabs;
		if (have_surf) {
			if (adda_ext_cross_gen(op, wavenum, inc_scale, pv, einc, &cext)) die("ExtCross");
		}
		else if (adda_ext_cross(op, &p, wavenum, gridspace, pv, prop, inc, &cext)) die("ExtCross");
		if (adda_abs_cross(op, &p, wavenum, cc, pv, &cabs)) die("AbsCross");
		cabs *= inc_scale;
		if (!have_surf &&
			adda_eplane_yz(op, &p, wavenum, gridspace, pv, prop, incY, incX, ntheta, npoints, pol == 0 ? eplY : eplX))
			die("CalcEplaneYZ");
		if (store_dip_pol || store_int_field) { /* CalculateE, src/CalculateE.c:875-877 */
			adda_cplx *P = malloc(3 * p.nd * sizeof(adda_cplx)), *chi_inv = malloc(p.nmat * sizeof(adda_cplx));
			if (!P || !chi_inv || adda_hip_vec_download(op, ADDA_F64, P, pv)) die("downloading P");
			if (store_int_field) { /* chi_inv = 4 pi/(V (m^2-1)), src/calculator.c:275-278 */
				const double V = gridspace * gridspace * gridspace;
				for (int i = 0; i < p.nmat; i++) {
					const double ar = m[i].re * m[i].re - m[i].im * m[i].im - 1, ai = 2 * m[i].re * m[i].im;
					const double d = V * (ar * ar + ai * ai);
					chi_inv[i].re = 4 * PI * ar / d;
					chi_inv[i].im = -4 * PI * ai / d;
				}
				store_field(dir, "IntField", nm_, "E", &p, gridspace, P, chi_inv);
			}
			if (store_dip_pol) store_field(dir, "DipPol", nm_, "P", &p, gridspace, P, NULL);
			free(P);
			free(chi_inv);
		}
		snprintf(fn, sizeof fn, "%s/CrossSec-%s", dir, nm_);
		FILE *f = fopen(fn, "w");
		if (!f) die("writing CrossSec");
		fprintf(f, "Cext\t= %.10g\nQext\t= %.10g\nCabs\t= %.10g\nQabs\t= %.10g\n", cext, cext * inv_G, cabs, cabs * inv_G);
		fclose(f);
		printf("%s: iterations %d, matvecs %d (+%d FP64), residual %.3E, Cext = %.10g, Cabs = %.10g, %.3f s\n", nm_,
			st.niter, st.nmatvec, st.nmatvec_fp64, st.resid, cext, cabs, ts);
	}
	double *mu = malloc((size_t)npoints * 16 * sizeof(double));
	if (!have_surf) {
		adda_mueller_yz(eplX, eplY, npoints, mu);
		snprintf(fn, sizeof fn, "%s/mueller", dir);
		FILE *f = fopen(fn, "w");
		if (!f) die("writing mueller");
		fprintf(f, "theta s11 s12 s13 s14 s21 s22 s23 s24 s31 s32 s33 s34 s41 s42 s43 s44\n");
		for (int i = 0; i < npoints; i++) {
			fprintf(f, "%.2f", i * 180.0 / (ntheta - 1));
			for (int j = 0; j < 16; j++) fprintf(f, " %.10E", mu[16 * i + j]);
			fprintf(f, "\n");
		}
		fclose(f);
	}
	fprintf(log, "\nTotal number of iterations: %d\nTotal number of matrix-vector products: %d\n", total_iter, total_mv);
	fprintf(log, "Total wall time: %.4f\n  init Dmatrix: %.4f\n  solves: %.4f\n", now() - t0, tD, t_solve);
	fclose(log);
	adda_hip_vec_free(op, einc);
	adda_hip_vec_free(op, x);
	adda_hip_vec_free(op, pv);
	adda_hip_op_destroy(op);
	adda_free_particle(&p);
	free(eplX); free(eplY); free(cc); free(ccs); free(mu);
	return 0;
}
