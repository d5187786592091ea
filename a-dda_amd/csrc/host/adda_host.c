/* adda_host.c -- host side of the MI355X ADDA solver: particle builder (MakeParticle subset), polarizability
 * (CoupleConstant), plane wave, scattering quantities (CalcFieldFree, ExtCross, AbsCross, CalcEplaneYZ, Mueller).
 * Bulk sums over dipoles run on the device (adda_hip_field_sums, adda_hip_weighted_norm2); the O(1) prefactors and
 * projections stay here in FP64.
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

#define PI 3.14159265358979323846264338327950288
#define TWO_PI 6.283185307179586476925286766559
#define FOUR_PI 12.566370614359172953850573533118
/* src/const.h:234-236 */
#define LDR_B1 1.8915316
#define LDR_B2 -0.1648469
#define LDR_B3 1.7700004

static char g_err[600];

int host_err(const char *file, int line, const char *fmt, ...)
{
	char msg[480];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(msg, sizeof msg, fmt, ap);
	va_end(ap);
	snprintf(g_err, sizeof g_err, "%s:%d: %s", file, line, msg);
	return -1;
}

int host_err_hip(const char *file, int line)
{
	const char *e = adda_hip_last_error();
	if (e && e[0]) snprintf(g_err, sizeof g_err, "%s:%d: %s", file, line, e);
	else if (!g_err[0]) snprintf(g_err, sizeof g_err, "%s:%d: backend call failed", file, line);
	return -1;
}

const char *adda_last_error(void) { return g_err; }

/* ---- particle ------------------------------------------------------------------------------------------------- */
/* FitBox / FitBox_yz with jagged = 1, src/make_particle.c:1284-1307 */
static int fit_box(int box) { return 2 * ((box + 1) / 2); }
static int fit_box_yz(double size) { return 2 * (int)floor((size + 1) / 2); }

int adda_make_particle(const char *shape, int grid, double aspect_y, double aspect_z, adda_particle *p)
{
	if (!shape || !p || grid <= 0) return host_err(__FILE__, __LINE__, "bad particle arguments");
	memset(p, 0, sizeof *p);
	int sph = !strcmp(shape, "sphere"), ell = !strcmp(shape, "ellipsoid"), box = !strcmp(shape, "box");
	if (!sph && !ell && !box) return host_err(__FILE__, __LINE__, "unsupported shape '%s'", shape);
	int bx = fit_box(grid), by, bz;
	if (sph) by = bz = bx;
	else {
		by = fit_box(fit_box_yz(aspect_y * bx));
		bz = fit_box(fit_box_yz(aspect_z * bx));
	}
	if (bx > 65535 || by > 65535 || bz > 65535) return host_err(__FILE__, __LINE__, "box too large");
	size_t cap = (size_t)bx * by * bz;
	p->position = malloc(3 * cap * sizeof(unsigned short));
	p->material = malloc(cap * sizeof(int));
	if (!p->position || !p->material) return host_err(__FILE__, __LINE__, "out of host memory");
	size_t n = 0;
	/* src/make_particle.c:1996-2007 with jagged=1: centre (i - box/2 + 0.5)/boxX; SH_SPHERE 2136-2137,
	 * SH_ELLIPSOID (invsqY/invsqZ, 1635-1637), SH_BOX (haspY/haspZ, 1480-1481) */
	for (int k = 0; k < bz; k++)
		for (int j = 0; j < by; j++)
			for (int i = 0; i < bx; i++) {
				double xr = (i - bx / 2 + 0.5) / bx, yr = (j - by / 2 + 0.5) / bx, zr = (k - bz / 2 + 0.5) / bx;
				int in;
				if (sph) in = xr * xr + yr * yr + zr * zr <= 0.25;
				else if (ell) in = xr * xr + yr * yr / (aspect_y * aspect_y) + zr * zr / (aspect_z * aspect_z) <= 0.25;
				else in = fabs(yr) <= aspect_y / 2 && fabs(zr) <= aspect_z / 2;
				if (!in) continue;
				p->position[3 * n] = (unsigned short)i;
				p->position[3 * n + 1] = (unsigned short)j;
				p->position[3 * n + 2] = (unsigned short)k;
				p->material[n] = 0;
				n++;
			}
	if (n == 0) return host_err(__FILE__, __LINE__, "All dipoles of the scatterer are void");
	p->boxX = bx;
	p->boxY = by;
	p->boxZ = bz;
	p->nd = n;
	p->nmat = 1;
	return 0;
}

double adda_volume_ratio(const char *shape, double aspect_y, double aspect_z)
{
	if (!strcmp(shape, "sphere")) return PI / 6;                       /* src/make_particle.c:1553 */
	if (!strcmp(shape, "ellipsoid")) return PI / 6 * aspect_y * aspect_z; /* src/make_particle.c:1637 */
	if (!strcmp(shape, "box")) return aspect_y * aspect_z;               /* src/make_particle.c:1482 */
	return -1;
}

int adda_read_particle(const char *fname, adda_particle *p)
{
	if (!fname || !p) return host_err(__FILE__, __LINE__, "null argument");
	memset(p, 0, sizeof *p);
	FILE *f = fopen(fname, "r");
	if (!f) return host_err(__FILE__, __LINE__, "cannot open geometry file '%s'", fname);
	size_t cap = 1024, n = 0;
	p->position = malloc(3 * cap * sizeof(unsigned short));
	p->material = malloc(cap * sizeof(int));
	char line[512];
	int maxc[3] = {0, 0, 0}, nmat = 1;
	while (fgets(line, sizeof line, f)) {
		if (line[0] == '#' || line[0] == '\n' || line[0] == '\r') continue;
		int x, y, z, m = 1;
		int got = sscanf(line, "%d %d %d %d", &x, &y, &z, &m);
		if (got < 3) continue;
		if (x < 0 || y < 0 || z < 0 || x > 65534 || y > 65534 || z > 65534 || m < 1) {
			fclose(f);
			return host_err(__FILE__, __LINE__, "bad geometry line '%s'", line);
		}
		if (n == cap) {
			cap *= 2;
			p->position = realloc(p->position, 3 * cap * sizeof(unsigned short));
			p->material = realloc(p->material, cap * sizeof(int));
			if (!p->position || !p->material) { fclose(f); return host_err(__FILE__, __LINE__, "out of host memory"); }
		}
		p->position[3 * n] = (unsigned short)x;
		p->position[3 * n + 1] = (unsigned short)y;
		p->position[3 * n + 2] = (unsigned short)z;
		p->material[n] = m - 1;
		if (m > nmat) nmat = m;
		if (x > maxc[0]) maxc[0] = x;
		if (y > maxc[1]) maxc[1] = y;
		if (z > maxc[2]) maxc[2] = z;
		n++;
	}
	fclose(f);
	if (n == 0) return host_err(__FILE__, __LINE__, "All dipoles of the scatterer are void");
	/* box = even size covering the dipoles (FitBox of the extent, src/make_particle.c:1194-1203) */
	p->boxX = fit_box(maxc[0] + 1);
	p->boxY = fit_box(maxc[1] + 1);
	p->boxZ = fit_box(maxc[2] + 1);
	p->nd = n;
	p->nmat = nmat;
	return 0;
}

void adda_free_particle(adda_particle *p)
{
	if (!p) return;
	free(p->position);
	free(p->material);
	memset(p, 0, sizeof *p);
}

/* ---- physics ------------------------------------------------------------------------------------------------- */
/* polCM, polM, polMplusRR, pol3coef: src/calculator.c:100-133; LDR S-term 214-221 */
adda_cplx adda_couple_constant(adda_cplx mo, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3])
{
	dcx m = mo.re + I * mo.im, m2 = m * m, res;
	double dipvol = gridspace * gridspace * gridspace, kd = wavenum * gridspace;
	dcx cm = (3 / FOUR_PI) * dipvol * (m2 - 1) / (m2 + 2);
	if (polrel == ADDA_POL_CM) res = cm;
	else {
		dcx M = 0;
		if (polrel == ADDA_POL_LDR) {
			double S = 0;
			for (int i = 0; i < 3; i++) S += prop[i] * prop[i] * incpol[i] * incpol[i];
			M = (LDR_B1 + (LDR_B2 + LDR_B3 * S) * m2) * kd * kd;
		}
		M += I * 2 * kd * kd * kd / 3;
		res = cm / (1 - (cm / dipvol) * M);
	}
	adda_cplx r = {creal(res), cimag(res)};
	return r;
}

int adda_generate_b(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const double prop[3],
	const double incpol[3], void *d_einc)
{
	(void)p;
	if (adda_hip_plane_wave(op, wavenum, gridspace, prop, incpol, d_einc)) return host_err_hip(__FILE__, __LINE__);
	return 0;
}

/* ---- particle above a substrate ------------------------------------------------------------------------------ */
int adda_set_surface(adda_hip_op *op, const adda_particle *p, double gridspace, const adda_surface *s)
{
	if (!op || !p || !s) return host_err(__FILE__, __LINE__, "null argument");
	if (!(s->hsub > 0)) return host_err(__FILE__, __LINE__, "height above surface %g is not positive", s->hsub);
	/* lowest dipole coordinate relative to the centre, src/make_particle.c:2268-2284 */
	const double cz = (p->boxZ - 1) / 2.0;
	int zmin = p->boxZ;
	for (size_t i = 0; i < p->nd; i++)
		if (p->position[3 * i + 2] < zmin) zmin = p->position[3 * i + 2];
	const double minZco = (zmin - cz) * gridspace;
	if (s->hsub <= -minZco)
		return host_err(__FILE__, __LINE__, "The particle must be entirely above the substrate. There exist a dipole "
			"with z=%g (relative to the center), making specified height of the center (%g) too small", minZco, s->hsub);
	const double zsum_shift = 2 * (s->hsub / gridspace - cz);
	dcx rcn = -1;
	if (!s->inf) {
		const dcx m = s->msub.re + I * s->msub.im;
		rcn = (1 - m * m) / (1 + m * m);
	}
	const adda_cplx r = {creal(rcn), cimag(rcn)};
	if (adda_hip_set_surface(op, zsum_shift, r)) return host_err_hip(__FILE__, __LINE__);
	return 0;
}

/* cSqrtCut (src/cmplx.h:58-82): the branch cut does not depend on the sign of a zero imaginary part */
static dcx csqrt_cut(dcx a)
{
	if (cimag(a) == 0) return creal(a) >= 0 ? sqrt(creal(a)) : I * sqrt(-creal(a));
	return csqrt(a);
}

/* Fresnel coefficients, src/cmplx.h:672-709 (normal wave-vector components ki, kt; mr = mt/mi) */
static dcx fr_rs(dcx ki, dcx kt) { return (ki - kt) / (ki + kt); }
static dcx fr_ts(dcx ki, dcx kt) { return 2 * ki / (ki + kt); }
static dcx fr_rp(dcx ki, dcx kt, dcx mr) { return (mr * mr * ki - kt) / (mr * mr * ki + kt); }
static dcx fr_tp(dcx ki, dcx kt, dcx mr) { return 2 * mr * ki / (mr * mr * ki + kt); }

int adda_generate_b_surf(adda_hip_op *op, const adda_particle *p, const adda_surface *s, double wavenum,
	double gridspace, const double prop[3], const double incpol_x[3], const double incpol_y[3], int which_x,
	void *d_einc, double *inc_scale)
{
	(void)p;
	if (!op || !s || !prop || !incpol_x || !incpol_y || !d_einc) return host_err(__FILE__, __LINE__, "null argument");
	if (prop[2] == 0)
		return host_err(__FILE__, __LINE__, "Ambiguous setting of beam propagating along the surface. Please specify "
			"the incident direction to have (arbitrary) small positive or negative z-component");
	if (s->inf && prop[2] > 0)
		return host_err(__FILE__, __LINE__, "Perfectly reflecting surface ('-surf ... inf') is incompatible with "
			"incident direction from below (including the default one)");
	const dcx msub = s->msub.re + I * s->msub.im;
	const double pp = prop[0] * prop[0] + prop[1] * prop[1];
	/* InitBeam: ki, kt, ktVec (ROUND_ERR 1e-15, SQRT_RND_ERR 3e-8: src/const.h:337-338) */
	dcx ki, kt = 0, ktv[3] = {0, 0, 0};
	double scale = 1;
	if (prop[2] > 0) {
		scale = 1 / creal(msub);
		ki = msub * prop[2];
		kt = (cabs(msub - 1) < 1e-15 && cabs(ki) < 3e-8) ? ki : csqrt_cut(1 - msub * msub * pp);
		ktv[0] = msub * prop[0];
		ktv[1] = msub * prop[1];
		ktv[2] = kt;
	}
	else {
		ki = -prop[2];
		if (!s->inf) {
			kt = (cabs(msub - 1) < 1e-15 && cabs(ki) < 3e-8) ? ki : csqrt_cut(msub * msub - pp);
			ktv[0] = prop[0];
			ktv[1] = prop[1];
			ktv[2] = -kt;
		}
	}
	/* the beam's frame: ex = incPolY, ey = -incPolX (Y); ex = incPolX, ey = incPolY (X), src/GenerateB.c:247-254 */
	double ex[3], ey[3];
	for (int c = 0; c < 3; c++) {
		ex[c] = which_x ? incpol_x[c] : incpol_y[c];
		ey[c] = which_x ? incpol_y[c] : -incpol_x[c];
	}
	adda_cplx kv[6], am[6];
	int nw;
	if (prop[2] > 0) { /* transmitted wave only */
		dcx et[3], tc;
		if (!which_x) {
			for (int c = 0; c < 3; c++) et[c] = ex[c];
			tc = fr_ts(ki, kt);
		}
		else { /* crCrossProd(ey, ktVec) */
			et[0] = ey[1] * ktv[2] - ey[2] * ktv[1];
			et[1] = ey[2] * ktv[0] - ey[0] * ktv[2];
			et[2] = ey[0] * ktv[1] - ey[1] * ktv[0];
			tc = fr_tp(ki, kt, 1 / msub);
		}
		const dcx f = tc * cexp(I * wavenum * (kt - ki) * s->hsub);
		for (int c = 0; c < 3; c++) {
			kv[c].re = creal(ktv[c]);
			kv[c].im = cimag(ktv[c]);
			am[c].re = creal(f * et[c]);
			am[c].im = cimag(f * et[c]);
		}
		nw = 1;
	}
	else { /* incident plus reflected wave (prIncRefl = prop with z inverted) */
		dcx er[3], rc;
		if (!which_x) {
			for (int c = 0; c < 3; c++) er[c] = ex[c];
			rc = s->inf ? -1 : fr_rs(ki, kt);
		}
		else { /* vInvRefl_cr */
			er[0] = -ex[0];
			er[1] = -ex[1];
			er[2] = ex[2];
			rc = s->inf ? 1 : fr_rp(ki, kt, msub);
		}
		const double ph = 2 * wavenum * creal(ki) * s->hsub;
		const dcx f = rc * (cos(ph) + I * sin(ph));
		for (int c = 0; c < 3; c++) {
			kv[c].re = prop[c];
			kv[c].im = 0;
			am[c].re = ex[c];
			am[c].im = 0;
			kv[3 + c].re = c == 2 ? -prop[c] : prop[c];
			kv[3 + c].im = 0;
			am[3 + c].re = creal(f * er[c]);
			am[3 + c].im = cimag(f * er[c]);
		}
		nw = 2;
	}
	if (adda_hip_plane_waves(op, wavenum, gridspace, nw, kv, am, d_einc)) return host_err_hip(__FILE__, __LINE__);
	if (inc_scale) *inc_scale = scale;
	return 0;
}

int adda_ext_cross_gen(adda_hip_op *op, double wavenum, double inc_scale, const void *d_pvec, const void *d_einc,
	double *cext)
{
	adda_cplx d;
	/* sum P.conj(Einc): the conjugating inner product (nDotProd, src/linalg.c:75-88) */
	if (adda_hip_dot(op, ADDA_F64, d_pvec, d_einc, ADDA_DOT_CONJ, &d, NULL)) return host_err_hip(__FILE__, __LINE__);
	*cext = FOUR_PI * wavenum * d.im * inc_scale;
	return 0;
}

/* ---- scattering quantities -------------------------------------------------------------------------------------- */
/* CalcFieldFree, src/crosssec.c:506-595 (DRAINE scattering relation, free space), for K solutions with ndir
 * directions each (K == 1: a single device vector; K > 1: a batched one) */
static int calc_field(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	int K, const double *dirs, int ndir, adda_cplx *ebuff)
{
	const double kd = wavenum * gridspace;
	const size_t nj = (size_t)K * ndir;
	adda_cplx *sums = malloc((nj ? nj : 1) * 3 * sizeof(adda_cplx));
	if (!sums) return host_err(__FILE__, __LINE__, "out of host memory");
	const int rc = K == 1 ? adda_hip_field_sums(op, ADDA_F64, d_pvec, kd, dirs, ndir, sums)
	                      : adda_hip_field_sums_batch(op, ADDA_F64, d_pvec, K, kd, dirs, ndir, sums);
	if (rc) {
		free(sums);
		return host_err_hip(__FILE__, __LINE__);
	}
	/* r0 = box_origin_unif = -d*(cX,cY,cZ), src/make_particle.c:2305-2311 */
	const double r0[3] = {-gridspace * (p->boxX - 1) / 2.0, -gridspace * (p->boxY - 1) / 2.0,
		-gridspace * (p->boxZ - 1) / 2.0};
	const double kkk = wavenum * wavenum * wavenum;
	for (size_t i = 0; i < nj; i++) {
		const double *n = dirs + 3 * i;
		dcx s[3], tb[3];
		for (int c = 0; c < 3; c++) s[c] = sums[3 * i + c].re + I * sums[3 * i + c].im;
		dcx dpr = s[0] * n[0] + s[1] * n[1] + s[2] * n[2];
		for (int c = 0; c < 3; c++) tb[c] = s[c] - dpr * n[c];
		double ph = -wavenum * (r0[0] * n[0] + r0[1] * n[1] + r0[2] * n[2]);
		dcx a = cos(ph) + I * sin(ph);
		dcx tmp = -I * a * kkk;
		for (int c = 0; c < 3; c++) {
			dcx e = tmp * tb[c];
			ebuff[3 * i + c].re = creal(e);
			ebuff[3 * i + c].im = cimag(e);
		}
	}
	free(sums);
	return 0;
}

int adda_calc_field(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	const double *dirs, int ndir, adda_cplx *ebuff)
{
	return calc_field(op, p, wavenum, gridspace, d_pvec, 1, dirs, ndir, ebuff);
}

int adda_calc_field_batch(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace,
	const void *d_pvec, int K, const double *dirs, int ndir, adda_cplx *ebuff)
{
	if (K < 1) return host_err(__FILE__, __LINE__, "batch of %d right-hand sides", K);
	return calc_field(op, p, wavenum, gridspace, d_pvec, K, dirs, ndir, ebuff);
}

/* ExtCross, plane wave: src/crosssec.c:811-823 (K solutions, beam k along props[k] polarized along incpols[k]) */
static int ext_cross(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	int K, const double *props, const double *incpols, double *cext)
{
	adda_cplx *e = malloc((size_t)K * 3 * sizeof(adda_cplx));
	if (!e) return host_err(__FILE__, __LINE__, "out of host memory");
	const int rc = calc_field(op, p, wavenum, gridspace, d_pvec, K, props, 1, e);
	if (!rc)
		for (int k = 0; k < K; k++) {
			const adda_cplx *ek = e + 3 * k;
			const double *pol = incpols + 3 * k;
			double sum = ek[0].re * pol[0] + ek[1].re * pol[1] + ek[2].re * pol[2];
			cext[k] = sum * FOUR_PI / (wavenum * wavenum);
		}
	free(e);
	return rc;
}

int adda_ext_cross(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	const double prop[3], const double incpol[3], double *cext)
{
	return ext_cross(op, p, wavenum, gridspace, d_pvec, 1, prop, incpol, cext);
}

int adda_ext_cross_batch(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace,
	const void *d_pvec, int K, const double *props, const double *incpols, double *cext)
{
	if (K < 1) return host_err(__FILE__, __LINE__, "batch of %d right-hand sides", K);
	return ext_cross(op, p, wavenum, gridspace, d_pvec, K, props, incpols, cext);
}

/* AbsCross, SQ_DRAINE: src/crosssec.c:873-884,916 (K solutions, coupling constants cc[k][nmat][3]) */
static int abs_cross(adda_hip_op *op, const adda_particle *p, double wavenum, const adda_cplx *cc, const void *d_pvec,
	int K, double *cabs)
{
	const double temp1 = 2 * wavenum * wavenum * wavenum / 3;
	const size_t nw = (size_t)K * p->nmat * 3;
	double *w = malloc(nw * sizeof(double));
	if (!w) return host_err(__FILE__, __LINE__, "out of host memory");
	for (size_t i = 0; i < nw; i++) {
		dcx c = cc[i].re + I * cc[i].im;
		w[i] = -cimag(1 / c) - temp1;
	}
	int rc = K == 1 ? adda_hip_weighted_norm2(op, ADDA_F64, d_pvec, w, cabs)
	                : adda_hip_weighted_norm2_batch(op, ADDA_F64, d_pvec, K, w, cabs);
	free(w);
	if (rc) return host_err_hip(__FILE__, __LINE__);
	for (int k = 0; k < K; k++) cabs[k] *= FOUR_PI * wavenum;
	return 0;
}

int adda_abs_cross(adda_hip_op *op, const adda_particle *p, double wavenum, const adda_cplx *cc, const void *d_pvec,
	double *cabs)
{
	return abs_cross(op, p, wavenum, cc, d_pvec, 1, cabs);
}

int adda_abs_cross_batch(adda_hip_op *op, const adda_particle *p, double wavenum, const adda_cplx *cc,
	const void *d_pvec, int K, double *cabs)
{
	if (K < 1) return host_err(__FILE__, __LINE__, "batch of %d right-hand sides", K);
	return abs_cross(op, p, wavenum, cc, d_pvec, K, cabs);
}

/* CalcEplaneYZ, CE_NORMAL, no orientation averaging: src/CalculateE.c:423-507 (K solutions, beam frame k given by
 * props[k], pars[k], pers[k]) */
static int eplane_yz(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	int K, const double *props, const double *pars, const double *pers, int ntheta, int npoints, adda_cplx *eplane)
{
	if (ntheta < 2 || npoints < 1 || K < 1) return host_err(__FILE__, __LINE__, "ntheta must be >= 2");
	const double dth = PI / (double)(ntheta - 1); /* dtheta_rad, src/calculator.c:671-672 */
	const size_t nj = (size_t)K * npoints;
	double *dirs = malloc(nj * 3 * sizeof(double));
	adda_cplx *e = malloc(nj * 3 * sizeof(adda_cplx));
	if (!dirs || !e) { free(dirs); free(e); return host_err(__FILE__, __LINE__, "out of host memory"); }
	for (int k = 0; k < K; k++) {
		const double *prop = props + 3 * k, *par = pars + 3 * k;
		for (int i = 0; i < npoints; i++) {
			double th = i * dth, co = cos(th), si = sin(th);
			for (int c = 0; c < 3; c++) dirs[3 * ((size_t)k * npoints + i) + c] = co * prop[c] + si * par[c];
		}
	}
	int rc = calc_field(op, p, wavenum, gridspace, d_pvec, K, dirs, npoints, e);
	if (!rc)
		for (int k = 0; k < K; k++) {
			const double *prop = props + 3 * k, *par = pars + 3 * k, *per = pers + 3 * k;
			for (int i = 0; i < npoints; i++) {
				const size_t j = (size_t)k * npoints + i;
				double th = i * dth, co = cos(th), si = sin(th), epar[3];
				for (int c = 0; c < 3; c++) epar[c] = -si * prop[c] + co * par[c];
				dcx eper = 0, epa = 0;
				for (int c = 0; c < 3; c++) {
					dcx ec = e[3 * j + c].re + I * e[3 * j + c].im;
					eper += ec * per[c];
					epa += ec * epar[c];
				}
				eplane[2 * j].re = creal(eper);
				eplane[2 * j].im = cimag(eper);
				eplane[2 * j + 1].re = creal(epa);
				eplane[2 * j + 1].im = cimag(epa);
			}
		}
	free(dirs);
	free(e);
	return rc;
}

int adda_eplane_yz(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	const double prop[3], const double par[3], const double per[3], int ntheta, int npoints, adda_cplx *eplane)
{
	return eplane_yz(op, p, wavenum, gridspace, d_pvec, 1, prop, par, per, ntheta, npoints, eplane);
}

int adda_eplane_yz_batch(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace,
	const void *d_pvec, int K, const double *props, const double *pars, const double *pers, int ntheta, int npoints,
	adda_cplx *eplane)
{
	if (K < 1) return host_err(__FILE__, __LINE__, "batch of %d right-hand sides", K);
	return eplane_yz(op, p, wavenum, gridspace, d_pvec, K, props, pars, pers, ntheta, npoints, eplane);
}

static inline double ca2(dcx a) { return creal(a) * creal(a) + cimag(a) * cimag(a); }

/* ComputeMuellerMatrix (src/CalculateE.c:83-116) with the yz-plane amplitudes of MuellerMatrix (236-246) */
void adda_mueller_yz(const adda_cplx *eX, const adda_cplx *eY, int npoints, double *out)
{
	for (int i = 0; i < npoints; i++) {
		dcx s1 = eX[2 * i].re + I * eX[2 * i].im, s2 = eY[2 * i + 1].re + I * eY[2 * i + 1].im;
		dcx s3 = eX[2 * i + 1].re + I * eX[2 * i + 1].im, s4 = eY[2 * i].re + I * eY[2 * i].im;
		double *m = out + 16 * i;
		m[0] = 0.5 * (ca2(s1) + ca2(s2) + ca2(s3) + ca2(s4));
		m[1] = 0.5 * (ca2(s2) - ca2(s1) + ca2(s4) - ca2(s3));
		m[2] = creal(s2 * conj(s3) + s1 * conj(s4));
		m[3] = cimag(s2 * conj(s3) - s1 * conj(s4));
		m[4] = 0.5 * (ca2(s2) - ca2(s1) + ca2(s3) - ca2(s4));
		m[5] = 0.5 * (ca2(s2) + ca2(s1) - ca2(s3) - ca2(s4));
		m[6] = creal(s2 * conj(s3) - s1 * conj(s4));
		m[7] = cimag(s2 * conj(s3) + s1 * conj(s4));
		m[8] = creal(s2 * conj(s4) + s1 * conj(s3));
		m[9] = creal(s2 * conj(s4) - s1 * conj(s3));
		m[10] = creal(s1 * conj(s2) + s3 * conj(s4));
		m[11] = cimag(s2 * conj(s1) + s4 * conj(s3));
		m[12] = cimag(s4 * conj(s2) + s1 * conj(s3));
		m[13] = cimag(s4 * conj(s2) - s1 * conj(s3));
		m[14] = cimag(s1 * conj(s2) - s3 * conj(s4));
		m[15] = creal(s1 * conj(s2) - s3 * conj(s4));
	}
}
