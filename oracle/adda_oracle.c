/* oracle/adda_oracle.c -- CPU FP64 restatement of ADDA's FFT MatVec + Krylov path (reference: ADDA 1.3b6).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library. It is written for clarity, not speed: full-grid 3D FFTs (no
 * zero-plane skipping), a textbook recursive mixed-radix FFT, plain loops.
 *
 * Pinning summary (tests/test_oracle.py):
 *   or_fftfit            - grid sizes of the reference sample (box 16 -> grid 32) + rule enumeration
 *   or_green_poi         - closed form at hand-computed offsets; SPARSE reference runs (whole-solve observables)
 *   or_fft1d             - brute-force DFT on sizes 1..40 and composite sizes, both signs
 *   or_init_dmatrix      - FFT matvec == direct-sum matvec (brute force) for random vectors; reduced == non-reduced
 *   or_matvec_fft/direct - same identity; her == conj(A conj x); A symmetric (x.Ay == y.Ax)
 *   or_couple_constant   - sample log KAT (CoupleConstant 0.005259037197+1.843854148e-05i); CM closed form
 *   or_iterative_solver  - converged Cext/Cabs/Mueller/internal fields vs the reference SPARSE build (golden files)
 *   or_ext/abs_cross, or_eplane_yz, or_mueller_yz - sample golden CrossSec-Y/mueller + SPARSE goldens
 *   or_refl_img          - closed form of the image dipole's field at hand-computed offsets; SPARSE -surf goldens
 *   or_init_rmatrix, surface MatVec - FFT MatVec == direct sum with the reflected term; A symmetric; SPARSE -surf
 *   or_generate_b_surf, or_ext_cross_gen - SPARSE -surf goldens (from above and below, finite and perfect reflector)
 */
#include "adda_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex dcx;

#define TWO_PI 6.283185307179586476925286766559
#define FOUR_PI 12.566370614359172953850573533118
/* LDR coefficients, src/const.h:234-236 */
#define LDR_B1 1.8915316
#define LDR_B2 -0.1648469
#define LDR_B3 1.7700004

static inline dcx C(or_cplx a) { return a.re + I * a.im; }
static inline or_cplx O(dcx a) { or_cplx r = {creal(a), cimag(a)}; return r; }
static void *xmalloc(size_t n) { void *p = calloc(n ? n : 1, 1); if (!p) abort(); return p; }

/*====================================================================================================================*/
/* fftFit: src/fft.c:528-562 (the weird_nprocs branch is irrelevant for one process) */
int or_fftfit(int x, int divis, int rule)
{
	for (;; x++) {
		if (x % 2 != 0 || x % divis != 0) continue;
		int y = x;
		while (y % 2 == 0) y /= 2;
		while (y % 3 == 0) y /= 3;
		while (y % 5 == 0) y /= 5;
		if (rule == 0) {
			while (y % 7 == 0) y /= 7;
			if (y % 11 == 0) y /= 11;
			else if (y % 13 == 0) y /= 13;
		}
		if (y == 1) return x;
	}
}

/*====================================================================================================================*/
/* InterTerm_poi via InterParams + InterTerm_core (non-SSE branch): src/interaction.c:159-180,239-266,290-300 */
void or_green_poi(int i, int j, int k, double wavenum, double gridspace, or_cplx res[6])
{
	double q[3] = {i, j, k};
	double rn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
	double invrn = 1.0 / rn;
	for (int c = 0; c < 3; c++) q[c] *= invrn;
	double rr = rn * gridspace;
	double invr3 = 1 / (rr * rr * rr);
	double kr = wavenum * rr, kr2 = kr * kr;
	double qmunu[6] = {q[0] * q[0], q[0] * q[1], q[0] * q[2], q[1] * q[1], q[1] * q[2], q[2] * q[2]};
	double t1 = 3 - kr2, t2 = -3 * kr, t3 = kr2 - 1;
	dcx expval = invr3 * (cos(kr) + I * sin(kr));
	for (int c = 0; c < 6; c++) {
		dcx r;
		if (c == 0 || c == 3 || c == 5) r = ((t1 * qmunu[c] + t3) + I * (kr + t2 * qmunu[c])) * expval;
		else r = (t1 + I * t2) * qmunu[c] * expval;
		res[c] = O(r);
	}
}

/*====================================================================================================================*/
/* Textbook recursive mixed-radix DFT (decimation in time over the smallest prime factor). Unnormalized;
 * sign -1 forward, +1 backward, as FFT_FORWARD/FFT_BACKWARD in src/fft.h:31-32. */
static void fft_rec(const dcx *in, size_t is, dcx *out, size_t n, const dcx *tw, size_t ntop, dcx *tmp)
{
	if (n == 1) { out[0] = in[0]; return; }
	size_t p = 2;
	while (n % p) p++;
	size_t m = n / p, step = ntop / n;
	for (size_t q = 0; q < p; q++) fft_rec(in + q * is, is * p, out + q * m, m, tw, ntop, tmp);
	/* combine: y[k + s*m] = sum_q w_n^{q(k+s m)} out[q*m + k], w_n^e = tw[e*(ntop/n)] */
	for (size_t k = 0; k < m; k++) {
		for (size_t q = 0; q < p; q++) tmp[q] = out[q * m + k];
		for (size_t s = 0; s < p; s++) {
			dcx acc = 0;
			for (size_t q = 0; q < p; q++) acc += tmp[q] * tw[((q * (k + s * m)) % n) * step];
			out[k + s * m] = acc;
		}
	}
}

void or_fft1d(or_cplx *data, size_t n, size_t stride, size_t howmany, size_t dist, int sign)
{
	dcx *line = xmalloc(n * sizeof(dcx)), *out = xmalloc(n * sizeof(dcx)), *tmp = xmalloc(n * sizeof(dcx));
	dcx *tw = xmalloc(n * sizeof(dcx));
	for (size_t e = 0; e < n; e++) {
		double ang = sign * TWO_PI * (double)e / (double)n;
		tw[e] = cos(ang) + I * sin(ang);
	}
	for (size_t h = 0; h < howmany; h++) {
		or_cplx *d = data + h * dist;
		for (size_t i = 0; i < n; i++) line[i] = C(d[i * stride]);
		fft_rec(line, 1, out, n, tw, n, tmp);
		for (size_t i = 0; i < n; i++) d[i * stride] = O(out[i]);
	}
	free(line); free(out); free(tmp); free(tw);
}

/* 3D FFT of one component stored [z][y][x] (x fastest) */
static void fft3d(or_cplx *a, size_t gx, size_t gy, size_t gz, int sign)
{
	or_fft1d(a, gx, 1, gy * gz, gx, sign);                              /* along x */
	for (size_t z = 0; z < gz; z++) or_fft1d(a + z * gx * gy, gy, gx, gx, 1, sign); /* along y */
	or_fft1d(a, gz, gx * gy, gx * gy, 1, sign);                         /* along z */
}

/*====================================================================================================================*/
/* D-matrix sizes: src/fft.c:955-970 */
static void dsizes(const or_grid *g, size_t *dy, size_t *dz)
{
	if (g->reduced_fft) { *dy = g->gridY / 2 + 1; *dz = g->gridZ / 2 + 1; }
	else { *dy = g->gridY; *dz = g->gridZ; }
}

size_t or_dmatrix_size(const or_grid *g)
{
	size_t dy, dz;
	dsizes(g, &dy, &dz);
	return 6 * (size_t)g->gridX * dy * dz;
}

/* InitDmatrix: src/fft.c:918-1416. G is placed on the periodic grid for all offsets |i|<boxX, |j|<boxY, |k|<boxZ
 * except 0 (src/fft.c:1226-1238; the mirror rules of 1275-1290 are the parity of G under reflection, which direct
 * evaluation reproduces exactly), transformed by a forward 3D FFT and stored as -G~/Ngrid with IndexDmatrix
 * (src/fft.c:144-150, 1310-1314). */
void or_init_dmatrix(const or_grid *g, or_cplx *dmat)
{
	size_t gx = g->gridX, gy = g->gridY, gz = g->gridZ, n = gx * gy * gz;
	size_t dy, dz;
	dsizes(g, &dy, &dz);
	double invN = 1.0 / (double)n;
	or_cplx *gt = xmalloc(6 * n * sizeof(or_cplx));
	for (int k = 1 - g->boxZ; k < g->boxZ; k++)
		for (int j = 1 - g->boxY; j < g->boxY; j++)
			for (int i = 1 - g->boxX; i < g->boxX; i++) {
				if (i == 0 && j == 0 && k == 0) continue;
				or_cplx r[6];
				or_green_poi(i, j, k, g->wavenum, g->gridspace, r);
				size_t x = (size_t)((i + (int)gx) % (int)gx), y = (size_t)((j + (int)gy) % (int)gy),
					z = (size_t)((k + (int)gz) % (int)gz);
				for (int c = 0; c < 6; c++) gt[c * n + (z * gy + y) * gx + x] = r[c];
			}
	for (int c = 0; c < 6; c++) {
		or_cplx *a = gt + c * n;
		fft3d(a, gx, gy, gz, -1);
		for (size_t x = 0; x < gx; x++)
			for (size_t z = 0; z < dz; z++)
				for (size_t y = 0; y < dy; y++) {
					dcx v = -invN * C(a[(z * gy + y) * gx + x]);
					dmat[6 * ((x * dz + z) * dy + y) + c] = O(v);
				}
	}
	free(gt);
}

/*====================================================================================================================*/
/* ReflTerm_img via vCopyIntRealShift, ReflParams and ReflTerm_core: src/interaction.c:1081-1153. The image of the
 * source lies at the mirror height, so the separation is (i, j, ksum + zsum_shift); the result is InterTerm_core's
 * tensor times rcn with the xz, yz and zz components negated. */
void or_refl_img(int i, int j, int ksum, double zsum_shift, or_cplx rcn, double wavenum, double gridspace,
	or_cplx res[6])
{
	double q[3] = {i, j, ksum + zsum_shift};
	double rn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
	for (int c = 0; c < 3; c++) q[c] /= rn;
	double rr = rn * gridspace;
	double invr3 = 1 / (rr * rr * rr);
	double kr = wavenum * rr, kr2 = kr * kr;
	double qmunu[6] = {q[0] * q[0], q[0] * q[1], q[0] * q[2], q[1] * q[1], q[1] * q[2], q[2] * q[2]};
	double t1 = 3 - kr2, t2 = -3 * kr, t3 = kr2 - 1;
	dcx scale = C(rcn) * invr3 * (cos(kr) + I * sin(kr));
	for (int c = 0; c < 6; c++) {
		dcx r;
		if (c == 0 || c == 3 || c == 5) r = ((t1 * qmunu[c] + t3) + I * (kr + t2 * qmunu[c])) * scale;
		else r = (t1 + I * t2) * qmunu[c] * scale;
		if (c == 2 || c == 4 || c == 5) r = -r;
		res[c] = O(r);
	}
}

static size_t rsize_y(const or_grid *g) { return g->reduced_fft ? (size_t)g->gridY / 2 + 1 : (size_t)g->gridY; }

size_t or_rmatrix_size(const or_grid *g) { return 6 * (size_t)g->gridX * g->gridZ * rsize_y(g); }

/* InitRmatrix: src/fft.c:819-891. GR is placed on the periodic grid for |i| < boxX, |j| < boxY (the y mirror of
 * 864-873 is GR's parity in y, which direct evaluation reproduces) and height sums 0 <= k <= 2*boxZ-2 (not periodic:
 * gridZ >= 2*boxZ), forward 3D FFT, stored as -GR~/Ngrid with IndexRmatrix_mv's layout. */
void or_init_rmatrix(const or_grid *g, or_cplx *rmat)
{
	const or_surface *sf = g->surf;
	size_t gx = g->gridX, gy = g->gridY, gz = g->gridZ, n = gx * gy * gz, ry = rsize_y(g);
	double invN = 1.0 / (double)n;
	or_cplx *gt = xmalloc(6 * n * sizeof(or_cplx));
	for (int k = 0; k <= 2 * g->boxZ - 2; k++)
		for (int j = 1 - g->boxY; j < g->boxY; j++)
			for (int i = 1 - g->boxX; i < g->boxX; i++) {
				or_cplx r[6];
				or_refl_img(i, j, k, sf->zsum_shift, sf->rcn, g->wavenum, g->gridspace, r);
				size_t x = (size_t)((i + (int)gx) % (int)gx), y = (size_t)((j + (int)gy) % (int)gy), z = (size_t)k;
				for (int c = 0; c < 6; c++) gt[c * n + (z * gy + y) * gx + x] = r[c];
			}
	for (int c = 0; c < 6; c++) {
		or_cplx *a = gt + c * n;
		fft3d(a, gx, gy, gz, -1);
		for (size_t x = 0; x < gx; x++)
			for (size_t z = 0; z < gz; z++)
				for (size_t y = 0; y < ry; y++)
					rmat[6 * ((x * gz + z) * ry + y) + c] = O(-invN * C(a[(z * gy + y) * gx + x]));
	}
	free(gt);
}

/*====================================================================================================================*/
/* cSymMatrVec: src/cmplx.h:369-376 */
static inline void sym_mv(const dcx m[6], const dcx x[3], dcx y[3])
{
	y[0] = m[0] * x[0] + m[1] * x[1] + m[2] * x[2];
	y[1] = m[1] * x[0] + m[3] * x[1] + m[4] * x[2];
	y[2] = m[2] * x[0] + m[4] * x[1] + m[5] * x[2];
}

/* cReflMatrVec: src/cmplx.h:391-401 (M21 = M12, M31 = -M13, M32 = -M23) */
static inline void refl_mv(const dcx m[6], const dcx x[3], dcx y[3])
{
	y[0] = m[0] * x[0] + m[1] * x[1] + m[2] * x[2];
	y[1] = m[1] * x[0] + m[3] * x[1] + m[4] * x[2];
	y[2] = -m[2] * x[0] - m[4] * x[1] + m[5] * x[2];
}

/* MatVec: src/matvec.c:128-425 (FFT mode, one process, no surface) */
void or_matvec_fft(const or_grid *g, const or_cplx *dmat, const or_cplx *cc_sqrt, const or_cplx *arg, or_cplx *res,
	double *inprod, int her)
{
	size_t gx = g->gridX, gy = g->gridY, gz = g->gridZ, n = gx * gy * gz;
	size_t dy, dz;
	dsizes(g, &dy, &dz);
	int transposed = (!g->reduced_fft) && her; /* src/matvec.c:181 */
	or_cplx *X = xmalloc(3 * n * sizeof(or_cplx));
	/* scatter S.x (conj for her): src/matvec.c:203-212 */
	for (size_t i = 0; i < g->nd; i++) {
		const unsigned short *p = g->pos + 3 * i;
		size_t idx = ((size_t)p[2] * gy + p[1]) * gx + p[0];
		for (int c = 0; c < 3; c++) {
			dcx a = C(arg[3 * i + c]);
			if (her) a = conj(a);
			X[c * n + idx] = O(C(cc_sqrt[3 * g->mat[i] + c]) * a);
		}
	}
	/* the reflected term transforms a copy with the z direction backward: src/matvec.c:249, src/fft.c:383-415 */
	or_cplx *XR = NULL;
	size_t ry = rsize_y(g);
	if (g->surf) {
		XR = xmalloc(3 * n * sizeof(or_cplx));
		memcpy(XR, X, 3 * n * sizeof(or_cplx));
		for (int c = 0; c < 3; c++) {
			or_cplx *a = XR + c * n;
			or_fft1d(a, gx, 1, gy * gz, gx, -1);
			for (size_t z = 0; z < gz; z++) or_fft1d(a + z * gx * gy, gy, gx, gx, 1, -1);
			or_fft1d(a, gz, gx * gy, gx * gy, 1, +1);
		}
	}
	for (int c = 0; c < 3; c++) fft3d(X + c * n, gx, gy, gz, -1);
	/* D~ product with reduced-FFT sign flips: src/matvec.c:270-305, IndexDmatrix_mv 87-103 */
	for (size_t z = 0; z < gz; z++)
		for (size_t y = 0; y < gy; y++)
			for (size_t x = 0; x < gx; x++) {
				size_t xi = x, yi = y, zi = z;
				if (transposed) {
					if (xi > 0) xi = gx - xi;
					if (yi > 0) yi = gy - yi;
					if (zi > 0) zi = gz - zi;
				}
				else {
					if (yi >= dy) yi = gy - yi;
					if (zi >= dz) zi = gz - zi;
				}
				const or_cplx *dm = dmat + 6 * ((xi * dz + zi) * dy + yi);
				dcx f[6], xv[3], yv[3];
				for (int c = 0; c < 6; c++) f[c] = C(dm[c]);
				if (g->reduced_fft) {
					if (y >= dy) {
						f[1] = -f[1];
						if (z >= dz) f[2] = -f[2];
						else f[4] = -f[4];
					}
					else if (z >= dz) {
						f[2] = -f[2];
						f[4] = -f[4];
					}
				}
				size_t idx = (z * gy + y) * gx + x;
				for (int c = 0; c < 3; c++) xv[c] = C(X[c * n + idx]);
				sym_mv(f, xv, yv);
				if (XR) { /* src/matvec.c:288-302, IndexRmatrix_mv 107-121 */
					size_t xr = x, yr = y;
					if (transposed) {
						if (xr > 0) xr = gx - xr;
						if (yr > 0) yr = gy - yr;
					}
					else if (yr >= ry) yr = gy - yr;
					const or_cplx *rm = g->surf->rmat + 6 * ((xr * gz + z) * ry + yr);
					dcx r[6], xr3[3], yr3[3];
					for (int c = 0; c < 6; c++) r[c] = C(rm[c]);
					if (g->reduced_fft && y >= ry) { r[1] = -r[1]; r[4] = -r[4]; }
					if (transposed) { r[2] = -r[2]; r[4] = -r[4]; }
					for (int c = 0; c < 3; c++) xr3[c] = C(XR[c * n + idx]);
					refl_mv(r, xr3, yr3);
					for (int c = 0; c < 3; c++) yv[c] += yr3[c];
				}
				for (int c = 0; c < 3; c++) X[c * n + idx] = O(yv[c]);
			}
	for (int c = 0; c < 3; c++) fft3d(X + c * n, gx, gy, gz, +1);
	/* gather: src/matvec.c:352-364 */
	double s = 0;
	for (size_t i = 0; i < g->nd; i++) {
		const unsigned short *p = g->pos + 3 * i;
		size_t idx = ((size_t)p[2] * gy + p[1]) * gx + p[0];
		for (int c = 0; c < 3; c++) {
			dcx a = C(arg[3 * i + c]);
			if (her) a = conj(a);
			dcx r = a + C(cc_sqrt[3 * g->mat[i] + c]) * C(X[c * n + idx]);
			s += creal(r) * creal(r) + cimag(r) * cimag(r);
			if (her) r = conj(r);
			res[3 * i + c] = O(r);
		}
	}
	if (inprod) *inprod = s;
	free(X);
	free(XR);
}

/* SPARSE MatVec: src/matvec.c:434-465 with AijProd/DiagProd/CcMul of src/sparse_ops.h (non-SSE branch) */
void or_matvec_direct(const or_grid *g, const or_cplx *cc_sqrt, const or_cplx *arg, or_cplx *res, double *inprod,
	int her)
{
	size_t nd = g->nd;
	dcx *sx = xmalloc(3 * nd * sizeof(dcx));
	for (size_t j = 0; j < nd; j++)
		for (int c = 0; c < 3; c++) {
			dcx a = C(arg[3 * j + c]);
			if (her) a = conj(a);
			sx[3 * j + c] = a * C(cc_sqrt[3 * g->mat[j] + c]);
		}
	double s = 0;
	for (size_t i = 0; i < nd; i++) {
		dcx acc[3] = {0, 0, 0};
		const unsigned short *pi = g->pos + 3 * i;
		for (size_t j = 0; j < nd; j++) {
			const unsigned short *pj = g->pos + 3 * j;
			or_cplx t[6];
			dcx f[6], yv[3];
			if (j != i) {
				or_green_poi((int)pi[0] - pj[0], (int)pi[1] - pj[1], (int)pi[2] - pj[2], g->wavenum, g->gridspace, t);
				for (int c = 0; c < 6; c++) f[c] = C(t[c]);
				sym_mv(f, sx + 3 * j, yv);
				for (int c = 0; c < 3; c++) acc[c] += yv[c];
			}
			if (g->surf) { /* src/sparse_ops.h:160-165: always, the dipole's own image included */
				or_refl_img((int)pi[0] - pj[0], (int)pi[1] - pj[1], (int)pi[2] + pj[2], g->surf->zsum_shift,
					g->surf->rcn, g->wavenum, g->gridspace, t);
				for (int c = 0; c < 6; c++) f[c] = C(t[c]);
				refl_mv(f, sx + 3 * j, yv);
				for (int c = 0; c < 3; c++) acc[c] += yv[c];
			}
		}
		for (int c = 0; c < 3; c++) {
			dcx a = C(arg[3 * i + c]);
			if (her) a = conj(a);
			dcx r = a - acc[c] * C(cc_sqrt[3 * g->mat[i] + c]);
			s += creal(r) * creal(r) + cimag(r) * cimag(r);
			if (her) r = conj(r);
			res[3 * i + c] = O(r);
		}
	}
	if (inprod) *inprod = s;
	free(sx);
}

/*====================================================================================================================*/
/* polCM / polM / polMplusRR / pol3coef: src/calculator.c:100-133; LDR case of CoupleConstant 214-221 */
or_cplx or_couple_constant(or_cplx mo, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3])
{
	dcx m = C(mo), m2 = m * m;
	double dipvol = gridspace * gridspace * gridspace, kd = wavenum * gridspace;
	dcx cm = (3 / FOUR_PI) * dipvol * (m2 - 1) / (m2 + 2);
	dcx M;
	if (polrel == 1) return O(cm);                        /* CM */
	if (polrel == 2) M = 0;                               /* RRC: polMplusRR(0,m) */
	else {                                                /* LDR */
		double S = 0;
		for (int i = 0; i < 3; i++) S += prop[i] * prop[i] * incpol[i] * incpol[i];
		M = (LDR_B1 + (LDR_B2 + LDR_B3 * S) * m2) * kd * kd;
	}
	M += I * 2 * kd * kd * kd / 3;
	return O(cm / (1 - (cm / dipvol) * M));
}

/* GenerateB, plane wave: src/GenerateB.c:324-330; DipoleCoord src/make_particle.c:2266-2275 */
void or_generate_b(const or_grid *g, const double prop[3], const double incpol[3], or_cplx *einc)
{
	double cx = (g->boxX - 1) / 2.0, cy = (g->boxY - 1) / 2.0, cz = (g->boxZ - 1) / 2.0;
	for (size_t i = 0; i < g->nd; i++) {
		const unsigned short *p = g->pos + 3 * i;
		double r[3] = {(p[0] - cx) * g->gridspace, (p[1] - cy) * g->gridspace, (p[2] - cz) * g->gridspace};
		double ph = g->wavenum * (r[0] * prop[0] + r[1] * prop[1] + r[2] * prop[2]);
		dcx e = cos(ph) + I * sin(ph);
		for (int c = 0; c < 3; c++) einc[3 * i + c] = O(e * incpol[c]);
	}
}

/* cSqrtCut: src/cmplx.h:58-82 (branch cut not depending on the sign of a zero imaginary part) */
static dcx csqrt_cut(dcx a)
{
	if (cimag(a) == 0) return creal(a) >= 0 ? sqrt(creal(a)) : I * sqrt(-creal(a));
	return csqrt(a);
}

/* Fresnel coefficients: src/cmplx.h:672-709 */
static dcx fresnel_rs(dcx ki, dcx kt) { return (ki - kt) / (ki + kt); }
static dcx fresnel_ts(dcx ki, dcx kt) { return 2 * ki / (ki + kt); }
static dcx fresnel_rp(dcx ki, dcx kt, dcx mr) { return (mr * mr * ki - kt) / (mr * mr * ki + kt); }
static dcx fresnel_tp(dcx ki, dcx kt, dcx mr) { return 2 * mr * ki / (mr * mr * ki + kt); }

/* InitBeam (B_PLANE with surface, src/GenerateB.c:96-137) and GenerateB (src/GenerateB.c:258-322) */
double or_generate_b_surf(const or_grid *g, or_cplx msub_o, int msub_inf, double hsub, const double prop[3],
	const double incpol_x[3], const double incpol_y[3], int which_x, or_cplx *einc)
{
	const dcx msub = C(msub_o);
	const double k = g->wavenum;
	double inc_scale = 1;
	dcx ki, kt = 0, ktv[3] = {0, 0, 0};
	const double pp = prop[0] * prop[0] + prop[1] * prop[1];
	if (prop[2] > 0) { /* from the substrate */
		inc_scale = 1 / creal(msub);
		ki = msub * prop[2];
		if (cabs(msub - 1) < 1e-15 && cabs(ki) < 3e-8) kt = ki;
		else kt = csqrt_cut(1 - msub * msub * pp);
		ktv[0] = msub * prop[0];
		ktv[1] = msub * prop[1];
		ktv[2] = kt;
	}
	else {
		ki = -prop[2];
		if (!msub_inf) {
			if (cabs(msub - 1) < 1e-15 && cabs(ki) < 3e-8) kt = ki;
			else kt = csqrt_cut(msub * msub - pp);
			ktv[0] = prop[0];
			ktv[1] = prop[1];
			ktv[2] = -kt;
		}
	}
	const double prefl[3] = {prop[0], prop[1], -prop[2]}; /* vRefl */
	/* ex = incPolY, ey = -incPolX for the Y polarization; ex = incPolX, ey = incPolY for X (src/GenerateB.c:247-254) */
	double ex[3], ey[3];
	for (int c = 0; c < 3; c++) {
		ex[c] = which_x ? incpol_x[c] : incpol_y[c];
		ey[c] = which_x ? incpol_y[c] : -incpol_x[c];
	}
	dcx erefl[3], etran[3], rc, tc = 0;
	if (prop[2] > 0) {
		if (!which_x) { /* s-polarized */
			for (int c = 0; c < 3; c++) erefl[c] = etran[c] = ex[c];
			rc = fresnel_rs(ki, kt);
			tc = fresnel_ts(ki, kt);
		}
		else {
			erefl[0] = -ex[0]; erefl[1] = -ex[1]; erefl[2] = ex[2];          /* vInvRefl_cr */
			etran[0] = ey[1] * ktv[2] - ey[2] * ktv[1];                       /* crCrossProd(ey, ktVec) */
			etran[1] = ey[2] * ktv[0] - ey[0] * ktv[2];
			etran[2] = ey[0] * ktv[1] - ey[1] * ktv[0];
			rc = fresnel_rp(ki, kt, 1 / msub);
			tc = fresnel_tp(ki, kt, 1 / msub);
		}
		dcx st = tc * cexp(I * k * (kt - ki) * hsub);
		for (int c = 0; c < 3; c++) etran[c] *= st;
		for (size_t i = 0; i < g->nd; i++) {
			const unsigned short *p = g->pos + 3 * i;
			double r[3] = {(p[0] - (g->boxX - 1) / 2.0) * g->gridspace, (p[1] - (g->boxY - 1) / 2.0) * g->gridspace,
				(p[2] - (g->boxZ - 1) / 2.0) * g->gridspace};
			dcx e = cexp(I * k * (ktv[0] * r[0] + ktv[1] * r[1] + ktv[2] * r[2]));
			for (int c = 0; c < 3; c++) einc[3 * i + c] = O(e * etran[c]);
		}
	}
	else {
		if (!which_x) {
			for (int c = 0; c < 3; c++) erefl[c] = ex[c];
			rc = msub_inf ? -1 : fresnel_rs(ki, kt);
		}
		else {
			erefl[0] = -ex[0]; erefl[1] = -ex[1]; erefl[2] = ex[2];
			rc = msub_inf ? 1 : fresnel_rp(ki, kt, msub);
		}
		double ph = 2 * k * creal(ki) * hsub;
		dcx sr = rc * (cos(ph) + I * sin(ph));
		for (int c = 0; c < 3; c++) erefl[c] *= sr;
		for (size_t i = 0; i < g->nd; i++) {
			const unsigned short *p = g->pos + 3 * i;
			double r[3] = {(p[0] - (g->boxX - 1) / 2.0) * g->gridspace, (p[1] - (g->boxY - 1) / 2.0) * g->gridspace,
				(p[2] - (g->boxZ - 1) / 2.0) * g->gridspace};
			double p1 = k * (r[0] * prop[0] + r[1] * prop[1] + r[2] * prop[2]);
			double p2 = k * (r[0] * prefl[0] + r[1] * prefl[1] + r[2] * prefl[2]);
			dcx e1 = cos(p1) + I * sin(p1), e2 = cos(p2) + I * sin(p2);
			for (int c = 0; c < 3; c++) einc[3 * i + c] = O(e1 * ex[c] + e2 * erefl[c]);
		}
	}
	return inc_scale;
}

/* ExtCross, general branch: src/crosssec.c:824-832 (cDotProd_Im = Im(P.conj(Einc))) */
double or_ext_cross_gen(const or_grid *g, const or_cplx *pvec, const or_cplx *einc)
{
	double s = 0;
	for (size_t i = 0; i < 3 * g->nd; i++) s += cimag(C(pvec[i]) * conj(C(einc[i])));
	return FOUR_PI * g->wavenum * s;
}

/*====================================================================================================================*/
/* Krylov solvers: src/iterative.c. Vector kernels restate src/linalg.c. */
typedef struct {
	const or_grid *g;
	const or_cplx *dmat, *cc_sqrt;
	size_t n;       /* local_nRows */
	int nmv;
} solv;

static void mv(solv *s, const dcx *x, dcx *y, double *inprod, int her)
{
	if (s->dmat) or_matvec_fft(s->g, s->dmat, s->cc_sqrt, (const or_cplx *)x, (or_cplx *)y, inprod, her);
	else or_matvec_direct(s->g, s->cc_sqrt, (const or_cplx *)x, (or_cplx *)y, inprod, her);
	s->nmv++;
}
static double norm2(const dcx *a, size_t n) { double s = 0; for (size_t i = 0; i < n; i++) s += creal(a[i]) * creal(a[i]) + cimag(a[i]) * cimag(a[i]); return s; }
static dcx dot(const dcx *a, const dcx *b, size_t n) { dcx s = 0; for (size_t i = 0; i < n; i++) s += a[i] * conj(b[i]); return s; }   /* nDotProd */
static dcx dotc(const dcx *a, const dcx *b, size_t n) { dcx s = 0; for (size_t i = 0; i < n; i++) s += a[i] * b[i]; return s; }        /* nDotProd_conj */

int or_iterative_solver(const or_grid *g, const or_cplx *dmat, const or_cplx *cc_sqrt, const or_cplx *einc_o,
	int method, double eps, int maxiter, int init_field, or_cplx *x_o, or_cplx *pvec_o, double *resid_hist,
	int hist_len, int *niter_out, int *nmatvec_out)
{
	solv S = {g, dmat, cc_sqrt, 3 * g->nd, 0};
	size_t n = S.n;
	const dcx *einc = (const dcx *)einc_o;
	dcx *x = (dcx *)x_o;
	dcx *b = xmalloc(n * sizeof(dcx)), *r = xmalloc(n * sizeof(dcx)), *Av = xmalloc(n * sizeof(dcx));
	dcx *p = xmalloc(n * sizeof(dcx)), *v1 = xmalloc(n * sizeof(dcx)), *v2 = xmalloc(n * sizeof(dcx));
	dcx *v3 = xmalloc(n * sizeof(dcx)), *v4 = xmalloc(n * sizeof(dcx));
	const int mc_tab[7] = {50000, 30000, 10, 50000, 10, 15000, 50000}; /* params[].mc, src/iterative.c:118-132 */
	int mc = mc_tab[method];
	int rc = 0;
	/* b = S.Einc (nMult_mat), src/iterative.c:1421-1425 */
	for (size_t i = 0; i < g->nd; i++)
		for (int c = 0; c < 3; c++) b[3 * i + c] = C(cc_sqrt[3 * g->mat[i] + c]) * einc[3 * i + c];
	double temp = norm2(b, n);
	double resid_scale = 1 / temp, epsB = eps * eps * temp, inprodR, inprodRp1 = 0;
	int matvec_ready = 0;
	/* CalcInitField, src/iterative.c:1342-1394 */
	if (init_field == 0) {
		mv(&S, b, Av, NULL, 0);
		for (size_t i = 0; i < n; i++) r[i] = b[i] - Av[i];
		inprodR = norm2(r, n);
		if (temp < inprodR) {
			memset(x, 0, n * sizeof(dcx));
			memcpy(r, b, n * sizeof(dcx));
			inprodR = temp;
			matvec_ready = 1;
		}
		else memcpy(x, b, n * sizeof(dcx));
	}
	else if (init_field == 1) {
		memset(x, 0, n * sizeof(dcx));
		memcpy(r, b, n * sizeof(dcx));
		inprodR = temp;
	}
	else {
		memcpy(x, b, n * sizeof(dcx));
		mv(&S, x, Av, NULL, 0);
		for (size_t i = 0; i < n; i++) r[i] = b[i] - Av[i];
		inprodR = norm2(r, n);
	}
	if (resid_hist && hist_len > 0) resid_hist[0] = sqrt(resid_scale * inprodR);
	int niter = 1, counter = 0;

	/* solver state (static variables of each ITER_FUNC) */
	double c_old = 0, c_new = 0, omega_old = 0, omega_new = 0, ro_old_d = 0;
	dcx beta = 0, tautilda = 0, s_new = 0, s_old = 0, ro_old = 0, omega = 0, alpha = 0, tau = 0;
	double beta_d = 0;
	dcx *v = v1, *vtilda = v2, *p_new = p, *p_old = v3;   /* QMR_CS renames, src/iterative.c:916-921 */
	dcx *q_new = r, *q_old = v1, *pc_new = p, *pc_old = v2; /* CSYM renames, src/iterative.c:795-799 */
	/* BCGS2 (l = 2) renames, src/iterative.c:365-377: R[0..2] residuals, U[0..2] search directions, p = r~_0 */
	dcx *R[3] = {r, v1, v3}, *U[3] = {v2, Av, v4};
	dcx rho0 = 0;
	/* QMR_CS_2 state, src/iterative.c:1050-1053 (v = vec1, d = vec2, p = pvec) */
	double q2_ro = 0, q2_c = 0, q2_theta = 0;
	dcx q2_eps = 0, q2_eta = 0;
	/* PHASE_INIT */
	if (method == 0) {
		omega_old = 0.0;
		beta = csqrt(dotc(r, r, n));
		omega_new = sqrt(inprodR) / cabs(beta);
		for (size_t i = 0; i < n; i++) v[i] = r[i] / beta;
		tautilda = omega_new * beta;
		c_new = c_old = 1.0;
		s_new = s_old = 0.0;
	}
	else if (method == 1) memcpy(v3, r, n * sizeof(dcx)); /* r~ = r_0 */
	else if (method == 5) { memcpy(p, r, n * sizeof(dcx)); rho0 = -1; } /* src/iterative.c:384-387 */
	else if (method == 6) { /* src/iterative.c:1070-1080 */
		q2_ro = sqrt(inprodR);
		memcpy(v1, r, n * sizeof(dcx));
		q2_c = 1;
		q2_eps = 1;
		q2_theta = 0;
		q2_eta = -1;
	}
	else if (method == 4) {
		tau = sqrt(inprodR);
		for (size_t i = 0; i < n; i++) q_new[i] = conj(q_new[i]) * (1 / creal(tau));
		c_new = 1;
		c_old = 0;
		s_new = s_old = 0;
	}
	while (inprodR > epsB && niter <= maxiter && counter <= mc) {
		if (method == 0) { /* QMR_CS, src/iterative.c:962-1030 */
			double dtmp1 = 1 / (omega_new * omega_new), dtmp2, zetaabs;
			dcx theta, eta, zetatilda, zeta, temp1, temp2, temp4;
			if (dtmp1 < 1e-10) { rc = -1; break; }
			if (niter == 1 && matvec_ready) { temp1 = 1 / beta; for (size_t i = 0; i < n; i++) Av[i] *= temp1; }
			else mv(&S, v, Av, NULL, 0);
			alpha = dotc(v, Av, n);
			temp2 = -alpha;
			if (niter == 1) for (size_t i = 0; i < n; i++) vtilda[i] = temp2 * v[i] + Av[i];
			else { temp1 = -beta; for (size_t i = 0; i < n; i++) vtilda[i] = temp1 * vtilda[i] + temp2 * v[i] + Av[i]; }
			theta = conj(s_old) * omega_old * beta;
			eta = c_old * c_new * omega_old * beta;
			eta += alpha * conj(s_new) * omega_new;
			zetatilda = c_new * omega_new * alpha - s_new * c_old * omega_old * beta;
			omega_old = omega_new;
			{ /* nDotProdSelf_conj_Norm2 */
				double b0 = 0, b1 = 0, b2 = 0;
				for (size_t i = 0; i < n; i++) {
					b0 += creal(vtilda[i]) * creal(vtilda[i]);
					b1 += cimag(vtilda[i]) * cimag(vtilda[i]);
					b2 += creal(vtilda[i]) * cimag(vtilda[i]);
				}
				dtmp1 = b0 + b1;
				temp1 = b0 - b1 + I * 2 * b2;
			}
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
			if (niter == 1) for (size_t i = 0; i < n; i++) p_new[i] = temp4 * v[i];
			else {
				temp2 = -eta * temp4;
				if (niter == 2) for (size_t i = 0; i < n; i++) p_old[i] = temp2 * p_new[i] + temp4 * v[i];
				else {
					temp1 = -theta * temp4;
					for (size_t i = 0; i < n; i++) p_old[i] = temp1 * p_old[i] + temp2 * p_new[i] + temp4 * v[i];
				}
				dcx *t = p_old; p_old = p_new; p_new = t;
			}
			tau = c_new * tautilda;
			tautilda = -s_new * tautilda;
			for (size_t i = 0; i < n; i++) x[i] += tau * p_new[i];
			temp1 = 1 / beta;
			for (size_t i = 0; i < n; i++) vtilda[i] *= temp1;
			{ dcx *t = v; v = vtilda; vtilda = t; }
			temp1 = (c_new / omega_new) * tautilda;
			double cs = creal(s_new) * creal(s_new) + cimag(s_new) * cimag(s_new);
			double sum = 0;
			for (size_t i = 0; i < n; i++) {
				r[i] = cs * r[i] + temp1 * v[i];
				sum += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
			}
			inprodRp1 = sum;
		}
		else if (method == 1) { /* BiCGStab, src/iterative.c:675-718; v=vec1 s=vec2 rtilda=vec3 */
			dcx *vv = v1, *sv = v2, *rt = v3;
			dcx ro_new = dot(r, rt, n), temp1, temp2;
			if (niter == 1) memcpy(p, r, n * sizeof(dcx));
			else {
				temp1 = ro_new * alpha;
				temp2 = ro_old * omega;
				if (cabs(temp2) / cabs(temp1) < 1e-10) { rc = -1; break; }
				beta = temp1 / temp2;
				temp1 = -beta * omega;
				for (size_t i = 0; i < n; i++) p[i] = beta * p[i] + temp1 * vv[i] + r[i];
			}
			if (niter == 1 && matvec_ready) memcpy(vv, Av, n * sizeof(dcx));
			else mv(&S, p, vv, NULL, 0);
			temp1 = dot(vv, rt, n);
			if (cabs(temp1) / cabs(ro_new) < 1e-10) { rc = -1; break; }
			alpha = ro_new / temp1;
			temp1 = -alpha;
			double sum = 0;
			for (size_t i = 0; i < n; i++) {
				sv[i] = temp1 * vv[i] + r[i];
				sum += creal(sv[i]) * creal(sv[i]) + cimag(sv[i]) * cimag(sv[i]);
			}
			inprodRp1 = sum;
			if (inprodRp1 < epsB) for (size_t i = 0; i < n; i++) x[i] += alpha * p[i];
			else {
				double denumOmega;
				mv(&S, sv, Av, &denumOmega, 0);
				omega = dot(sv, Av, n) / denumOmega;
				for (size_t i = 0; i < n; i++) x[i] += alpha * p[i] + omega * sv[i];
				temp1 = -omega;
				sum = 0;
				for (size_t i = 0; i < n; i++) {
					r[i] = temp1 * Av[i] + sv[i];
					sum += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
				}
				inprodRp1 = sum;
				ro_old = ro_new;
			}
		}
		else if (method == 2) { /* CGNR, src/iterative.c:745-766 */
			double ro_new, denum, a;
			if (niter == 1) mv(&S, r, p, &ro_new, 1);
			else {
				mv(&S, r, Av, &ro_new, 1);
				double bt = ro_new / ro_old_d;
				for (size_t i = 0; i < n; i++) p[i] = bt * p[i] + Av[i];
			}
			mv(&S, p, Av, &denum, 0);
			a = ro_new / denum;
			for (size_t i = 0; i < n; i++) x[i] += a * p[i];
			double sum = 0;
			for (size_t i = 0; i < n; i++) {
				r[i] += -a * Av[i];
				sum += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
			}
			inprodRp1 = sum;
			ro_old_d = ro_new;
		}
		else if (method == 3) { /* BiCG_CS, src/iterative.c:549-630 */
			dcx ro_new = dotc(r, r, n), mu, temp;
			double abs_ro_new = cabs(ro_new);
			if (abs_ro_new / inprodR < 1e-10) { rc = -1; break; }
			if (niter == 1) memcpy(p, r, n * sizeof(dcx));
			else {
				beta = ro_new / ro_old;
				for (size_t i = 0; i < n; i++) p[i] = beta * p[i] + r[i];
			}
			if (!(niter == 1 && matvec_ready)) mv(&S, p, Av, NULL, 0);
			mu = dotc(p, Av, n);
			if (cabs(mu) / abs_ro_new < 1e-10) { rc = -1; break; }
			alpha = ro_new / mu;
			for (size_t i = 0; i < n; i++) x[i] += alpha * p[i];
			temp = -alpha;
			double sum = 0;
			for (size_t i = 0; i < n; i++) {
				r[i] += temp * Av[i];
				sum += creal(r[i]) * creal(r[i]) + cimag(r[i]) * cimag(r[i]);
			}
			inprodRp1 = sum;
			ro_old = ro_new;
		}
		else if (method == 5) { /* BCGS2, src/iterative.c:391-483 */
			dcx rho1, sigma;
			for (int j = 0; j < 2; j++) { /* BiCG part */
				rho1 = dot(R[j], p, n);
				if (niter == 1 && j == 0) memcpy(U[0], R[0], n * sizeof(dcx));
				else {
					if (cabs(rho0) / (cabs(rho1) * cabs(alpha)) < 1e-10) { rc = -1; break; }
					beta = alpha * rho1 / rho0;
					for (int i = 0; i <= j; i++)
						for (size_t e = 0; e < n; e++) U[i][e] = -beta * U[i][e] + R[i][e];
				}
				rho0 = rho1;
				if (!(niter == 1 && j == 0 && matvec_ready)) mv(&S, U[j], U[j + 1], NULL, 0);
				sigma = dot(U[j + 1], p, n);
				if (cabs(sigma) / cabs(rho1) < 1e-10) { rc = -1; break; }
				alpha = rho1 / sigma;
				for (size_t e = 0; e < n; e++) x[e] += alpha * U[0][e];
				for (int i = 0; i <= j; i++)
					for (size_t e = 0; e < n; e++) R[i][e] += -alpha * U[i + 1][e];
				mv(&S, R[j], R[j + 1], NULL, 0);
			}
			if (rc) break;
			/* minimal-residual polynomial of degree 2 over Z = R^H R, with Z[i][j] = R_j . conj(R_i) */
			dcx Z[3][3], y0[3], yl[3], zy0[3], zyl[3], vr = 0;
			for (int i = 0; i < 3; i++)
				for (int j = 0; j <= i; j++) {
					Z[i][j] = dot(R[j], R[i], n);
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
			dcx t = -hg * k0 / kl;
			for (int i = 0; i < 3; i++) y0[i] += t * yl[i];
			omega = y0[2];
			for (int i = 1; i <= 2; i++)
				for (size_t e = 0; e < n; e++) {
					U[0][e] += -y0[i] * U[i][e];
					x[e] += y0[i] * R[i - 1][e];
					R[0][e] += -y0[i] * R[i][e];
				}
			inprodRp1 = 0;
			for (int i = 0; i < 3; i++) {
				dcx z = Z[i][0] * y0[0] + Z[i][1] * y0[1] + Z[i][2] * y0[2];
				inprodRp1 += creal(z * conj(y0[i]));
			}
			rho0 *= -omega;
		}
		else if (method == 6) { /* QMR_CS_2, src/iterative.c:1082-1136 (v = v1, d = v2, p_k = p) */
			dcx *vq = v1, *d = v2, delta, bt, temp1;
			double inv = 1 / q2_ro, nv2 = 0;
			for (size_t e = 0; e < n; e++) vq[e] *= inv;
			delta = dotc(vq, vq, n);
			if (cabs(delta) < 1e-10) { rc = -1; break; }
			if (niter == 1) memcpy(p, vq, n * sizeof(dcx));
			else {
				temp1 = -q2_ro * delta / q2_eps;
				for (size_t e = 0; e < n; e++) p[e] = temp1 * p[e] + vq[e];
			}
			if (niter == 1 && matvec_ready) for (size_t e = 0; e < n; e++) Av[e] *= inv;
			else mv(&S, p, Av, NULL, 0);
			q2_eps = dotc(p, Av, n);
			bt = q2_eps / delta;
			if (cabs(q2_eps) < 1e-10) { rc = -1; break; }
			for (size_t e = 0; e < n; e++) {
				vq[e] = -bt * vq[e] + Av[e];
				nv2 += creal(vq[e]) * creal(vq[e]) + cimag(vq[e]) * cimag(vq[e]);
			}
			double ro_new = sqrt(nv2);
			double th = ro_new / (q2_c * cabs(bt));
			double c_nw = 1 / sqrt(1 + th * th), sabs2 = th * th / (1 + th * th), cr = c_nw / q2_c;
			q2_eta = -q2_ro * cr * cr * q2_eta / bt;
			if (niter == 1) for (size_t e = 0; e < n; e++) d[e] = q2_eta * p[e];
			else {
				double f = q2_theta * c_nw;
				for (size_t e = 0; e < n; e++) d[e] = f * f * d[e] + q2_eta * p[e];
			}
			double sum = 0;
			for (size_t e = 0; e < n; e++) {
				x[e] += d[e];
				r[e] = sabs2 * r[e] - q2_eta * vq[e];
				sum += creal(r[e]) * creal(r[e]) + cimag(r[e]) * cimag(r[e]);
			}
			inprodRp1 = sum;
			q2_ro = ro_new;
			q2_theta = th;
			q2_c = c_nw;
		}
		else { /* CSYM, src/iterative.c:822-888 */
			dcx eta, gamma, theta, invksi, temp1, temp2;
			double dtmp;
			mv(&S, q_new, Av, NULL, 0);
			alpha = dotc(q_new, Av, n);
			eta = c_old * c_new * beta_d + alpha * conj(s_new);
			gamma = c_new * alpha - c_old * s_new * beta_d;
			theta = beta_d * conj(s_old);
			temp1 = -alpha;
			double sum = 0;
			if (niter == 1) for (size_t i = 0; i < n; i++) {
				q_old[i] = temp1 * conj(q_new[i]) + Av[i];
				sum += creal(q_old[i]) * creal(q_old[i]) + cimag(q_old[i]) * cimag(q_old[i]);
			}
			else for (size_t i = 0; i < n; i++) {
				q_old[i] = -beta_d * conj(q_old[i]) + temp1 * conj(q_new[i]) + Av[i];
				sum += creal(q_old[i]) * creal(q_old[i]) + cimag(q_old[i]) * cimag(q_old[i]);
			}
			beta_d = sqrt(sum);
			c_old = c_new;
			s_old = s_new;
			dtmp = cabs(gamma);
			if (dtmp == 0) {
				if (beta_d == 0) { rc = -1; break; }
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
			if (niter == 1) for (size_t i = 0; i < n; i++) pc_new[i] = invksi * q_new[i];
			else {
				temp1 = -eta * invksi;
				if (niter == 2) for (size_t i = 0; i < n; i++) pc_old[i] = temp1 * pc_new[i] + invksi * q_new[i];
				else {
					temp2 = -theta * invksi;
					for (size_t i = 0; i < n; i++) pc_old[i] = temp2 * pc_old[i] + temp1 * pc_new[i] + invksi * q_new[i];
				}
				dcx *t = pc_old; pc_old = pc_new; pc_new = t;
			}
			temp1 = c_new * tau;
			for (size_t i = 0; i < n; i++) x[i] += temp1 * pc_new[i];
			for (size_t i = 0; i < n; i++) q_old[i] = conj(q_old[i]) * (1 / beta_d);
			{ dcx *t = q_old; q_old = q_new; q_new = t; }
			tau *= -s_new;
			inprodRp1 = creal(tau) * creal(tau) + cimag(tau) * cimag(tau);
		}
		/* ProgressReport, src/iterative.c:282-318 */
		if (inprodRp1 <= inprodR) { inprodR = inprodRp1; counter = 0; }
		else counter++;
		if (resid_hist && niter < hist_len) resid_hist[niter] = sqrt(resid_scale * inprodRp1);
		niter++;
	}
	if (rc == 0 && inprodR > epsB && niter <= maxiter && counter > mc) rc = -2; /* src/iterative.c:1509-1510 */
	/* P = S.x, src/iterative.c:1527 */
	for (size_t i = 0; i < g->nd; i++)
		for (int c = 0; c < 3; c++) pvec_o[3 * i + c] = O(C(cc_sqrt[3 * g->mat[i] + c]) * x[3 * i + c]);
	if (niter_out) *niter_out = niter - 1;
	if (nmatvec_out) *nmatvec_out = S.nmv;
	free(b); free(r); free(Av); free(p); free(v1); free(v2); free(v3); free(v4);
	return rc;
}

/*====================================================================================================================*/
/* CalcFieldFree: src/crosssec.c:506-595 (DRAINE scattering relation, non-surface) */
void or_calc_field(const or_grid *g, const or_cplx *pvec, const double nv[3], or_cplx ebuff[3])
{
	double kd = g->wavenum * g->gridspace;
	dcx sum[3] = {0, 0, 0};
	for (size_t j = 0; j < g->nd; j++) {
		const unsigned short *p = g->pos + 3 * j;
		double ph = -kd * (nv[0] * p[0] + nv[1] * p[1] + nv[2] * p[2]);
		dcx a = cos(ph) + I * sin(ph);
		for (int c = 0; c < 3; c++) sum[c] += C(pvec[3 * j + c]) * a;
	}
	dcx dpr = sum[0] * nv[0] + sum[1] * nv[1] + sum[2] * nv[2];
	dcx tb[3];
	for (int c = 0; c < 3; c++) tb[c] = sum[c] - dpr * nv[c];
	/* r0 = box_origin_unif = -d*(cX,cY,cZ), src/make_particle.c:2305-2311 */
	double r0[3] = {-g->gridspace * (g->boxX - 1) / 2.0, -g->gridspace * (g->boxY - 1) / 2.0,
		-g->gridspace * (g->boxZ - 1) / 2.0};
	double ph = -g->wavenum * (r0[0] * nv[0] + r0[1] * nv[1] + r0[2] * nv[2]);
	dcx a = cos(ph) + I * sin(ph);
	double kkk = g->wavenum * g->wavenum * g->wavenum;
	dcx tmp = -I * a * kkk;
	for (int c = 0; c < 3; c++) ebuff[c] = O(tmp * tb[c]);
}

/* ExtCross, plane wave: src/crosssec.c:811-823 */
double or_ext_cross(const or_grid *g, const or_cplx *pvec, const double prop[3], const double incpol[3])
{
	or_cplx e[3];
	or_calc_field(g, pvec, prop, e);
	double sum = e[0].re * incpol[0] + e[1].re * incpol[1] + e[2].re * incpol[2];
	return sum * FOUR_PI / (g->wavenum * g->wavenum);
}

/* AbsCross, SQ_DRAINE: src/crosssec.c:873-884,916 */
double or_abs_cross(const or_grid *g, const or_cplx *pvec, const or_cplx *cc)
{
	double k = g->wavenum, temp1 = 2 * k * k * k / 3, sum = 0;
	for (size_t d = 0; d < g->nd; d++)
		for (int c = 0; c < 3; c++) {
			double mult = -cimag(1 / C(cc[3 * g->mat[d] + c])) - temp1;
			dcx p = C(pvec[3 * d + c]);
			sum += mult * (creal(p) * creal(p) + cimag(p) * cimag(p));
		}
	return FOUR_PI * k * sum;
}

/* CalcEplaneYZ, CE_NORMAL, no orientation averaging: src/CalculateE.c:423-507 */
void or_eplane_yz(const or_grid *g, const or_cplx *pvec, const double prop[3], const double par[3],
	const double per[3], int ntheta, int npoints, or_cplx *eplane)
{
	double dth = TWO_PI / 2 / (double)(ntheta - 1); /* dtheta_rad, src/calculator.c:671-672 */
	for (int i = 0; i < npoints; i++) {
		double th = i * dth, co = cos(th), si = sin(th), robs[3], epar[3];
		for (int c = 0; c < 3; c++) { robs[c] = co * prop[c] + si * par[c]; epar[c] = -si * prop[c] + co * par[c]; }
		or_cplx e[3];
		or_calc_field(g, pvec, robs, e);
		dcx eper = 0, epa = 0;
		for (int c = 0; c < 3; c++) { eper += C(e[c]) * per[c]; epa += C(e[c]) * epar[c]; }
		eplane[2 * i] = O(eper);
		eplane[2 * i + 1] = O(epa);
	}
}

static inline double ca2(dcx a) { return creal(a) * creal(a) + cimag(a) * cimag(a); }

/* ComputeMuellerMatrix (src/CalculateE.c:83-116) with s1=EX[2i], s2=EY[2i+1], s3=EX[2i+1], s4=EY[2i]
 * (MuellerMatrix, yz plane, src/CalculateE.c:236-246) */
void or_mueller_yz(const or_cplx *eplX, const or_cplx *eplY, int ntheta, double *out)
{
	for (int i = 0; i < ntheta; i++) {
		dcx s1 = C(eplX[2 * i]), s2 = C(eplY[2 * i + 1]), s3 = C(eplX[2 * i + 1]), s4 = C(eplY[2 * i]);
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

/*====================================================================================================================*/
/* Adjoint gradient of Cext and Cabs with respect to each dipole's permittivity eps = m^2. Not a reference function
 * (ADDA has no gradient); it is the derivative of the reference's own quantities: CoupleConstant
 * (src/calculator.c:100-133), ExtCross (src/crosssec.c:811-823, equal to 4 pi k Im sum conj(Einc).P for a plane wave)
 * and AbsCross/DRAINE (src/crosssec.c:852-884). Pinned by central finite differences of the oracle's own forward
 * solves and by finite differences of the reference's outputs (tests/golden/adj_*). */

/* d(cc)/d(eps) of or_couple_constant: cc = a/(1 - a M/V), a = (3V/4pi)(eps-1)/(eps+2), M = M(eps) */
or_cplx or_couple_constant_deps(or_cplx mo, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3])
{
	dcx m = C(mo), e = m * m;
	double V = gridspace * gridspace * gridspace, kd = wavenum * gridspace;
	dcx a = (3 / FOUR_PI) * V * (e - 1) / (e + 2), da = (3 / FOUR_PI) * V * 3 / ((e + 2) * (e + 2));
	if (polrel == 1) return O(da);
	dcx M = I * 2 * kd * kd * kd / 3, dM = 0;
	if (polrel != 2) {
		double S = 0;
		for (int i = 0; i < 3; i++) S += prop[i] * prop[i] * incpol[i] * incpol[i];
		M += (LDR_B1 + (LDR_B2 + LDR_B3 * S) * e) * kd * kd;
		dM = (LDR_B2 + LDR_B3 * S) * kd * kd;
	}
	dcx den = 1 - a * M / V;
	return O((da + a * a * dM / V) / (den * den));
}

/* Assembly: P solves M P = Einc with M = diag(1/cc) - G; lam = M^-1 conj(Einc), mu = M^-1 conj(W P) (M is
 * symmetric); q = d(1/cc)/d eps = -cc'/cc^2 per material and component; wabs = -Im(1/cc) - 2k^3/3.
 *   Cext = ce Im(sum conj(Einc).P):   dCext = ce Im(h de),  h = -sum_c q lam_c P_c
 *   Cabs = ca sum wabs |P|^2:         dCabs = ca [-Im(q de)|P|^2 + 2 Re(z de)],  z = -sum_c q mu_c P_c
 * out[2j] = dC/dRe(eps_j), out[2j+1] = dC/dIm(eps_j). */
void or_adjoint_grad(const or_grid *g, const or_cplx *pvec, const or_cplx *lam, const or_cplx *mu, const or_cplx *q,
	const double *wabs, double ce, double ca, double *gext, double *gabs)
{
	(void)wabs;
	for (size_t j = 0; j < g->nd; j++) {
		const int mt = g->mat[j];
		dcx h = 0, z = 0, loc = 0;
		for (int c = 0; c < 3; c++) {
			const dcx qq = C(q[3 * mt + c]), p = C(pvec[3 * j + c]);
			h -= qq * C(lam[3 * j + c]) * p;
			z -= qq * C(mu[3 * j + c]) * p;
			loc += qq * (creal(p) * creal(p) + cimag(p) * cimag(p));
		}
		gext[2 * j] = ce * cimag(h);
		gext[2 * j + 1] = ce * creal(h);
		gabs[2 * j] = ca * (-cimag(loc) + 2 * creal(z));
		gabs[2 * j + 1] = ca * (-creal(loc) - 2 * cimag(z));
	}
}
