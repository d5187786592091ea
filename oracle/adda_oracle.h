/* oracle/adda_oracle.h -- CPU FP64 restatement of ADDA's FFT MatVec + Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is compared against. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference checkout).
 * Pinning (which check fixes each function) is listed next to each declaration; see tests/test_oracle.py.
 */
#ifndef ADDA_ORACLE_H
#define ADDA_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* complex numbers cross the boundary as interleaved (re,im) doubles, layout-compatible with C99 double complex */
typedef struct { double re, im; } or_cplx;

/* Particle above a plane substrate (-surf h {mre mim | inf} with -int_surf img): the image-dipole reflected term */
typedef struct {
	double zsum_shift;           /* ZsumShift = 2*(hsub/d - (boxZ-1)/2), src/make_particle.c:2312 */
	or_cplx rcn;                 /* surfRCn: -1 for a perfect reflector, else (1-msub^2)/(1+msub^2), src/interaction.c:1411 */
	const or_cplx *rmat;         /* R-matrix of or_init_rmatrix, used by the FFT MatVec (the direct sum ignores it) */
} or_surface;

typedef struct {
	int boxX, boxY, boxZ;        /* particle box (src/vars.h boxX..) */
	int gridX, gridY, gridZ;     /* padded FFT grid (src/comm.c:489-494) */
	size_t nd;                   /* nvoid_Ndip */
	const unsigned short *pos;   /* position[3*nd] (src/vars.h:94) */
	const int *mat;              /* material[nd] (index into the cc_sqrt table) */
	double wavenum;              /* WaveNum = 2*pi/lambda */
	double gridspace;            /* d = lambda/dpl */
	int reduced_fft;             /* reduced_FFT (src/param.c:531) */
	const or_surface *surf;      /* NULL: free space */
} or_grid;

/* fftFit (src/fft.c:528-562): first x' >= x, even, divisible by divis, with prime factors {2,3,5} (rule 1, the
 * Temperton/clAmdFft rule) or {2,3,5,7} plus at most one 11 or 13 (rule 0, the FFTW3 rule). Pinned: tests compare with
 * the grids in the reference sample log (16 -> 32) and enumerations of the rules. */
int or_fftfit(int x, int divis, int rule);

/* InterTerm_poi (src/interaction.c:159-180,290-331): point-dipole Green's tensor between grid offset (i,j,k)
 * (units of d), 6 components xx,xy,xz,yy,yz,zz. Pinned by the closed form and by SPARSE-build runs. */
void or_green_poi(int i, int j, int k, double wavenum, double gridspace, or_cplx res[6]);

/* Batched unnormalized 1D DFT, forward sign -1 / backward +1 (src/fft.h:31-32). Pinned by brute-force DFT. */
void or_fft1d(or_cplx *data, size_t n, size_t stride, size_t howmany, size_t dist, int sign);

/* Size of the D-matrix array in complex elements (6*gridX*DsizeY*DsizeZ, src/fft.c:955-970,1113) */
size_t or_dmatrix_size(const or_grid *g);

/* InitDmatrix (src/fft.c:918-1416): D = -FFT3(G)/(gX*gY*gZ), stored with IndexDmatrix (src/fft.c:144-150). Pinned by the
 * FFT-convolution = direct-sum identity and by reduced = non-reduced equality. */
void or_init_dmatrix(const or_grid *g, or_cplx *dmat);

/* MatVec (src/matvec.c:128-425) restated on the full (not zero-skipped) grid: y = x + S.F^-1[D.F(S.x)];
 * her => conj(A conj x); inprod (nullable) => ||y||^2. cc_sqrt is [nmat][3]. With g->surf the reflected term
 * R.F'(S.x) is added to the product (src/matvec.c:249,288-302): F' is the 3D transform with the z direction backward
 * (src/fft.c:383-415), cReflMatrVec (src/cmplx.h:391-401) applies R. */
void or_matvec_fft(const or_grid *g, const or_cplx *dmat, const or_cplx *cc_sqrt, const or_cplx *arg, or_cplx *res,
	double *inprod, int her);

/* SPARSE MatVec (src/matvec.c:434-465, src/sparse_ops.h AijProd/DiagProd): y_i = x_i - S_i sum_{j!=i} G_ij S_j x_j;
 * with g->surf also - S_i sum_j GR(x_i-x_j, y_i-y_j, z_i+z_j) S_j x_j over every j, i included
 * (src/sparse_ops.h:160-165). */
void or_matvec_direct(const or_grid *g, const or_cplx *cc_sqrt, const or_cplx *arg, or_cplx *res, double *inprod,
	int her);

/* ReflTerm_img through its integer wrapper (src/interaction.c:1081-1153, vCopyIntRealShift): the image-dipole
 * reflected Green's tensor for grid offset (i, j) and height sum ksum (units of d; qvec = (i, j, ksum + zsum_shift)),
 * components as InterTerm with xz, yz, zz negated and scaled by rcn. Pinned by the closed form at hand-computed
 * offsets (an independent evaluation of the image dipole's field) and by SPARSE -surf runs. */
void or_refl_img(int i, int j, int ksum, double zsum_shift, or_cplx rcn, double wavenum, double gridspace,
	or_cplx res[6]);

/* Size of the R-matrix in complex elements: 6*gridX*gridZ*RsizeY, RsizeY = gridY/2+1 reduced, else gridY
 * (src/fft.c:982-1000). */
size_t or_rmatrix_size(const or_grid *g);

/* InitRmatrix (src/fft.c:819-891): R = -FFT3(GR)/Ngrid with GR[i][j][k] = ReflTerm for |i|<boxX, |j|<boxY,
 * 0 <= k <= 2*boxZ-2, stored as IndexRmatrix_mv (src/matvec.c:107-121): 6*((x*gridZ + z)*RsizeY + y). Needs
 * g->surf. Pinned by FFT MatVec == direct sum with the reflected term (brute force). */
void or_init_rmatrix(const or_grid *g, or_cplx *rmat);

/* CoupleConstant for scalar m (src/calculator.c:100-133,170-262). polrel: 0 = LDR, 1 = CM, 2 = RRC.
 * prop/incpol: unit vectors (LDR S-term = sum_i (prop_i*incpol_i)^2, src/calculator.c:214-221).
 * KAT: 0.005259037197+1.843854148e-05i for m=1.5, dpl=15 (sample log). */
or_cplx or_couple_constant(or_cplx m, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3]);

/* GenerateB plane wave (src/GenerateB.c:324-330): Einc_i = exp(i k r_i.prop) e_pol, with DipoleCoord
 * r = (pos - (box-1)/2)*d (src/make_particle.c:2266-2275). */
void or_generate_b(const or_grid *g, const double prop[3], const double incpol[3], or_cplx *einc);

/* GenerateB plane wave above a substrate (src/GenerateB.c:258-322) with InitBeam's ki, kt, ktVec, prIncRefl
 * (src/GenerateB.c:96-137): from above (prop[2] < 0) incident plus reflected wave, from below (prop[2] > 0) the
 * transmitted wave; which_x = 0 takes ex = incPolY (s-polarized), 1 takes ex = incPolX (p-polarized). msub_inf: a
 * perfectly reflecting substrate (msub ignored). Returns inc_scale (1/Re(msub) from below, else 1). Pinned by SPARSE
 * -surf runs (incidence from above and below, finite and perfectly reflecting substrate). */
double or_generate_b_surf(const or_grid *g, or_cplx msub, int msub_inf, double hsub, const double prop[3],
	const double incpol_x[3], const double incpol_y[3], int which_x, or_cplx *einc);

/* ExtCross, general form (src/crosssec.c:824-846): 4*pi*k*sum Im(P.conj(Einc)) (times inc_scale near a substrate). */
double or_ext_cross_gen(const or_grid *g, const or_cplx *pvec, const or_cplx *einc);

/* IterativeSolver (src/iterative.c:1398-1530) with QMR_CS (898-1032), BiCGStab (640-724), CGNR (730-770),
 * BiCG_CS (495-634), CSYM (774-894), BCGS2 (342-488) and QMR_CS_2 (1039-1143) restated; matvec is the FFT one if
 * dmat != NULL, else the direct sum.
 * method: 0=qmr 1=bicgstab 2=cgnr 3=bicg 4=csym 5=bcgs2 6=qmr2. init_field: 0=auto 1=zero 2=inc (src/iterative.c:1342-1394).
 * On return: x (solution of the symmetrized system), pvec = S.x (polarization), resid_hist[0..niter] relative
 * residuals (may be NULL), *niter_out iterations, *nmatvec_out matvec count. Returns 0, or -1 on breakdown,
 * -2 on stagnation. */
int or_iterative_solver(const or_grid *g, const or_cplx *dmat, const or_cplx *cc_sqrt, const or_cplx *einc,
	int method, double eps, int maxiter, int init_field, or_cplx *x, or_cplx *pvec, double *resid_hist,
	int hist_len, int *niter_out, int *nmatvec_out);

/* CalcFieldFree (src/crosssec.c:506-595): scattering amplitude along unit vector n from polarization P. */
void or_calc_field(const or_grid *g, const or_cplx *pvec, const double n[3], or_cplx ebuff[3]);

/* ExtCross for a plane wave (src/crosssec.c:811-823). */
double or_ext_cross(const or_grid *g, const or_cplx *pvec, const double prop[3], const double incpol[3]);

/* AbsCross, DRAINE formula (src/crosssec.c:852-884,916); cc is [nmat][3]. */
double or_abs_cross(const or_grid *g, const or_cplx *pvec, const or_cplx *cc);

/* CalcEplaneYZ (src/CalculateE.c:423-507), one polarization, CE_NORMAL: eplane[2*i] = E.incPolper,
 * eplane[2*i+1] = E.epar for theta_i = i*180/(ntheta-1), i < npoints (npoints = 2*(ntheta-1) when the range is
 * extended to 360 deg, src/calculator.c:674). */
void or_eplane_yz(const or_grid *g, const or_cplx *pvec, const double prop[3], const double incpol_par[3],
	const double incpol_per[3], int ntheta, int npoints, or_cplx *eplane);

/* ComputeMuellerMatrix (src/CalculateE.c:83-116) applied as in MuellerMatrix for the yz plane
 * (src/CalculateE.c:236-246): out[16*i..] row-major 4x4 per angle. */
void or_mueller_yz(const or_cplx *eplX, const or_cplx *eplY, int npoints, double *out);

/* d(cc)/d(eps) of or_couple_constant at eps = m^2 (derivative of its closed form). Pinned: central differences of
 * or_couple_constant in eps (tests/test_oracle.py). */
or_cplx or_couple_constant_deps(or_cplx m, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3]);

/* Adjoint gradient assembly of Cext = ce Im(sum conj(Einc).P) and Cabs = ca sum wabs|P|^2 with respect to each
 * dipole's permittivity: pvec = P, lam = M^-1 conj(Einc), mu = M^-1 conj(W.P) (M = diag(1/cc) - G, symmetric),
 * q[nmat][3] = d(1/cc)/d eps; gext/gabs[2j] = dC/dRe(eps_j), [2j+1] = dC/dIm(eps_j). Pinned: central differences of
 * the oracle's forward solves and of the reference's outputs (tests/golden/adj_*). */
void or_adjoint_grad(const or_grid *g, const or_cplx *pvec, const or_cplx *lam, const or_cplx *mu, const or_cplx *q,
	const double *wabs, double ce, double ca, double *gext, double *gabs);

#ifdef __cplusplus
}
#endif
#endif
