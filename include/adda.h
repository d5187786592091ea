/* adda.h -- host-side C API of the MI355X ADDA solver: particle builder, polarizability, plane-wave beam, the
 * Krylov driver (IterativeSolver with QMR_CS, BiCGStab, CGNR, BiCG_CS, CSYM), FP32-inner/FP64 defect correction, and
 * the scattering quantities. Everything numeric runs through include/adda_hip.h; this layer holds the scalar
 * recurrences and control flow, in C, the way ADDA's src/iterative.c holds them around MatVec.
 */
#ifndef ADDA_H
#define ADDA_H

#include <stddef.h>

#include "adda_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* iterative methods, ADDA's enum iter subset (src/const.h; option -iter, src/param.c:1230-1245) */
enum adda_iter {
	ADDA_IT_QMR_CS = 0, ADDA_IT_BICGSTAB = 1, ADDA_IT_CGNR = 2, ADDA_IT_BICG_CS = 3, ADDA_IT_CSYM = 4,
	ADDA_IT_BCGS2 = 5, ADDA_IT_QMR_CS_2 = 6
};
/* initial field (option -init_field, src/param.c:1146-1161) */
enum adda_init { ADDA_IF_AUTO = 0, ADDA_IF_ZERO = 1, ADDA_IF_INC = 2 };
/* polarizability relation (option -pol, subset) */
enum adda_pol { ADDA_POL_LDR = 0, ADDA_POL_CM = 1, ADDA_POL_RRC = 2 };
/* precision form: FP64; FP32 Krylov iterations (FP32 vectors and transforms) inside FP64 defect correction; FP64
 * Krylov iterations (FP64 vectors and recurrences) around FP32 transforms, inside the same FP64 defect correction */
enum adda_prec { ADDA_PREC_FP64 = 0, ADDA_PREC_MIXED = 1, ADDA_PREC_MIXED_V64 = 2 };

/* Error text of the last failing call of this library or of the HIP backend. */
const char *adda_last_error(void);

/* ---- particle (MakeParticle subset, src/make_particle.c:1952-2319) -------------------------------------------- */
typedef struct {
	int boxX, boxY, boxZ;
	size_t nd;                /* nvoid_Ndip */
	unsigned short *position; /* [3*nd], dipoles ordered z-major, then y, then x (the loop order of MakeParticle) */
	int *material;            /* [nd] */
	int nmat;
} adda_particle;

/* shape: "sphere", "ellipsoid" (aspect_y, aspect_z = semi-axis ratios y/x, z/x), "box" (aspect = side ratios).
 * grid = boxX (-grid option); boxY/boxZ follow FitBox_yz (src/make_particle.c:1295-1307). */
int adda_make_particle(const char *shape, int grid, double aspect_y, double aspect_z, adda_particle *p);
/* "-shape read" geometry file: lines "x y z [mat]" (1-based mat, comments '#'), src/make_particle.c ReadDipFile. */
int adda_read_particle(const char *fname, adda_particle *p);
void adda_free_particle(adda_particle *p);
/* Volume ratio of the shape (volume_ratio, e.g. pi/6 for a sphere, src/make_particle.c:1553), for -size. */
double adda_volume_ratio(const char *shape, double aspect_y, double aspect_z);

/* ---- physics ------------------------------------------------------------------------------------------------- */
/* CoupleConstant for scalar m (src/calculator.c:100-133,170-262). */
adda_cplx adda_couple_constant(adda_cplx m, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3]);
/* Plane wave Einc_i = exp(i k r_i.prop) e_pol on the device (GenerateB, src/GenerateB.c:324-330). */
int adda_generate_b(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const double prop[3],
	const double incpol[3], void *d_einc /* FP64 device vector */);

/* Particle above a plane substrate parallel to the xy plane (-surf <h> <mre> <mim> | -surf <h> inf, src/param.c:1510-1528)
 * with the image-dipole reflection (-int_surf img, ReflTerm_img, src/interaction.c:1136-1153); the Sommerfeld-integral
 * reflection (-int_surf som) is not implemented. */
typedef struct {
	double hsub;      /* height of the particle centre above the surface (um) */
	adda_cplx msub;   /* refractive index of the substrate (ignored with inf) */
	int inf;          /* perfectly reflecting substrate (msubInf) */
} adda_surface;

/* Puts the operator's particle above the substrate: ZsumShift = 2*(hsub/d - (boxZ-1)/2) (src/make_particle.c:2312),
 * surfRCn = -1 (inf) or (1-msub^2)/(1+msub^2) (src/interaction.c:1411), then adda_hip_set_surface. Fails, as
 * src/make_particle.c:2282-2284 does, unless every dipole is above the surface. */
int adda_set_surface(adda_hip_op *op, const adda_particle *p, double gridspace, const adda_surface *s);
/* The incident plane wave near the substrate (InitBeam, src/GenerateB.c:96-137; GenerateB, 258-322): from above
 * (prop[2] < 0) the incident plus the reflected wave, from below (prop[2] > 0) the transmitted wave (evanescent
 * beyond the critical angle). which_x = 0: polarization ex = incPolY (s-polarized), 1: ex = incPolX (p-polarized);
 * incpol_x, incpol_y: the incident frame of src/param.c:2116-2131. *inc_scale = 1/Re(msub) from below, else 1. */
int adda_generate_b_surf(adda_hip_op *op, const adda_particle *p, const adda_surface *s, double wavenum,
	double gridspace, const double prop[3], const double incpol_x[3], const double incpol_y[3], int which_x,
	void *d_einc, double *inc_scale);
/* ExtCross, general form (src/crosssec.c:824-846), used near a substrate: 4 pi k inc_scale sum Im(P.conj(Einc)). */
int adda_ext_cross_gen(adda_hip_op *op, double wavenum, double inc_scale, const void *d_pvec, const void *d_einc,
	double *cext);

/* ---- solver ---------------------------------------------------------------------------------------------------- */
typedef struct {
	int method;          /* enum adda_iter */
	double eps;          /* stopping tolerance on ||r||/||b|| (-eps k => 10^-k) */
	int maxiter;         /* -maxiter */
	int init_field;      /* enum adda_init */
	int recalc_resid;    /* -recalc_resid: recompute the true residual at the end */
	int precision;       /* enum adda_prec */
	double inner_eps;    /* mixed: relative tolerance of each FP32 inner solve (default 1e-4 when 0) */
	int max_outer;       /* mixed: maximum FP64 correction steps (default 30 when 0) */
	int verbose;         /* bit 0: print RE_ lines to stdout; bit 1: a progress line (iteration, residual) about every
	                        10 s of a device-resident solve, inner solves of the mixed forms included */
	int host_scalars;    /* 1: scalar recurrences on the host with a synchronising reduction per inner product (the
	                      * reference's control flow, src/iterative.c); 0: device-resident recurrences (QMR_CS,
	                      * every method; checkpoints always use the host form) */
	int max_inner;       /* mixed: iteration cap of one FP32 inner solve before the FP64 correction (0 = no cap) */
	int graph;           /* device-resident recurrences: record the steady-state chunks of iterations once as a HIP
	                      * graph and replay it (1), or enqueue every iteration (0) */
	/* iteration checkpoints (-chpoint, -chp_type, -chp_dir, -chp_load; src/param.c:412-417, src/iterative.c:165-318);
	 * FP64 single-RHS solves, which then run the host form of the recurrences */
	int chp_type;        /* enum adda_chp: when to save <chp_dir>/chp.0 */
	double chp_time;     /* -chpoint: seconds after program start (> 0) -- the load of this library -- at which to save
	                      * (normal: then stop, regular: every chp_time seconds after the last save); <= 0 with
	                      * ADDA_CHP_ALWAYS: save at the end only. Set-up and earlier solves count, as the reference's
	                      * clock starts at program start (src/timing.c:102) */
	const char *chp_dir; /* checkpoint directory (NULL: "chpoint", the reference's default) */
	int chp_load;        /* 1: resume from <chp_dir>/chp.0 instead of starting from the initial field */
	const char *run_dir; /* output directory named in chp.log (NULL: ".") */
	int inner_method;    /* mixed forms: Krylov method of the FP32 inner solves, enum adda_iter + 1. 0 (auto): BiCGStab
	                      * for the complex-symmetric Lanczos methods (QMR_CS, BiCG_CS, CSYM, QMR_CS_2) -- they rely
	                      * on A^T = A, which the FP32 operator meets only to rounding (about 1e-7), and on a large
	                      * particle their inner solves stall while the recurrence residual keeps falling --, the
	                      * requested method otherwise. The outer FP64 residual decides every correction, and the FP64
	                      * finishing solve (a rejected correction) runs the requested method. */
} adda_solver_opts;

/* checkpoint types, ADDA's enum chpoint (src/const.h:313-318) */
enum adda_chp { ADDA_CHP_NONE = 0, ADDA_CHP_NORMAL = 1, ADDA_CHP_REGULAR = 2, ADDA_CHP_ALWAYS = 3 };

typedef struct {
	int rc;              /* 0 ok, -1 breakdown/error, -2 stagnation, 1 maxiter reached (warning), 2 stopped after
	                      * saving a checkpoint (CHP_EXIT of the reference, src/iterative.c:1528) */
	int niter;           /* Krylov iterations (summed over inner solves in mixed mode) */
	int nmatvec;         /* matvecs in the working precision of the iterations */
	int nmatvec_fp64;    /* FP64 matvecs of the outer loop (mixed mode) */
	int nouter;          /* outer correction steps (mixed mode) */
	double resid;        /* final relative residual of the recurrence (FP64 mode) or of the true residual (mixed) */
	double true_resid;   /* true ||b-Ax||/||b|| when recalculated (recalc_resid or mixed), else -1 */
	int chp_saved;       /* 1 when the solve wrote an iteration checkpoint */
} adda_solve_stats;

/* IterativeSolver (src/iterative.c:1398-1530) on device vectors: solves (I+S.D.S).x = S.Einc, returns x and the
 * polarization P = S.x (src/iterative.c:1527). d_einc, d_x, d_pvec: FP64 device vectors. resid_hist (may be NULL)
 * receives the relative residual after each iteration (RE_ lines), hist_len entries at most. */
int adda_iterative_solver(adda_hip_op *op, const adda_solver_opts *opt, const void *d_einc, void *d_x, void *d_pvec,
	double *resid_hist, int hist_len, adda_solve_stats *st);

/* The same for K right-hand sides sharing the operator (the polarizations and orientations ADDA solves one after
 * another, src/CalculateE.c:836-870 and src/crosssec.c:82-138): d_einc, d_x, d_pvec are batched FP64 device vectors
 * (adda_hip_vec_alloc_batch, stride adda_hip_batch_stride), the cc_sqrt table of RHS k set by
 * adda_hip_set_cc_sqrt_batch. Every RHS converges on its own (own scalars, tolerance test, status); in mixed
 * precision each runs its own defect correction, and the inner FP32 solves of the RHS still correcting share their
 * launches. Needs the device-resident recurrences (host_scalars 0, no checkpoints) when K > 1.
 * resid_hist: [K][hist_len] or NULL; st: [K]. Returns the most severe per-RHS code. */
int adda_iterative_solver_batch(adda_hip_op *op, const adda_solver_opts *opt, int K, const void *d_einc, void *d_x,
	void *d_pvec, double *resid_hist, int hist_len, adda_solve_stats *st);

/* ---- iteration checkpoint files (SaveIterChpoint / LoadIterChpoint, src/iterative.c:165-278) ----------------- */
/* The common scalars of a checkpoint, in the order of the file. */
typedef struct {
	int niter;           /* the next iteration */
	int counter;         /* iterations without residual decrease */
	double inprodR;      /* ||r||^2 of the last accepted iteration */
	double prev_err;     /* relative residual of the last iteration */
	double resid_scale;  /* 1/||b||^2 */
} adda_chp_head;
/* For method (enum adda_iter): its index in the reference's solver table (written first), its own scalars counted in
 * doubles (a doublecomplex counts two), its own vectors beyond xvec, rvec and pvec. */
int adda_chp_layout(int method, int *ind_m, int *nscal, int *nvec);
/* Write <dir>/chp.0 (creating dir) and <dir>/chp.log; vecs: xvec, rvec, pvec and the method's own vectors, each
 * nrows values on the host. */
int adda_chp_write(const char *dir, const char *run_dir, int method, size_t nrows, const adda_chp_head *h,
	const double *scal, const adda_cplx *const *vecs);
/* Read <dir>/chp.0 written for method and nrows (the reference's errors for another method, size or length). */
int adda_chp_read(const char *dir, int method, size_t nrows, adda_chp_head *h, double *scal, adda_cplx *const *vecs);

/* ---- scattering quantities (src/crosssec.c, src/CalculateE.c) ------------------------------------------------ */
/* CalcFieldFree for ndir directions (src/crosssec.c:506-595): ebuff[3*i..3*i+2] for unit vector dirs[3*i..]. */
int adda_calc_field(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	const double *dirs, int ndir, adda_cplx *ebuff);
/* ExtCross, plane wave (src/crosssec.c:811-823). */
int adda_ext_cross(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	const double prop[3], const double incpol[3], double *cext);
/* AbsCross, DRAINE (src/crosssec.c:852-884); cc: [nmat][3]. */
int adda_abs_cross(adda_hip_op *op, const adda_particle *p, double wavenum, const adda_cplx *cc, const void *d_pvec,
	double *cabs);
/* CalcEplaneYZ for one solution (src/CalculateE.c:423-507): eplane[2i] = E.per, eplane[2i+1] = E.par at
 * theta_i = i*180/(ntheta-1), i < npoints (npoints = 2*(ntheta-1) for the 360-degree range of an asymmetric
 * scatterer, src/calculator.c:671-674). */
int adda_eplane_yz(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace, const void *d_pvec,
	const double prop[3], const double par[3], const double per[3], int ntheta, int npoints, adda_cplx *eplane);
/* ComputeMuellerMatrix over the yz plane (src/CalculateE.c:83-116,236-246); out[16*npoints]. */
void adda_mueller_yz(const adda_cplx *eplX, const adda_cplx *eplY, int npoints, double *out);

/* The same quantities for the K right-hand sides of a batched solve (d_pvec a batched device vector), each RHS with
 * its own beam: props, incpols, pars, pers are [K][3], cc is [K][nmat][3], ebuff [K][ndir][3], eplane
 * [K][npoints][2]. ADDA computes them once per orientation (calculate_one_orientation, src/calculator.c:293-346);
 * here each kernel runs once for the whole batch, and every RHS's result equals the single-RHS function's bit for
 * bit. */
int adda_calc_field_batch(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace,
	const void *d_pvec, int K, const double *dirs, int ndir, adda_cplx *ebuff);
int adda_ext_cross_batch(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace,
	const void *d_pvec, int K, const double *props, const double *incpols, double *cext);
int adda_abs_cross_batch(adda_hip_op *op, const adda_particle *p, double wavenum, const adda_cplx *cc,
	const void *d_pvec, int K, double *cabs);
int adda_eplane_yz_batch(adda_hip_op *op, const adda_particle *p, double wavenum, double gridspace,
	const void *d_pvec, int K, const double *props, const double *pars, const double *pers, int ntheta, int npoints,
	adda_cplx *eplane);

/* ---- adjoint gradient (no counterpart in the reference: derivatives of its Cext and Cabs) ---------------------- */
/* d(cc)/d(eps) of adda_couple_constant at eps = m^2. */
adda_cplx adda_couple_constant_deps(adda_cplx m, int polrel, double wavenum, double gridspace, const double prop[3],
	const double incpol[3]);
/* Gradient of Cext and Cabs with respect to every dipole's permittivity eps_j = m_j^2 (the permittivity of the
 * dipole's material; with one material per dipole, nmat = nd, a per-voxel gradient) for the converged forward
 * solution d_pvec (P) of incident field d_einc with polarization (prop, incpol): two adjoint solves with the same
 * operator -- A is complex symmetric, so A^T = A -- for the right-hand sides conj(Einc) and conj(W.P), solved as one
 * batch with opt (method, eps, precision), then per-dipole assembly on the GPU. m: refractive index per material.
 * grad_ext, grad_abs: [2*nd] = (dC/dRe eps_j, dC/dIm eps_j); st: [2] stats of the two adjoint solves.
 * Resets the operator's cc_sqrt tables 0 and 1 to this polarization's. */
int adda_adjoint_gradient(adda_hip_op *op, const adda_particle *p, const adda_solver_opts *opt, const adda_cplx *m,
	int polrel, double wavenum, double gridspace, const double prop[3], const double incpol[3], const void *d_einc,
	const void *d_pvec, double *grad_ext, double *grad_abs, adda_solve_stats *st);

#ifdef __cplusplus
}
#endif
#endif
