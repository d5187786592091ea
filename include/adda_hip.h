/* adda_hip.h -- C ABI of the MI355X (gfx950) backend for ADDA's FFT matrix-vector product and Krylov vector kernels.
 *
 * Plain C: opaque handles, plain pointers and sizes, no HIP or torch types. Complex numbers are interleaved
 * (re,im) pairs, layout-compatible with ADDA's `doublecomplex` (C99 double complex, src/types.h:30).
 *
 * Every entry returns 0 on success or a negative code; adda_hip_last_error() gives "file:line: message" for the last
 * failure in the calling thread (the role of LogError/CL_CH_ERR, src/io.c:274-290, src/oclcore.h:58-75).
 *
 * Device vectors are `void *` device addresses of 3*nvoid_Ndip complex numbers (AoS per dipole: x,y,z), stored in
 * FP64 (ADDA_F64, 16 B/element) or FP32 (ADDA_F32, 8 B/element). All calls are ordered on the operator's single HIP
 * stream (one in-order queue, like src/oclcore.c:353-355); calls that return scalars synchronise that stream.
 */
#ifndef ADDA_HIP_H
#define ADDA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } adda_cplx;

enum { ADDA_F64 = 0, ADDA_F32 = 1 };

/* dot-product modes (src/linalg.c) */
enum {
	ADDA_DOT_CONJ = 0,       /* sum a*conj(b)                 nDotProd            src/linalg.c:75-88   */
	ADDA_DOT_BILIN = 1,      /* sum a*b                       nDotProd_conj       src/linalg.c:92-105  */
	ADDA_DOT_SELF_BILIN = 2, /* sum a*a (+ ||a||^2 in norm2)  nDotProdSelf_conj(_Norm2) src/linalg.c:109-144 */
	ADDA_DOT_NORM2 = 3       /* ||a||^2 in norm2              nNorm2              src/linalg.c:59-71   */
};

/* lincomb conjugation flags */
enum { ADDA_CONJ_A = 1, ADDA_CONJ_B = 2 };

typedef struct adda_hip_op adda_hip_op;

/* Last error message of the calling thread ("" if none). */
const char *adda_hip_last_error(void);

/* Number of visible HIP devices (0 when there is none). */
int adda_hip_device_count(void);

/* fftFit (src/fft.c:528-562) under the 2/3/5 rule that ADDA applies for clAmdFft (CLFFT_AMD, src/fft.c:542-545):
 * the first x' >= x that is even, divisible by divis and has no prime factor other than 2, 3, 5. Negative (error set)
 * when divis has another prime factor, so that no such size exists. */
int adda_hip_fftfit(int x, int divis);
/* FFT grid size chosen for speed (no reference counterpart; ADDA itself always takes fftFit): the smallest power of two
 * >= x divisible by divis when it is at most 25 % larger than adda_hip_fftfit(x, divis), else that size. The
 * convolution is exact for any grid >= 2*box, and power-of-two lines run on the radix-16/8 plans: a 108-point y-grid
 * (2^2*3^3, four radix-3/4 stages) costs twice as much per point as 128. Results differ only by rounding. A divis
 * with an odd factor gives adda_hip_fftfit's size; negative where adda_hip_fftfit is. */
int adda_hip_fastgrid(int x, int divis);

/* Creates the operator of one particle on `device` and builds its D-matrix on the GPU in FP64.
 * Replaces InitDmatrix (src/fft.c:918-1416, with InterTerm_poi, src/interaction.c:290-331) together with oclinit
 * (src/oclcore.c:327-437). position: ushort[3*nd] grid coordinates (src/vars.h:94); material: int[nd] indices into the
 * cc_sqrt table of nmat rows (src/vars.h:65; nmat may equal nd for per-dipole materials). grid*: padded FFT grid
 * (ParSetup, src/comm.c:489-494), each >= 2*box and even, of one of ADDA's fftFit rules (src/fft.c:528-562): 2/3/5-smooth
 * (Temperton FFT, clAmdFft), or also with factors of 7 and one 11 or 13 (FFTW3). flags: bit0 = also keep an FP32
 * copy of the D-matrix and FP32 work buffers (mixed precision).  Returns NULL on failure. */
adda_hip_op *adda_hip_op_create(int device, int boxX, int boxY, int boxZ, int gridX, int gridY, int gridZ, size_t nd,
	const unsigned short *position, const int *material, int nmat, double wavenum, double gridspace, int reduced_fft,
	int flags);

/* Free_FFT_Dmat (src/fft.c:1420-1493) + oclunload (src/oclcore.c:476). */
void adda_hip_op_destroy(adda_hip_op *op);

/* Uploads cc_sqrt[nmat][3] (InitCC's device write, src/calculator.c:283-288); may be called between solves. */
int adda_hip_set_cc_sqrt(adda_hip_op *op, const adda_cplx *cc_sqrt);

/* Copies the D-matrix to the host in ADDA's layout Dmatrix[6*((x*DsizeZ+z)*DsizeY+y)+comp] (IndexDmatrix,
 * src/fft.c:144-150); dst holds 6*gridX*DsizeY*DsizeZ complex numbers. For checks against a CPU build. With the
 * reduced FFT the device stores only x <= gridX/2 (G is even or odd in x as in y and z); columns x > gridX/2 are
 * filled from column gridX - x with the xy and xz components negated. */
int adda_hip_get_dmatrix(adda_hip_op *op, adda_cplx *dst);
size_t adda_hip_dmatrix_size(const adda_hip_op *op);

/* Particle above a plane substrate, image-dipole reflection (-surf with -int_surf img): builds the R-matrix on the GPU
 * (InitRmatrix, src/fft.c:819-891, with ReflTerm_img, src/interaction.c:1081-1153) and from then on every matvec adds
 * the reflected term (src/matvec.c:249,288-302). zsum_shift: ZsumShift = 2*(hsub/gridspace - (boxZ-1)/2)
 * (src/make_particle.c:2312), > 0; rcn: surfRCn = -1 for a perfect reflector, else (1-msub^2)/(1+msub^2)
 * (src/interaction.c:1411). Needs reduced_fft and gridZ >= 2*boxZ. The z-backward transform of the reflected source
 * is taken from the forward one at plane -kz, which P3 already holds (no extra transforms). Tuned matvec
 * configurations whose P3 variant has no reflected term are replaced by the defaults. */
int adda_hip_set_surface(adda_hip_op *op, double zsum_shift, adda_cplx rcn);

/* Copies the R-matrix to the host in ADDA's layout Rmatrix[6*((x*gridZ+z)*RsizeY+y)+comp], RsizeY = gridY/2+1
 * (IndexRmatrix_mv, src/matvec.c:107-121); the device stores x <= gridX/2 (columns beyond are filled from gridX - x
 * with xy and xz negated). For checks against a CPU build. */
int adda_hip_get_rmatrix(adda_hip_op *op, adda_cplx *dst);
size_t adda_hip_rmatrix_size(const adda_hip_op *op);

/* Incident field as a sum of nw (1 or 2) plane waves with complex wave vectors kvec[nw][3] (units of wavenum; an
 * evanescent wave has a complex one) and complex amplitudes amp[nw][3]: Einc_i = sum_w amp_w exp(i k kvec_w.r_i),
 * r_i as in adda_hip_plane_wave. The device half of GenerateB's substrate branch (src/GenerateB.c:258-322): incident
 * plus reflected wave from above, transmitted wave from below; FP64 device vector; synchronises. */
int adda_hip_plane_waves(adda_hip_op *op, double wavenum, double gridspace, int nw, const adda_cplx *kvec,
	const adda_cplx *amp, void *einc);

/* Device vectors: allocation (zero-filled), host <-> device copies (FP64 host data), precision conversion. A freed
 * vector is kept by the operator for the next allocation of the same size (reuse is ordered on the operator's stream,
 * so no synchronisation is needed) and released at adda_hip_op_destroy; freeing a pointer the operator did not hand
 * out is an error. */
void *adda_hip_vec_alloc(adda_hip_op *op, int prec);
int adda_hip_vec_free(adda_hip_op *op, void *v);
int adda_hip_vec_upload(adda_hip_op *op, int prec, void *dst, const adda_cplx *src);
int adda_hip_vec_download(adda_hip_op *op, int prec, adda_cplx *dst, const void *src);
int adda_hip_vec_convert(adda_hip_op *op, void *dst, int dprec, const void *src, int sprec);
size_t adda_hip_nrows(const adda_hip_op *op);  /* local_nRows = 3*nd */

/* MatVec (src/matvec.c:128-425; OpenCL twin src/oclmatvec.c:35-184) on device vectors:
 * result = A.arg with A = I + S.D.S (her: conj(A.conj(arg))); when inprod != NULL, *inprod = ||result||^2
 * (synchronises). argvec is not modified. prec selects FP64 or FP32 arithmetic and storage. */
int adda_hip_matvec(adda_hip_op *op, int prec, const void *argvec, void *resultvec, double *inprod, int her);

/* The same with host vectors (the shape of src/oclmatvec.c:35-40: copy in, compute, copy out), FP64. */
int adda_hip_matvec_host(adda_hip_op *op, const adda_cplx *argvec, adda_cplx *resultvec, double *inprod, int her);

/* Fused vector update out = ca*a' + cb*b' + cc*c (a' = conj(a) if ADDA_CONJ_A, b' likewise) over 3*nd elements;
 * any of a, b, c may be NULL (term dropped); out may alias a, b or c. When norm2 != NULL, *norm2 = ||out||^2
 * (synchronises). Covers nIncrem*, nLinComb*, nSubtr, nMult*, nCopy, nInit (src/linalg.c:148-605). */
int adda_hip_lincomb(adda_hip_op *op, int prec, void *out, const void *a, adda_cplx ca, const void *b, adda_cplx cb,
	const void *c, adda_cplx cc, int conj_flags, double *norm2);

/* Inner products (src/linalg.c:59-144) with FP64 accumulation and a deterministic two-level reduction
 * (MyInnerProduct's role, src/comm.c:454-477); synchronises. res and/or norm2 may be NULL if not needed. */
int adda_hip_dot(adda_hip_op *op, int prec, const void *a, const void *b, int mode, adda_cplx *res, double *norm2);

/* nMult_mat (src/linalg.c:549-569): out[3i+c] = cc_sqrt[mat[i]][c] * in[3i+c]. */
int adda_hip_mult_mat(adda_hip_op *op, int prec, void *out, const void *in);

/* Plane-wave incident field on the dipoles (GenerateB, src/GenerateB.c:324-330): Einc_i = exp(i k r_i.prop) e_pol,
 * r_i = (position_i - (box-1)/2)*gridspace (DipoleCoord, src/make_particle.c:2266-2275); FP64 device vector. */
int adda_hip_plane_wave(adda_hip_op *op, double wavenum, double gridspace, const double prop[3],
	const double incpol[3], void *einc);

/* Scattering sums of CalcFieldFree (src/crosssec.c:548-584) for ndir unit vectors n:
 * sums[3*i+c] = sum_j P[3j+c] exp(-i kd n_i.position_j), kd = wavenum*gridspace; deterministic reduction. */
int adda_hip_field_sums(adda_hip_op *op, int prec, const void *pvec, double kd, const double *dirs, int ndir,
	adda_cplx *sums);

/* sum_j sum_c w[mat_j][c] |v[3j+c]|^2 (the AbsCross sum, src/crosssec.c:873-884); w is [nmat][3]. */
int adda_hip_weighted_norm2(adda_hip_op *op, int prec, const void *v, const double *w, double *out);

/* Batched forms of the two sums above for K right-hand sides (a batched device vector, K <= the reserved batch):
 * adda_hip_field_sums_batch takes ndir directions per RHS, dirs[K][ndir][3] -> sums[K][ndir][3];
 * adda_hip_weighted_norm2_batch takes one weight table per RHS, w[K][nmat][3] -> out[K].
 * ADDA evaluates CalcFieldFree (src/crosssec.c:506-595) and AbsCross (src/crosssec.c:873-884) once per orientation
 * (calculate_one_orientation, src/calculator.c:293-346, called by orient_integrand, src/calculator.c:348-358); these
 * replace that loop for the post-processing of a batched solve, in one launch per kernel. Each RHS's result equals
 * the single-RHS call bit for bit. */
int adda_hip_field_sums_batch(adda_hip_op *op, int prec, const void *pvec, int K, double kd, const double *dirs,
	int ndir, adda_cplx *sums);
int adda_hip_weighted_norm2_batch(adda_hip_op *op, int prec, const void *v, int K, const double *w, double *out);

/* Measurement: runs nrep matvecs and times each of the 5 passes with HIP events on the operator's stream;
 * pass_ms[5] receives the average duration of P1 (scatter + FFT-x), P2 (FFT-z), P3 (FFT-y, D-matrix product,
 * inverse FFT-y), P4 (inverse FFT-z), P5 (inverse FFT-x + gather); total_ms the average per matvec of another nrep
 * matvecs issued back to back with events at the two ends only (as a solver issues them; the events between the passes
 * add gaps of their own). (The role of ADDA's PRECISE_TIMING breakdown, src/matvec.c:147-422.) */
int adda_hip_profile_matvec(adda_hip_op *op, int prec, const void *argvec, void *resultvec, int nrep, double *pass_ms,
	double *total_ms);

/* Algorithmic HBM bytes of one matvec pass under the streaming model (each pass reads and writes its operand once,
 * zero half-planes skipped, D-matrix read once): pass 0..4 as above, or 5 for the whole matvec. */
double adda_hip_matvec_bytes(const adda_hip_op *op, int prec, int pass);

/* ---- device-resident Krylov iterations ------------------------------------------------------------------------
 * The FP64 scalar recurrences of QMR_CS (src/iterative.c:898-1032), BiCGStab (640-724), CGNR (730-770), BiCG_CS
 * (549-630), CSYM (775-888), QMR_CS_2 (1037-1136) and BCGS2 (342-483) run on the GPU: one call enqueues one whole
 * iteration (matvecs, fused vector updates with their reductions, one-block scalar kernels, ProgressReport 282-318)
 * on the operator's stream without synchronising. The state is an array of
 * ADDA_DK_NSLOTS doubles in device memory plus a residual history; slot 6 holds the status (0 running, 1 converged,
 * 2 maxiter, -1 breakdown, -2 stagnation), and once it is non-zero every later enqueued kernel returns at once.
 * Slots: 0 inprodR, 1 inprodRp1, 2 epsB, 3 1/||b||^2, 4 niter, 5 counter, 6 status, 7 maxiter, 8 mc, 9 history
 * length, 10 breakdown code, 16.. method scalars (set by the caller at PHASE_INIT, see host/iterative.c). niter is
 * the iteration number the caller is enqueueing (it selects the niter==1 and niter==2 forms); matvec_ready: the
 * first iteration reuses A.b held in Av (QMR_CS, BiCG_CS, QMR_CS_2; BiCGStab: in t; BCGS2: in u1; CSYM never
 * does). CSYM and
 * QMR_CS rename their vectors after every iteration (the caller swaps q_new/q_old, v/vtilda every time and
 * p_new/p_old from the second on). Vectors are in precision prec. */
enum { ADDA_DK_NSLOTS = 96, ADDA_DK_STATUS = 6, ADDA_DK_NITER = 4, ADDA_DK_SC = 16 };
void *adda_hip_dk_create(adda_hip_op *op, int hist_len);
int adda_hip_dk_free(adda_hip_op *op, void *state);
int adda_hip_dk_set(adda_hip_op *op, void *state, int slot, const double *vals, int n);
int adda_hip_dk_get(adda_hip_op *op, void *state, int slot, double *vals, int n); /* synchronises */
int adda_hip_dk_history(adda_hip_op *op, void *state, double *hist, int n);      /* synchronises */
int adda_hip_dk_qmr_iter(adda_hip_op *op, int prec, void *state, int niter, int matvec_ready, void *v, void *vtilda,
	void *Av, void *p_old, void *p_new, void *x, void *r);
int adda_hip_dk_bicgstab_iter(adda_hip_op *op, int prec, void *state, int niter, int matvec_ready, void *x, void *r,
	void *p, void *v, void *s, void *rtilda, void *t);
int adda_hip_dk_cgnr_iter(adda_hip_op *op, int prec, void *state, int niter, void *x, void *r, void *p, void *Av);
int adda_hip_dk_bicgcs_iter(adda_hip_op *op, int prec, void *state, int niter, int matvec_ready, void *x, void *r,
	void *p, void *Av);
int adda_hip_dk_qmr2_iter(adda_hip_op *op, int prec, void *state, int niter, int matvec_ready, void *v, void *d,
	void *p, void *Av, void *x, void *r);
int adda_hip_dk_csym_iter(adda_hip_op *op, int prec, void *state, int niter, void *q_new, void *q_old, void *p_new,
	void *p_old, void *Av, void *x);
/* Launch list of the batched matvecs of K right-hand sides (2 <= K <= 64), no counterpart in the reference (its
 * solves are one RHS at a time): from now on those matvecs launch blocks only for the RHS k with run[k] != 0 and
 * neither read nor write the others, as if their gate status were non-zero (a gate still applies to the listed ones).
 * A gated RHS costs the dispatch of its blocks even though they exit at once; a solver that knows which RHS are still
 * running (krylov_dev_batch after each chunk) saves that. run == NULL, or no entry set, clears the list. The batched
 * solver (adda_iterative_solver_batch) replaces any list with its own and clears it on return; autotuning and
 * adda_hip_profile_matvec_batch time full launches whatever is listed. */
int adda_hip_set_launch_list(adda_hip_op *op, int K, const int *run);
int adda_hip_dk_bcgs2_iter(adda_hip_op *op, int prec, void *state, int niter, int matvec_ready, void *x, void *rtilda,
	void *r0, void *r1, void *r2, void *u0, void *u1, void *u2);

/* Matvec configuration per precision: radix schedules along x, y, z and tile sizes, as a string
 * "x=16.16 y=16.16 z=16.16 LY=4 TC=16 LP=2" (x radices: P1/P5; y: P3; z: P2/P4; LY y-rows (all three components) per
 * P1/P5 block, TC columns per P2/P4 block, LP column units per P3 block), optionally followed by the P3 options
 * "R3=1" (D-matrix product in registers), "XP=0" (single columns instead of column pairs), "TG=1" (twiddles read
 * from global memory) and "PT=1" (one thread per D-matrix element). set_config validates it against the grid. */
int adda_hip_get_config(adda_hip_op *op, int prec, char *buf, size_t len);
int adda_hip_set_config(adda_hip_op *op, int prec, const char *cfg);

/* Transform precision of FP64 matvecs (no reference counterpart: ADDA's MatVec computes in the vector precision).
 * mode 0: FP64 matvecs run in FP64 (the default). mode 1: FP64 matvecs (adda_hip_matvec*, and the matvecs inside the
 * adda_hip_dk_* iterations with ADDA_F64) read and write FP64 vectors but run the FFTs and the D-matrix product in
 * FP32 (the operator's FP32 D-matrix, created with flags & 1), adding the FP32 correction S.X to the FP64 argument in
 * FP64 -- Krylov iterations with FP64 vectors and recurrences around an inexact matvec. FP32 matvecs are unaffected. */
int adda_hip_set_matvec_mode(adda_hip_op *op, int mode);

/* Autotuner: times every candidate radix schedule (radix-16/8/4/2 splits of the power of two, radix 3/5 stages
 * before or after) and tile size of each pass group (x: P1+P5, z: P2+P4, y: P3) with HIP events, reps timed runs
 * each, and keeps the fastest. With cache_path, a line {"key": ..., "config": ...} is looked up first (no timing if
 * found) and appended after tuning. report (may be NULL) receives the chosen configuration and pass times. */
int adda_hip_autotune(adda_hip_op *op, int prec, int reps, const char *cache_path, char *report, size_t report_len);

/* ---- batched multi-RHS solves ----------------------------------------------------------------------------------
 * K right-hand sides (polarizations x orientations of one particle) share the operator: the D-matrix, the geometry
 * and the plans. ADDA solves them one after another (the orientation loop of src/crosssec.c:82-138 and the two
 * polarizations of src/CalculateE.c:836-870, each an IterativeSolver call); here every launch of a matvec or Krylov
 * kernel covers all K. A batched device vector holds the K vectors at stride adda_hip_batch_stride(op) complex
 * elements (>= 3*nd). Each RHS has its own cc_sqrt table (the LDR polarizability depends on the incident direction
 * and polarization, src/calculator.c:170-289). gate_state (a state from adda_hip_dk_create_batch, or NULL): RHS k is
 * left untouched when its status slot is non-zero (converged or excluded). Single-RHS calls are K = 1. */
int adda_hip_reserve_batch(adda_hip_op *op, int K);    /* grows work buffers and tables to K RHS (synchronises) */
size_t adda_hip_batch_stride(const adda_hip_op *op);
int adda_hip_set_cc_sqrt_batch(adda_hip_op *op, int K, const adda_cplx *cc_sqrt); /* cc_sqrt[K][nmat][3] */
void *adda_hip_vec_alloc_batch(adda_hip_op *op, int prec, int K);
int adda_hip_vec_convert_batch(adda_hip_op *op, int K, void *dst, int dprec, const void *src, int sprec);
int adda_hip_matvec_batch(adda_hip_op *op, int prec, int K, const void *gate_state, const void *argvec,
	void *resultvec, double *inprod /* [K] or NULL */, int her);
int adda_hip_lincomb_batch(adda_hip_op *op, int prec, int K, const void *gate_state, void *out, const void *a,
	const adda_cplx *ca, const void *b, const adda_cplx *cb, const void *c, const adda_cplx *cc, int conj_flags,
	double *norm2); /* coefficients and norm2: [K] */
int adda_hip_dot_batch(adda_hip_op *op, int prec, int K, const void *gate_state, const void *a, const void *b,
	int mode, adda_cplx *res, double *norm2); /* res, norm2: [K] */
/* adda_hip_dot_batch without a gate, for callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc. No counterpart in the reference. It allocates nothing, copies nothing to the
 * host and does not synchronise: the results go to the caller's device buffer d_out, [K][3] doubles {re, im, norm2}
 * (the product and 0 for ADDA_DOT_CONJ / ADDA_DOT_BILIN; sum a*a and ||a||^2 for the self modes), equal bit for bit to
 * adda_hip_dot_batch's, on the operator's stream. The partial sums use a static device array of 3*2048*16 doubles
 * (768 KiB per GPU, held until the process ends): 3*nb*K doubles must fit, nb = min(2048, ceil(3*nd/256)) -- 16 RHS
 * at the most blocks; larger requests fail with an error before anything is launched. Every call on a GPU shares that
 * array: calls that overlap on one GPU (other streams or threads) race. */
int adda_hip_dot_noalloc(adda_hip_op *op, int prec, int K, const void *a, const void *b, int mode, double *d_out);
int adda_hip_mult_mat_batch(adda_hip_op *op, int prec, int K, void *out, const void *in);
int adda_hip_profile_matvec_batch(adda_hip_op *op, int prec, int K, const void *argvec, void *resultvec, int nrep,
	double *pass_ms, double *total_ms);
double adda_hip_matvec_bytes_batch(const adda_hip_op *op, int prec, int K, int pass);
/* Solver state of K RHS: slots of RHS k at k*ADDA_DK_NSLOTS (adda_hip_dk_set/get address them so), histories at
 * k*hist_len. The dk_*_iter calls then advance all K recurrences, with batched vectors. */
void *adda_hip_dk_create_batch(adda_hip_op *op, int K, int hist_len);
int adda_hip_autotune_batch(adda_hip_op *op, int prec, int K, int reps, const char *cache_path, char *report,
	size_t report_len);

/* HIP graphs: capture_begin starts recording the operator's stream (nothing runs), capture_end instantiates what
 * was enqueued since (NULL on failure); graph_launch replays it on the stream. Used for chunks of dk_*_iter calls,
 * which contain no allocation, copy, memset or synchronisation. */
int adda_hip_capture_begin(adda_hip_op *op);
void *adda_hip_capture_end(adda_hip_op *op);
int adda_hip_graph_launch(adda_hip_op *op, void *graph);
int adda_hip_graph_free(adda_hip_op *op, void *graph);

/* ---- adjoint gradient ------------------------------------------------------------------------------------------
 * No counterpart in the reference (ADDA computes no derivatives); these assemble d(Cext, Cabs)/d(eps_j) from the
 * solutions of the forward and adjoint systems, see adda_adjoint_gradient in adda.h. */
/* out[3j+c] = conj(w[mat_j][c] * in[3j+c]) (FP64; w = [nmat][3] host table, NULL: plain conjugate) -- the adjoint
 * right-hand sides conj(E_inc) and conj(W.P). */
int adda_hip_weighted_conj(adda_hip_op *op, void *out, const void *in, const double *w);
/* Per dipole j with material m: h = -sum_c q[m][c] lam_c P_c, z = -sum_c q[m][c] mu_c P_c,
 * loc = sum_c q[m][c] |P_c|^2; grad_ext[2j] = ce Im h, grad_ext[2j+1] = ce Re h, grad_abs[2j] = ca (2 Re z - Im loc),
 * grad_abs[2j+1] = ca (-2 Im z - Re loc) (host arrays of 2*nd). q: [nmat][3] = d(1/cc)/d eps. FP64 vectors. */
int adda_hip_adjoint_grad(adda_hip_op *op, const void *pvec, const void *lam, const void *mu, const adda_cplx *q,
	double ce, double ca, double *grad_ext, double *grad_abs);
/* Copies n complex elements starting at element `first` of a device vector to the host (FP64 host data). */
int adda_hip_vec_download_part(adda_hip_op *op, int prec, adda_cplx *dst, const void *src, size_t first, size_t n);

/* Waits for all work queued on the operator's stream. */
int adda_hip_sync(adda_hip_op *op);

/* Measurement, no counterpart in the reference: achievable HBM rates on `device` of a plain copy, a read-only pass, a
 * fill and a copy with nontemporal loads and stores, over nbytes (one 16-byte element per thread), timed with HIP
 * events over nrep launches each; gbps[0..3] = copy (read + write bytes), read, fill, nontemporal copy in GB/s. The
 * ceilings the matvec passes are compared with beside the 8 TB/s peak. */
int adda_hip_hbm_rates(int device, size_t nbytes, int nrep, double *gbps);

#ifdef __cplusplus
}
#endif
#endif
