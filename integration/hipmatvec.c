/* hipmatvec.c -- the HIP (MI355X) flavor of ADDA's FFT MatVec, for a sequential ADDA build.
 *
 * It replaces two files of ADDA's source tree at link time, as the OpenCL flavor replaces matvec.c with oclmatvec.c
 * (src/ocl/Makefile:80-82):
 *   src/matvec.c  MatVec                                            (src/matvec.c:128-425)
 *   src/fft.c     InitDmatrix, Free_FFT_Dmat, fftFit, CheckNprocs   (src/fft.c:918-1416, 1420-1493, 528-562, 566-620)
 * and calls the C ABI of include/adda_hip.h (a-dda_amd/libadda_hip.so) instead: the D-matrix (and the R-matrix of the
 * substrate) is built on the GPU, and every MatVec runs there on ADDA's host vectors (argvec and resultvec cross PCIe,
 * like the OpenCL flavor's MatVec without its device-resident buffers).
 *
 * Build (integration/Makefile): every ADDA source except matvec.c and fft.c, with -DFFT_TEMPERTON so that no file
 * includes fftw3.h (src/fft.h:22-24, src/memory.c:29-31: plain malloc instead of fftw_malloc), plus this file and
 * -ladda_hip.
 *
 * Without edits of ADDA's sources this file is complete: the GPU index comes from the environment variable
 * ADDA_HIP_GPU, and MatVec uploads cc_sqrt when InitCC has changed it. A maintainer who adds the flavor to ADDA makes
 * three edits, each next to the OpenCL flavor's code:
 *   src/param.c:435 and 1048  '#ifdef OPENCL' -> '#if defined(OPENCL) || defined(HIP)': the -gpu <index> option sets
 *                             gpuInd, defined below (the OpenCL flavor defines it in src/oclcore.c:64);
 *   src/calculator.c:283-288  after the OpenCL upload of cc_sqrt in InitCC: '#ifdef HIP HipUploadCC(); #endif';
 *   src/iterative.c:50,1249,1310  '#ifdef OPENCL' -> '#if defined(OPENCL) || defined(HIP)' (and '!defined(HIP)' in
 *                             line 50): -init_field wkb then takes its scratch from the solver's vectors instead of
 *                             Xmatrix (src/vars.c:144), which fft.c allocates and this flavor leaves NULL.
 * With -DHIP the environment variable and the per-call cc_sqrt check are not used.
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>

#include "cmplx.h"
#include "comm.h"
#include "const.h"
#include "fft.h"
#include "io.h"
#include "memory.h"
#include "timing.h"
#include "types.h"
#include "vars.h"

#include "adda_hip.h"

#ifdef ADDA_MPI
#	error "the HIP flavor is sequential: one process per GPU"
#endif

// GLOBAL VARIABLES
int gpuInd;                  // -gpu <index> (src/param.c:1048-1054 with the edit above)

// defined and initialized in timing.c
extern size_t TotalMatVec;
extern TIME_TYPE Timing_FFT_Init,Timing_Dm_Init;
// defined and initialized in make_particle.c: distance between the lowest dipoles and their images (units of d)
extern double ZsumShift;
// defined and initialized in param.c
extern const enum init_field InitField;

static adda_hip_op *hop;      // the particle's operator on the GPU
static doublecomplex cc_dev[MAX_NMAT][3]; // the cc_sqrt last uploaded
static bool cc_set;

#define HIP_CH_ERR(call) do { if ((call)!=0) LogError(ONE_POS,"HIP: %s",adda_hip_last_error()); } while (0)

//======================================================================================================================

int fftFit(int x,int divis)
/* the first even x' >= x divisible by divis with prime factors 2, 3 and 5 only: the clAmdFft rule of src/fft.c:542-545,
 * which the GPU transforms support
 */
{
	return adda_hip_fftfit(x,divis);
}

//======================================================================================================================

void CheckNprocs(void)
// the sequential build runs one process
{
	if (nprocs!=1) LogError(ONE_POS,"The HIP flavor of ADDA runs one process (nprocs=%d)",nprocs);
}

//======================================================================================================================

void HipUploadCC(void)
// called by InitCC (src/calculator.c:283-288) after cc_sqrt is computed, for every incident polarization
{
	HIP_CH_ERR(adda_hip_set_cc_sqrt(hop,(const adda_cplx *)cc_sqrt));
	memcpy(cc_dev,cc_sqrt,sizeof(cc_dev));
	cc_set=true;
}

//======================================================================================================================

void InitDmatrix(void)
/* the operator of the particle on the GPU: D-matrix (InitDmatrix, src/fft.c:918-1416) and, near a substrate, the
 * R-matrix of the image-dipole reflection (InitRmatrix, src/fft.c:819-891) with surfRCn of src/interaction.c:1411
 */
{
	if (prognosis) return;
	TIME_TYPE tstart=GET_TIME();
#ifndef HIP
	const char *env=getenv("ADDA_HIP_GPU");
	if (env!=NULL) gpuInd=atoi(env);
#endif
	if (InitField==IF_WKB) LogError(ONE_POS,"-init_field wkb needs the edit of src/iterative.c described in "
		"hipmatvec.c");
	int *mat=intVector(local_nvoid_Ndip,ALL_POS,"material indices");
	for (size_t i=0;i<local_nvoid_Ndip;i++) mat[i]=material[i];
	hop=adda_hip_op_create(gpuInd,boxX,boxY,boxZ,(int)gridX,(int)gridY,(int)gridZ,local_nvoid_Ndip,position,mat,Nmat,
		WaveNum,gridspace,reduced_FFT,0);
	Free_general(mat);
	if (hop==NULL) LogError(ONE_POS,"HIP: %s",adda_hip_last_error());
	if (surface) {
		if (ReflRelation!=GR_IMG) LogError(ONE_POS,"The HIP flavor computes the reflected interaction through the "
			"image dipole only ('-int_surf img')");
		const doublecomplex rcn = msubInf ? -1 : (1-msub*msub)/(1+msub*msub);
		const adda_cplx r={creal(rcn),cimag(rcn)};
		HIP_CH_ERR(adda_hip_set_surface(hop,ZsumShift,r));
	}
	cc_set=false;
	Timing_FFT_Init=0;
	Timing_Dm_Init=GET_TIME()-tstart;
}

//======================================================================================================================

void Free_FFT_Dmat(void)
{
	adda_hip_op_destroy(hop);
	hop=NULL;
}

//======================================================================================================================

void MatVec (doublecomplex * restrict argvec,    // the argument vector
             doublecomplex * restrict resultvec, // the result vector
             double *inprod,         // the resulting inner product
             const bool her,         // whether Hermitian transpose of the matrix is used
             TIME_TYPE *timing,      // this variable is incremented by total time
             TIME_TYPE *comm_timing) // this variable is incremented by communication time
/* resultvec = A.argvec (her: conj(A.conj(argvec))) with A = I + S.(D+R).S on the GPU; *inprod = ||resultvec||^2 when
 * inprod != NULL; argvec unchanged (the contract of src/matvec.c:128-138)
 */
{
	TIME_TYPE tstart=GET_TIME();
#ifndef HIP
	// without the InitCC edit: upload cc_sqrt when it differs from the device copy (Nmat*3 values)
	if (!cc_set || memcmp(cc_dev,cc_sqrt,sizeof(cc_dev))!=0) HipUploadCC();
#endif
	HIP_CH_ERR(adda_hip_matvec_host(hop,(const adda_cplx *)argvec,(adda_cplx *)resultvec,inprod,her));
	(*timing)+=GET_TIME()-tstart;
	TotalMatVec++;
}
