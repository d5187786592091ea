// fft_device.h -- device-side building blocks of the zero-skipping FFT convolution (CDNA4, wave64).
//
// Complex numbers are interleaved (re,im) pairs of T = double or float. Every FFT here is an unnormalized DFT with
// forward sign -1 and backward sign +1, the convention of ADDA's FFT layer (src/fft.h:31-32, src/fft.c:651-652).
//
// The transform of a batch of lines lives in LDS. Stages are Stockham autosort stages of radix 2, 3, 4, 5 or 8
// (natural order in, natural order out): every thread gathers the inputs of its butterflies into registers, the
// block synchronises, and the outputs are written back to the same LDS buffer -- one buffer per line, no ping-pong.
#pragma once
#include <hip/hip_runtime.h>

namespace adda {

template <typename T> struct alignas(2 * sizeof(T)) cx { T x, y; };

template <typename T> __host__ __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cx<T> add(cx<T> a, cx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cx<T> sub(cx<T> a, cx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cx<T> mul(cx<T> a, cx<T> b)
{
	return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ cx<T> conjc(cx<T> a) { return mk<T>(a.x, -a.y); }
// Loads and stores with the nontemporal hint (the line is not kept in the caches for reuse).
template <typename T> __device__ __forceinline__ cx<T> ld_nt(const cx<T> *p)
{
	typedef T v2 __attribute__((ext_vector_type(2)));
	const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
	return mk<T>(v.x, v.y);
}
template <typename T> __device__ __forceinline__ void st_nt(cx<T> *p, cx<T> a)
{
	typedef T v2 __attribute__((ext_vector_type(2)));
	__builtin_nontemporal_store(v2{a.x, a.y}, reinterpret_cast<v2 *>(p));
}
// The element at a 32-bit byte offset from a base pointer. With a base that is the same for the whole block the
// access takes the scalar-base form of the global load/store: one 32-bit add per access instead of a 64-bit product
// and add (an index scaled in 64 bits costs multiply-adds that issue at a quarter of the VALU rate).
template <typename P> __device__ __forceinline__ P &at_b(P *base, unsigned byte_off)
{
	return *reinterpret_cast<P *>(reinterpret_cast<char *>(base) + byte_off);
}
template <typename P> __device__ __forceinline__ const P &at_b(const P *base, unsigned byte_off)
{
	return *reinterpret_cast<const P *>(reinterpret_cast<const char *>(base) + byte_off);
}
// precision conversion (identity when the types agree)
template <typename T, typename U> __device__ __forceinline__ cx<T> cvt(cx<U> a) { return mk<T>((T)a.x, (T)a.y); }
template <typename T> __device__ __forceinline__ cx<T> scl(cx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __device__ __forceinline__ cx<T> mulmi(cx<T> a)
{
	return INV ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// Maximum number of Stockham stages of a plan and radices supported by the kernels (2, 3, 4, 5, 8 and 16; 7, 11 and 13
// for the grids of ADDA's FFTW rule).
constexpr int kMaxStages = 12;

struct FftPlan {
	int n;                    // line length
	int nst;                  // number of stages
	int rad[kMaxStages];      // radix of each stage (8, 4, 2, 3 or 5)
};

// ---- in-register DFTs of size R -----------------------------------------------------------------------------------
template <typename T, bool INV> __device__ __forceinline__ void dft2(cx<T> *v)
{
	cx<T> a = v[0], b = v[1];
	v[0] = add(a, b);
	v[1] = sub(a, b);
}

template <typename T, bool INV> __device__ __forceinline__ void dft4(cx<T> *v)
{
	cx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]), t2 = add(v[1], v[3]);
	cx<T> t3 = mulmi<T, INV>(sub(v[1], v[3]));
	v[0] = add(t0, t2);
	v[2] = sub(t0, t2);
	v[1] = add(t1, t3);
	v[3] = sub(t1, t3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft8(cx<T> *v)
{
	const T h = (T)0.70710678118654752440084436210484904;  // 1/sqrt(2)
	cx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
	dft4<T, INV>(e);
	dft4<T, INV>(o);
	// w8^s, s = 0..3 (forward: exp(-i pi s/4); inverse: conjugate)
	cx<T> o1 = INV ? mk<T>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)) : mk<T>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
	cx<T> o2 = mulmi<T, INV>(o[2]);
	cx<T> o3 = INV ? mk<T>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)) : mk<T>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
	v[0] = add(e[0], o[0]);
	v[4] = sub(e[0], o[0]);
	v[1] = add(e[1], o1);
	v[5] = sub(e[1], o1);
	v[2] = add(e[2], o2);
	v[6] = sub(e[2], o2);
	v[3] = add(e[3], o3);
	v[7] = sub(e[3], o3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft3(cx<T> *v)
{
	const T s3 = (T)0.86602540378443864676372317075293618;  // sqrt(3)/2
	cx<T> t = add(v[1], v[2]);
	cx<T> m = mk<T>(v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y);
	cx<T> d = mulmi<T, INV>(scl(sub(v[1], v[2]), s3));
	v[0] = add(v[0], t);
	v[1] = add(m, d);
	v[2] = sub(m, d);
}

template <typename T, bool INV> __device__ __forceinline__ void dft5(cx<T> *v)
{
	const T c1 = (T)0.30901699437494742410229341718281906;   // cos(2pi/5)
	const T c2 = (T)-0.80901699437494742410229341718281906;  // cos(4pi/5)
	const T s1 = (T)0.95105651629515357211643933337938214;   // sin(2pi/5)
	const T s2 = (T)0.58778525229247312916870595463907277;   // sin(4pi/5)
	cx<T> t1 = add(v[1], v[4]), t2 = add(v[2], v[3]), t3 = sub(v[1], v[4]), t4 = sub(v[2], v[3]);
	cx<T> a1 = mk<T>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
	cx<T> a2 = mk<T>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
	cx<T> b1 = mk<T>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
	cx<T> b2 = mk<T>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
	b1 = mulmi<T, INV>(b1);  // -i b (forward)
	b2 = mulmi<T, INV>(b2);
	v[0] = add(v[0], add(t1, t2));
	v[1] = add(a1, b1);
	v[4] = sub(a1, b1);
	v[2] = add(a2, b2);
	v[3] = sub(a2, b2);
}

// cos(2 pi m/R) and sin(2 pi m/R), m = 1..(R-1)/2, for the odd prime radices of FFTW-sized grids (40 digits)
template <int R> struct PrimeTw;
template <> struct PrimeTw<7> {
	static constexpr double C[3] = {0.6234898018587335305250048840042398106, -0.2225209339563144042889025644967947594,
		-0.9009688679024191262361023195074450511};
	static constexpr double S[3] = {0.7818314824680298087084445266740577502, 0.9749279121818236070181316829939312172,
		0.4338837391175581204757683328483587546};
};
template <> struct PrimeTw<11> {
	static constexpr double C[5] = {0.8412535328311811688618116489193677175, 0.4154150130018864255292741492296232035,
		-0.1423148382732851404437926686163696687, -0.6548607339452850640569250724662935531,
		-0.9594929736144973898903680570663276990};
	static constexpr double S[5] = {0.5406408174555975821076359543186916954, 0.9096319953545183714117153830790284601,
		0.9898214418809327323760920377767187874, 0.7557495743542582837740358439723444202,
		0.2817325568414296977114179153466168990};
};
template <> struct PrimeTw<13> {
	static constexpr double C[6] = {0.8854560256532098959003755220150988786, 0.5680647467311558025118075591275166245,
		0.1205366802553230533490676874525435823, -0.3546048870425356259696378926000184743,
		-0.7485107481711010986346305997013513838, -0.9709418174260520271569822762937892272};
	static constexpr double S[6] = {0.4647231720437685456560153351331047776, 0.8229838658936563945796174234393819907,
		0.9927088740980539928007516494925201793, 0.9350162426854148234397845998378307291,
		0.6631226582407952023767854926667662795, 0.2393156642875577671487537262602118952};
};

// DFT of odd prime size R (7, 11, 13) on the symmetric pairs s_j = v_j + v_{R-j}, d_j = v_j - v_{R-j}:
// X_k = v_0 + sum_j s_j cos(2 pi jk/R) -/+ i sum_j d_j sin(2 pi jk/R) (forward/inverse), X_{R-k} the conjugate
// combination -- the form of dft5 for any odd prime.
template <typename T, bool INV, int R> __device__ __forceinline__ void dftP(cx<T> *v)
{
	constexpr int H = (R - 1) / 2;
	cx<T> sp[H], dm[H];
	cx<T> s0 = v[0];
#pragma unroll
	for (int j = 1; j <= H; j++) {
		sp[j - 1] = add(v[j], v[R - j]);
		dm[j - 1] = sub(v[j], v[R - j]);
		s0 = add(s0, sp[j - 1]);
	}
	const cx<T> x0 = v[0];
#pragma unroll
	for (int k = 1; k <= H; k++) {
		cx<T> a = x0, b = mk<T>(0, 0);
#pragma unroll
		for (int j = 1; j <= H; j++) {
			const int m = (j * k) % R;  // cos is even and sin odd about R/2: fold m into 1..H
			const T c = (T)PrimeTw<R>::C[(m <= H ? m : R - m) - 1];
			const T sn = (T)(m <= H ? PrimeTw<R>::S[m - 1] : -PrimeTw<R>::S[R - m - 1]);
			a = mk<T>(a.x + c * sp[j - 1].x, a.y + c * sp[j - 1].y);
			b = mk<T>(b.x + sn * dm[j - 1].x, b.y + sn * dm[j - 1].y);
		}
		const cx<T> bi = mulmi<T, INV>(b);  // -i b (forward) or +i b (inverse)
		v[k] = add(a, bi);
		v[R - k] = sub(a, bi);
	}
	v[0] = s0;
}

template <typename T, bool INV, int R> __device__ __forceinline__ void dftR(cx<T> *v)
{
	if constexpr (R == 2) dft2<T, INV>(v);
	else if constexpr (R == 3) dft3<T, INV>(v);
	else if constexpr (R == 4) dft4<T, INV>(v);
	else if constexpr (R == 5) dft5<T, INV>(v);
	else if constexpr (R == 7 || R == 11 || R == 13) dftP<T, INV, R>(v);
	else dft8<T, INV>(v);
}

// ---- one Stockham stage over a batch of lines in LDS ------------------------------------------------------------
// Element k of line l is buf[k*KS + l*LS]. COLMAJ: consecutive threads take consecutive lines (columns of a tile,
// KS = number of columns, LS = 1); otherwise consecutive threads walk along one line (KS = 1).
// Butterfly j (j < N/R) of a stage with accumulated span Ns reads elements j + q*N/R, applies twiddles
// w_N^{q*(j%Ns)*N/(Ns*R)} (table tw[e] = exp(-2 pi i e/N)), runs a DFT of size R and writes element
// (j/Ns)*Ns*R + j%Ns + q*Ns.  MAXIT bounds the butterflies one thread owns per stage.
template <typename T, bool INV, bool COLMAJ, int MAXIT, int R>
__device__ __forceinline__ void stockham_stage(cx<T> *buf, int N, int nlines, int KS, int LS, int Ns, const cx<T> *tw)
{
	const int nb = N / R, total = nlines * nb, twstep = N / (Ns * R);
	cx<T> v[MAXIT][R];
	int dst[MAXIT];
#pragma unroll
	for (int it = 0; it < MAXIT; it++) {
		int item = threadIdx.x + it * blockDim.x;
		dst[it] = -1;
		if (item < total) {
			int line, j;
			if (COLMAJ) { line = item % nlines; j = item / nlines; }
			else { j = item % nb; line = item / nb; }
			const int base = line * LS;
#pragma unroll
			for (int q = 0; q < R; q++) v[it][q] = buf[(j + q * nb) * KS + base];
			const int k = j % Ns;
			if (Ns > 1) {
#pragma unroll
				for (int q = 1; q < R; q++) {
					cx<T> w = tw[q * k * twstep];
					v[it][q] = mul(v[it][q], INV ? conjc(w) : w);
				}
			}
			dftR<T, INV, R>(v[it]);
			dst[it] = ((j / Ns) * Ns * R + k) * KS + base;
		}
	}
	__syncthreads();
#pragma unroll
	for (int it = 0; it < MAXIT; it++) {
		if (dst[it] >= 0) {
#pragma unroll
			for (int q = 0; q < R; q++) buf[dst[it] + q * Ns * KS] = v[it][q];
		}
	}
	__syncthreads();
}

// Full in-LDS FFT of nlines lines of length plan.n. The caller synchronises before (data in LDS) and the routine
// ends synchronised.
template <typename T, bool INV, bool COLMAJ, int MAXIT>
__device__ void lds_fft(cx<T> *buf, const FftPlan &plan, int nlines, int KS, int LS, const cx<T> *tw)
{
	int Ns = 1;
	const int N = plan.n;
	for (int s = 0; s < plan.nst; s++) {
		switch (plan.rad[s]) {
			case 8: stockham_stage<T, INV, COLMAJ, MAXIT, 8>(buf, N, nlines, KS, LS, Ns, tw); break;
			case 4: stockham_stage<T, INV, COLMAJ, MAXIT, 4>(buf, N, nlines, KS, LS, Ns, tw); break;
			case 2: stockham_stage<T, INV, COLMAJ, MAXIT, 2>(buf, N, nlines, KS, LS, Ns, tw); break;
			case 3: stockham_stage<T, INV, COLMAJ, MAXIT, 3>(buf, N, nlines, KS, LS, Ns, tw); break;
			case 7: stockham_stage<T, INV, COLMAJ, MAXIT, 7>(buf, N, nlines, KS, LS, Ns, tw); break;
			case 11: stockham_stage<T, INV, COLMAJ, MAXIT, 11>(buf, N, nlines, KS, LS, Ns, tw); break;
			case 13: stockham_stage<T, INV, COLMAJ, MAXIT, 13>(buf, N, nlines, KS, LS, Ns, tw); break;
			default: stockham_stage<T, INV, COLMAJ, MAXIT, 5>(buf, N, nlines, KS, LS, Ns, tw); break;
		}
		Ns *= plan.rad[s];
	}
}

}  // namespace adda
