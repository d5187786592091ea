// fft_core.h -- register/LDS Stockham FFT core with the first stage loading from global memory and the last stage
// storing to global memory through caller functors (CDNA4, wave64).
//
// A line of length N = R_0*R_1*...*R_{s-1} (radices 16, 8, 4, 2, 3, 5) is owned by TPL = N/RMAX threads (RMAX = the
// largest radix of the plan). Stage s has N/R_s butterflies; thread t of a line owns butterflies j = t + u*TPL.
// Stage 0 reads its inputs straight from the Load functor (global memory, zero padding, gathers), the middle stages
// exchange through LDS, and the last stage hands each output (index j + q*N/R) to the Store functor. A 256-point
// transform therefore costs one LDS exchange (two radix-16 stages).
//
// Unnormalized DFT, forward sign -1, backward +1 (the convention of ADDA's FFT layer, src/fft.h:31-32).
#pragma once
#include "fft_device.h"

#include <type_traits>

namespace adda {

struct FftPlan2 {
	int n;                 // line length
	int nst;               // stages
	int rad[kMaxStages];   // radices
	int rmax;              // largest radix (threads per line = n / rmax)
};

// DFT-16 as 4 x 4 with the W16 twiddles between the two radix-4 layers.
template <typename T, bool INV> __device__ __forceinline__ void dft16(cx<T> *v)
{
	const T c1 = (T)0.92387953251128675612818318939678829;  // cos(pi/8)
	const T s1 = (T)0.38268343236508977172845998403039887;  // sin(pi/8)
	const T h = (T)0.70710678118654752440084436210484904;
	cx<T> a[4][4];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		cx<T> t[4] = {v[r], v[r + 4], v[r + 8], v[r + 12]};
		dft4<T, INV>(t);
#pragma unroll
		for (int s = 0; s < 4; s++) a[r][s] = t[s];
	}
	// twiddle a[r][s] *= W16^(r*s) (forward exp(-2 pi i rs/16))
	auto tw = [&](cx<T> x, T c, T s) -> cx<T> {
		return INV ? mk<T>(x.x * c - x.y * s, x.x * s + x.y * c) : mk<T>(x.x * c + x.y * s, x.y * c - x.x * s);
	};
	a[1][1] = tw(a[1][1], c1, s1);
	a[1][2] = tw(a[1][2], h, h);
	a[1][3] = tw(a[1][3], s1, c1);
	a[2][1] = tw(a[2][1], h, h);
	a[2][2] = mulmi<T, INV>(a[2][2]);
	a[2][3] = tw(a[2][3], -h, h);
	a[3][1] = tw(a[3][1], s1, c1);
	a[3][2] = tw(a[3][2], -h, h);
	a[3][3] = tw(a[3][3], -c1, -s1);
#pragma unroll
	for (int s = 0; s < 4; s++) {
		cx<T> t[4] = {a[0][s], a[1][s], a[2][s], a[3][s]};
		dft4<T, INV>(t);
#pragma unroll
		for (int q = 0; q < 4; q++) v[s + 4 * q] = t[q];
	}
}

template <typename T, bool INV, int R> __device__ __forceinline__ void dftR2(cx<T> *v)
{
	if constexpr (R == 16) dft16<T, INV>(v);
	else dftR<T, INV, R>(v);
}

// DFTs whose upper half of inputs (v[q], q >= R/2) is zero -- the first stage of a transform of zero-padded data,
// where every input index j + q*N/R with q >= R/2 lies at or beyond N/2 >= the box size. Only v[0..R/2) is read; the
// additions with the known zeros are skipped, the outputs are those of the full DFT.
template <typename T, bool INV> __device__ __forceinline__ void dft4h(cx<T> *v)
{
	const cx<T> a = v[0], b = v[1], c = mulmi<T, INV>(b);
	v[0] = add(a, b);
	v[2] = sub(a, b);
	v[1] = add(a, c);
	v[3] = sub(a, c);
}

template <typename T, bool INV> __device__ __forceinline__ void dft8h(cx<T> *v)
{
	const T h = (T)0.70710678118654752440084436210484904;
	cx<T> e[4] = {v[0], v[2]}, o[4] = {v[1], v[3]};
	dft4h<T, INV>(e);
	dft4h<T, INV>(o);
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

template <typename T, bool INV> __device__ __forceinline__ void dft16h(cx<T> *v)
{
	const T c1 = (T)0.92387953251128675612818318939678829;
	const T s1 = (T)0.38268343236508977172845998403039887;
	const T h = (T)0.70710678118654752440084436210484904;
	cx<T> a[4][4];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		cx<T> t[4] = {v[r], v[r + 4]};
		dft4h<T, INV>(t);
#pragma unroll
		for (int s = 0; s < 4; s++) a[r][s] = t[s];
	}
	auto tw = [&](cx<T> x, T c, T s) -> cx<T> {
		return INV ? mk<T>(x.x * c - x.y * s, x.x * s + x.y * c) : mk<T>(x.x * c + x.y * s, x.y * c - x.x * s);
	};
	a[1][1] = tw(a[1][1], c1, s1);
	a[1][2] = tw(a[1][2], h, h);
	a[1][3] = tw(a[1][3], s1, c1);
	a[2][1] = tw(a[2][1], h, h);
	a[2][2] = mulmi<T, INV>(a[2][2]);
	a[2][3] = tw(a[2][3], -h, h);
	a[3][1] = tw(a[3][1], s1, c1);
	a[3][2] = tw(a[3][2], -h, h);
	a[3][3] = tw(a[3][3], -c1, -s1);
#pragma unroll
	for (int s = 0; s < 4; s++) {
		cx<T> t[4] = {a[0][s], a[1][s], a[2][s], a[3][s]};
		dft4<T, INV>(t);
#pragma unroll
		for (int q = 0; q < 4; q++) v[s + 4 * q] = t[q];
	}
}

template <typename T, bool INV, int R> __device__ __forceinline__ void dftR2h(cx<T> *v)
{
	static_assert(R % 2 == 0, "half-zero inputs need an even radix");
	if constexpr (R == 16) dft16h<T, INV>(v);
	else if constexpr (R == 8) dft8h<T, INV>(v);
	else if constexpr (R == 4) dft4h<T, INV>(v);
	else v[1] = v[0];
}

// LDS position of element k of a line: lines stored as rows (KS == 1) get one pad element after every 16, so the
// stride-R writes of a Stockham stage (R up to 16 complex numbers apart) fall into different LDS banks; column tiles
// (KS = number of columns, consecutive threads on consecutive columns) need no padding.
__device__ __forceinline__ int lpos(int k, int KS) { return KS == 1 ? k + (k >> 4) : k * KS; }
// Row strides of padded lines of length n (n + n/16 elements at least), chosen by how the lanes of a wave walk the
// lines. Measured on gfx950 (ds_read_b64 for FP32 complex, ds_read_b128 for FP64 complex):
//  - line-major (16 consecutive lanes per line, several lines per wave): lines whose starts differ by 16 (mod 32)
//    elements are conflict-free; at 17 (mod 32) the accesses take 1.6x (b64) to 1.9x (b128) longer;
//  - line-fastest (consecutive lanes on consecutive lines at the same position): an odd stride spreads the lines over
//    all banks, while 16 (mod 32) puts every other line on the same banks.
// Kernels whose transforms run line-major throughout (FFT-y) use rowstride_lm; kernels that also switch to the
// line-fastest mapping (FFT-x with its transposed global access) use rowstride_odd.
__host__ __device__ __forceinline__ int rowstride_lm(int n)
{
	const int r = n + n / 16;
	return r + (48 - r % 32) % 32;
}
__host__ __device__ __forceinline__ int rowstride_odd(int n) { return (n + n / 16) | 1; }

// LDS layouts of a block's lines for the compile-time plans (element k of line l; KS > 1: column tiles, element k of
// column l at k*KS + l). LayPad: padded rows (k + k/16), line stride LS from rowstride_lm / rowstride_odd. LaySwz: rows
// without padding, the 16-element groups of a row permuted by XOR (offset o of group g at o ^ (g mod 16)), which keeps
// the stride-16 butterfly accesses on distinct banks; FLIP also swaps neighbouring groups of odd lines, so that two
// lines read by one half-wave fall on opposite bank halves (line-major kernels, LS = N). Without FLIP, LS = N + 1 for
// kernels that also walk across lines. Measured on gfx950 both are conflict-free for the patterns of the kernel they
// serve (FP32 and FP64), and they take N instead of N + N/16 elements per line. Require N % 16 == 0.
struct LayPad {
	__device__ static __forceinline__ int pos(int l, int k, int KS, int LS) { return KS == 1 ? l * LS + k + (k >> 4) : k * KS + l; }
};
template <bool FLIP> struct LaySwz {
	__device__ static __forceinline__ int pos(int l, int k, int KS, int LS)
	{
		if (KS != 1) return k * KS + l;
		const int g = k >> 4, o = (k & 15) ^ (g & 15);
		return l * LS + 16 * (FLIP ? (g ^ (l & 1)) : g) + o;
	}
};

// Number of butterflies per thread of a stage of radix R when the largest radix of the plan is at most RM: a line has
// N/RMAX >= N/RM threads, so ceil(RM/R) butterflies per thread suffice. Kernels are instantiated for RM = 16 and
// RM = 8; the radix-8 variant holds half as many values per thread (fewer registers, more waves per SIMD).
template <int R, int RM> constexpr int maxu() { return (RM + R - 1) / R; }

// One stage. FIRST: inputs come from ld(line, k); LAST: outputs go to st(line, k, v); otherwise through LDS
// (element k of line l at buf[lpos(k, KS) + l*LS]). Twiddles from the LDS table tws[e] = exp(-2 pi i e/N).
// Control flow is uniform: every thread of the block executes the same barriers; `active` only predicates the work
// (a wave must never reach a barrier from two branches). A non-first stage synchronises between reading LDS and
// writing its outputs (the LAST stage's Store may target the same LDS), and a non-last stage synchronises after
// writing.
template <typename T, bool INV, int R, bool FIRST, bool LAST, int RM, class Ld, class St>
__device__ __forceinline__ void stage2(cx<T> *buf, const cx<T> *tws, int N, int Ns, int KS, int LS, int line, int t,
	int TPL, bool active, Ld &ld, St &st)
{
	constexpr int MU = maxu<R, RM>();
	const int nb = N / R, twstep = N / (Ns * R);
	cx<T> v[MU][R];
#pragma unroll
	for (int u = 0; u < MU; u++) {
		const int j = t + u * TPL;
		if (active && (MU == 1 || j < nb)) {
#pragma unroll
			for (int q = 0; q < R; q++) v[u][q] = FIRST ? ld(line, j + q * nb) : buf[lpos(j + q * nb, KS) + line * LS];
			if (!FIRST) {
				const int k = j % Ns;
#pragma unroll
				for (int q = 1; q < R; q++) {
					const cx<T> w = tws[q * k * twstep];
					v[u][q] = mul(v[u][q], INV ? conjc(w) : w);
				}
			}
			dftR2<T, INV, R>(v[u]);
		}
	}
	__syncthreads();  // a FIRST stage's Load may read the LDS buffer it is about to overwrite
	if (LAST) {
		// Ns*R == N here: output index j + q*Ns
#pragma unroll
		for (int u = 0; u < MU; u++) {
			const int j = t + u * TPL;
			if (active && (MU == 1 || j < nb)) {
#pragma unroll
				for (int q = 0; q < R; q++) st(line, j + q * Ns, v[u][q]);
			}
		}
	}
	else {
#pragma unroll
		for (int u = 0; u < MU; u++) {
			const int j = t + u * TPL;
			if (active && (MU == 1 || j < nb)) {
				const int k = j % Ns;
				const int d = (j / Ns) * Ns * R + k;
#pragma unroll
				for (int q = 0; q < R; q++) buf[lpos(d + q * Ns, KS) + line * LS] = v[u][q];
			}
		}
		__syncthreads();
	}
}

template <typename T, bool INV, bool FIRST, bool LAST, int RM, class Ld, class St>
__device__ __forceinline__ void stage_dispatch(int R, cx<T> *buf, const cx<T> *tws, int N, int Ns, int KS, int LS,
	int line, int t, int TPL, bool active, Ld &ld, St &st)
{
	switch (R) {
		case 16:
			if constexpr (RM >= 16) stage2<T, INV, 16, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st);
			break;
		case 8: stage2<T, INV, 8, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st); break;
		case 4: stage2<T, INV, 4, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st); break;
		case 2: stage2<T, INV, 2, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st); break;
		case 3: stage2<T, INV, 3, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st); break;
		// the radices of FFTW-sized grids are compiled into the RM = 16 instances only (mv_find_dyn sends every plan
		// with one of them there; radices above 8 need them anyway: a line then has N/RMAX >= N/16 threads)
		case 7:
			if constexpr (RM >= 16) stage2<T, INV, 7, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st);
			break;
		case 11:
			if constexpr (RM >= 16) stage2<T, INV, 11, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st);
			break;
		case 13:
			if constexpr (RM >= 16) stage2<T, INV, 13, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st);
			break;
		default: stage2<T, INV, 5, FIRST, LAST, RM>(buf, tws, N, Ns, KS, LS, line, t, TPL, active, ld, st); break;
	}
}

// Full transform of one line by its TPL threads. All threads of the block call this together (uniform control flow).
// The first stage, which reads global memory, uses the thread mapping (lineA, tA, activeA); the other stages, the
// last of which writes global memory, use (lineB, tB, activeB) -- so each global access can be mapped for contiguity
// (e.g. x-fastest for a dipole gather, y-fastest for a transposed store). The plan is uniform across the block.
template <typename T, bool INV, int RM, class Ld, class St>
__device__ __forceinline__ void fft_line(cx<T> *buf, const cx<T> *tws, const FftPlan2 &p, int KS, int LS, int lineA, int tA,
	bool activeA, int lineB, int tB, bool activeB, Ld &ld, St &st)
{
	const int N = p.n, TPL = N / p.rmax;
	if (p.nst == 1) {
		stage_dispatch<T, INV, true, true, RM>(p.rad[0], buf, tws, N, 1, KS, LS, lineA, tA, TPL, activeA, ld, st);
		return;
	}
	stage_dispatch<T, INV, true, false, RM>(p.rad[0], buf, tws, N, 1, KS, LS, lineA, tA, TPL, activeA, ld, st);
	int Ns = p.rad[0];
	for (int s = 1; s < p.nst - 1; s++) {
		stage_dispatch<T, INV, false, false, RM>(p.rad[s], buf, tws, N, Ns, KS, LS, lineB, tB, TPL, activeB, ld, st);
		Ns *= p.rad[s];
	}
	stage_dispatch<T, INV, false, true, RM>(p.rad[p.nst - 1], buf, tws, N, Ns, KS, LS, lineB, tB, TPL, activeB, ld, st);
}

// ------------------------------------------------------------------------------------------------------------------
// Compile-time plans. The generic path above takes the length and the radix schedule at run time, so every index of
// a stage (j % Ns, j / Ns, twiddle indices, padded LDS positions, line offsets) costs integer divisions and
// multiplies; in the FFT-y pass that integer work outnumbers the butterflies. Fix<R0, R1, ...> fixes N and the
// schedule at compile time: the same Stockham stages with all of that folded into shifts and immediate offsets.
// Kernels are templated on a plan policy: Dyn<RM> (run-time plan, RM = bound on the radix) or Fix<...>.

// HIN (first stage, even R): inputs j + q*NB with q >= R/2 are zero and are not loaded. HOUT (last stage, even R):
// only outputs j + q*NS with q < R/2 (indices below N/2) are stored, and the arithmetic of the others is dropped.
// Twiddles come from this stage's own table tws[(q-1)*NS + k] = exp(-2 pi i q k / (NS R)) (see Fix::load_tw): the
// threads of a line read consecutive entries, where the natural table's q*k strides conflict in the LDS banks.
// GIN / GOUT: the first stage's Load reads global memory (not the LDS buffer) / the last stage's Store writes global
// memory. Such a stage needs no barrier between its reads and its writes: nothing it writes can clobber an LDS input
// of another thread. (With GIN the caller must have no pending reads of the buffer; with GOUT the transform ends
// without a barrier, and the caller synchronises before reusing the buffer.)
template <typename T, bool INV, int N, int TPL, int NS, bool FIRST, bool LAST, int R, int RMAX, bool HIN, bool HOUT,
	bool GIN, bool GOUT, class LAY, class Ld, class St>
__device__ __forceinline__ void sstage(cx<T> *buf, const cx<T> *tws, int KS, int LS, int line, int t, bool active,
	Ld &ld, St &st)
{
	constexpr int NB = N / R;                    // butterflies per line
	constexpr int MU = (RMAX + R - 1) / R;        // butterflies per thread (TPL = N / RMAX threads per line)
	constexpr bool HI = FIRST && HIN && R % 2 == 0, HO = LAST && HOUT && R % 2 == 0;
	cx<T> v[MU][R];
#pragma unroll
	for (int u = 0; u < MU; u++) {
		const int j = t + u * TPL;
		if (active && (NB % TPL == 0 || j < NB)) {
#pragma unroll
			for (int q = 0; q < (HI ? R / 2 : R); q++) {
				// a Load functor that takes a third argument also gets the input's slot u*R + q (the order of
				// Fix::for_first), e.g. to return a value it gathered before the transform
				if constexpr (!FIRST) v[u][q] = buf[LAY::pos(line, j + q * NB, KS, LS)];
				else if constexpr (std::is_invocable_v<Ld &, int, int, int>) v[u][q] = ld(line, j + q * NB, u * R + q);
				else v[u][q] = ld(line, j + q * NB);
			}
			if (!FIRST) {
				const int k = j % NS;
#pragma unroll
				for (int q = 1; q < R; q++) {
					const cx<T> w = tws[(q - 1) * NS + k];
					v[u][q] = mul(v[u][q], INV ? conjc(w) : w);
				}
			}
			if constexpr (HI) dftR2h<T, INV, R>(v[u]);
			else dftR2<T, INV, R>(v[u]);
		}
	}
	// every LDS read of this stage (by any thread) precedes its LDS writes
	if constexpr (!(FIRST && GIN) && !(LAST && GOUT)) __syncthreads();
	if (LAST) {
#pragma unroll
		for (int u = 0; u < MU; u++) {
			const int j = t + u * TPL;
			if (active && (NB % TPL == 0 || j < NB)) {
#pragma unroll
				for (int q = 0; q < (HO ? R / 2 : R); q++) {
					// a Store functor that takes a fourth argument also gets the output's slot u*R + q (the order of
					// Fix::for_last), e.g. to find data it prefetched for that output
					if constexpr (std::is_invocable_v<St &, int, int, cx<T>, int>) st(line, j + q * NS, v[u][q], u * R + q);
					else st(line, j + q * NS, v[u][q]);
				}
			}
		}
	}
	else {
#pragma unroll
		for (int u = 0; u < MU; u++) {
			const int j = t + u * TPL;
			if (active && (NB % TPL == 0 || j < NB)) {
				const int d = (j / NS) * NS * R + j % NS;
#pragma unroll
				for (int q = 0; q < R; q++) buf[LAY::pos(line, d + q * NS, KS, LS)] = v[u][q];
			}
		}
		__syncthreads();
	}
}

template <int A, int... B> constexpr int cmax()
{
	if constexpr (sizeof...(B) == 0) return A;
	else return A > cmax<B...>() ? A : cmax<B...>();
}

// ------------------------------------------------------------------------------------------------------------------
// Three lines per thread: the x, y and z components of one grid line, carried through the same butterflies so that a
// 3x3 product per frequency can be applied in registers. e[c][u*R + q] holds element (u, q) of component c in the
// butterfly layout of the current stage (thread t owns butterflies j_u = t + u*TPL, u < RMAX/R). A stage takes its
// inputs from global memory (Ld functor ld(c, k)), LDS (lines lb + c) or the registers left by the previous stage,
// and leaves its outputs in LDS, in registers (element (u, q) = frequency j_u + q*NS), or in global memory
// (St functor st(c, k, v)). Requires every radix to divide RMAX (power-of-two plans).
enum { IO_GLOBAL = 0, IO_LDS = 1, IO_REG = 2 };

template <typename T, bool INV, int N, int TPL, int RMAX, int NS, int R, int IN, int OUT, class Ld, class St>
__device__ __forceinline__ void stage3(cx<T> (&e)[3][RMAX], cx<T> *buf, const cx<T> *tws, int LS, int lb, int t,
	bool active, Ld &ld, St &st)
{
	static_assert(RMAX % R == 0, "three-line stages need every radix to divide the largest one");
	constexpr int NB = N / R, MU = RMAX / R, TW = N / (NS * R);
	if (active) {
#pragma unroll
		for (int u = 0; u < MU; u++) {
			const int j = t + u * TPL;
			if constexpr (IN == IO_GLOBAL) {
#pragma unroll
				for (int c = 0; c < 3; c++)
#pragma unroll
					for (int q = 0; q < R; q++) e[c][u * R + q] = ld(c, j + q * NB);
			}
			else if constexpr (IN == IO_LDS) {
#pragma unroll
				for (int c = 0; c < 3; c++)
#pragma unroll
					for (int q = 0; q < R; q++) e[c][u * R + q] = buf[(lb + c) * LS + lpos(j + q * NB, 1)];
			}
			if (NS > 1) {
				const int k = j % NS;
#pragma unroll
				for (int q = 1; q < R; q++) {
					const cx<T> w = tws[q * k * TW];
					const cx<T> wq = INV ? conjc(w) : w;
#pragma unroll
					for (int c = 0; c < 3; c++) e[c][u * R + q] = mul(e[c][u * R + q], wq);
				}
			}
#pragma unroll
			for (int c = 0; c < 3; c++) dftR2<T, INV, R>(&e[c][u * R]);
		}
	}
	if constexpr (OUT == IO_LDS) {
		__syncthreads();  // every LDS read of this stage (by any thread) precedes the writes
		if (active) {
#pragma unroll
			for (int u = 0; u < MU; u++) {
				const int j = t + u * TPL;
				const int d = (j / NS) * NS * R + j % NS;
#pragma unroll
				for (int c = 0; c < 3; c++)
#pragma unroll
					for (int q = 0; q < R; q++) buf[(lb + c) * LS + lpos(d + q * NS, 1)] = e[c][u * R + q];
			}
		}
		__syncthreads();
	}
	else if constexpr (OUT == IO_GLOBAL) {
		if (active) {
#pragma unroll
			for (int u = 0; u < MU; u++) {
				const int j = t + u * TPL;
#pragma unroll
				for (int c = 0; c < 3; c++)
#pragma unroll
					for (int q = 0; q < R; q++) st(c, j + q * NS, e[c][u * R + q]);
			}
		}
	}
}

template <int RM> struct Dyn {
	// the HIN / HOUT options are ignored: every input is loaded and every output stored
	static constexpr bool HALF_IN = false, HALF_OUT = false;
	static constexpr int RMAXC = 0;    // largest radix unknown at compile time
	static constexpr bool SWZ = false;  // padded LDS rows (LayPad)
	// the natural twiddle table tws[e] = exp(-2 pi i e/N), e < N. No barrier: the caller's next __syncthreads (every
	// transform reaches one before its first twiddle) publishes the table, so its loads overlap the caller's.
	template <typename T>
	__device__ static __forceinline__ void load_tw(cx<T> *tws, const cx<T> *__restrict__ twg, int N)
	{
		for (int i = threadIdx.x; i < N; i += blockDim.x) tws[i] = twg[i];
	}
	__device__ static __forceinline__ int n(const FftPlan2 &p) { return p.n; }
	__device__ static __forceinline__ int tpl(const FftPlan2 &p) { return p.n / p.rmax; }
	// HIN / HOUT (zero upper half of the input, unneeded upper half of the output) and GIN / GOUT (global Load /
	// Store, fewer barriers) only apply to compile-time plans: the run-time path keeps every barrier
	template <typename T, bool INV, bool HIN = false, bool HOUT = false, bool GIN = false, bool GOUT = false,
		class LAY = LayPad, class Ld, class St>
	__device__ static __forceinline__ void fft(cx<T> *buf, const cx<T> *tws, const FftPlan2 &p, int KS, int LS,
		int lineA, int tA, bool activeA, int lineB, int tB, bool activeB, Ld &ld, St &st)
	{
		static_assert(std::is_same_v<LAY, LayPad>, "run-time plans use the padded layout");
		fft_line<T, INV, RM>(buf, tws, p, KS, LS, lineA, tA, activeA, lineB, tB, activeB, ld, st);
	}
};

template <int... Rs> struct Fix {
	static constexpr int NST = sizeof...(Rs);
	static constexpr int RAD[NST] = {Rs...};
	static constexpr int N = (Rs * ...);
	static constexpr int RMAX = cmax<Rs...>();
	static constexpr int TPL = N / RMAX;
	// whether HIN skips the inputs at indices >= N/2 and HOUT the outputs there (even first / last radix)
	static constexpr bool HALF_IN = RAD[0] % 2 == 0, HALF_OUT = RAD[NST - 1] % 2 == 0;
	static constexpr int RMAXC = RMAX;
	static constexpr bool SWZ = N % 16 == 0;  // swizzled LDS rows (LaySwz) where the length allows
	// Offset of stage S's twiddles in the staged table: stages 1..NST-1 in order, (R_s - 1)*NS_s entries each.
	static constexpr int twoff(int S)
	{
		int o = 0, ns = RAD[0];
		for (int s = 1; s < S; s++) {
			o += (RAD[s] - 1) * ns;
			ns *= RAD[s];
		}
		return o;
	}
	// Staged twiddle table from the natural one twg[e] = exp(-2 pi i e/N): stage s (span NS, radix R) holds
	// w[(q-1)*NS + k] = twg[q*k*N/(NS*R)], q = 1..R-1, k < NS. N - RAD[0] entries in all. No barrier (as Dyn).
	template <typename T, int S, int NS>
	__device__ static __forceinline__ void fill_tw(cx<T> *tws, const cx<T> *__restrict__ twg)
	{
		if constexpr (S < NST) {
			constexpr int R = RAD[S], TW = N / (NS * R), OFF = twoff(S);
			for (int e = threadIdx.x; e < (R - 1) * NS; e += blockDim.x) {
				const int q = e / NS + 1, k = e % NS;
				tws[OFF + e] = twg[q * k * TW];
			}
			fill_tw<T, S + 1, NS * R>(tws, twg);
		}
	}
	template <typename T>
	__device__ static __forceinline__ void load_tw(cx<T> *tws, const cx<T> *__restrict__ twg, int)
	{
		fill_tw<T, 1, RAD[0]>(tws, twg);
	}
	static bool match(const FftPlan2 &p)
	{
		if (p.n != N || p.nst != NST) return false;
		for (int i = 0; i < NST; i++)
			if (p.rad[i] != RAD[i]) return false;
		return true;
	}
	__device__ static __forceinline__ constexpr int n(const FftPlan2 &) { return N; }
	__device__ static __forceinline__ constexpr int tpl(const FftPlan2 &) { return TPL; }
	// The outputs of the last stage that thread t of a line stores: f(slot, k) for element k, in the order and with
	// the slot numbers the last stage passes to a four-argument Store functor (HOUT: only k < N/2 with an even last
	// radix). NSLOT bounds the slot numbers.
	static constexpr int RL = RAD[NST - 1], NBL = N / RL, MUL = (RMAX + RL - 1) / RL, NSLOT = MUL * RL;
	// The same for the inputs of the first stage that thread t of a line loads (HIN: only k < N/2 with an even first
	// radix), with the slot numbers the first stage passes to a three-argument Load functor.
	static constexpr int RF = RAD[0], NBF = N / RF, MUF = (RMAX + RF - 1) / RF, NSLOTF = MUF * RF;
	template <bool HIN, class F> __device__ static __forceinline__ void for_first(int t, F &&f)
	{
		constexpr bool HI = HIN && RF % 2 == 0;
#pragma unroll
		for (int u = 0; u < MUF; u++) {
			const int j = t + u * TPL;
			if (NBF % TPL == 0 || j < NBF) {
#pragma unroll
				for (int q = 0; q < (HI ? RF / 2 : RF); q++) f(u * RF + q, j + q * NBF);
			}
		}
	}
	template <bool HOUT, class F> __device__ static __forceinline__ void for_last(int t, F &&f)
	{
		constexpr bool HO = HOUT && RL % 2 == 0;
#pragma unroll
		for (int u = 0; u < MUL; u++) {
			const int j = t + u * TPL;
			if (NBL % TPL == 0 || j < NBL) {
#pragma unroll
				for (int q = 0; q < (HO ? RL / 2 : RL); q++) f(u * RL + q, j + q * NBL);
			}
		}
	}
	template <typename T, bool INV, bool HIN, bool HOUT, bool GIN, bool GOUT, class LAY, int S, int NS, class Ld,
		class St>
	__device__ static __forceinline__ void run(cx<T> *buf, const cx<T> *tws, int KS, int LS, int lineA, int tA,
		bool activeA, int lineB, int tB, bool activeB, Ld &ld, St &st)
	{
		constexpr int R = RAD[S];
		constexpr bool FIRST = S == 0, LAST = S == NST - 1;
		sstage<T, INV, N, TPL, NS, FIRST, LAST, R, RMAX, HIN, HOUT, GIN, GOUT, LAY>(buf, tws + (FIRST ? 0 : twoff(S)),
			KS, LS, FIRST ? lineA : lineB, FIRST ? tA : tB, FIRST ? activeA : activeB, ld, st);
		if constexpr (!LAST)
			run<T, INV, HIN, HOUT, GIN, GOUT, LAY, S + 1, NS * R>(buf, tws, KS, LS, lineA, tA, activeA, lineB, tB,
				activeB, ld, st);
	}
	// HIN: inputs at indices >= N/2 are zero (zero-padded data); HOUT: outputs at indices >= N/2 are not needed.
	// Either may leave the Load / Store functor uncalled for those indices. GIN / GOUT: the Load reads / the Store
	// writes global memory, not the LDS buffer (see sstage). tws: the staged table of load_tw.
	template <typename T, bool INV, bool HIN = false, bool HOUT = false, bool GIN = false, bool GOUT = false,
		class LAY = LayPad, class Ld, class St>
	__device__ static __forceinline__ void fft(cx<T> *buf, const cx<T> *tws, const FftPlan2 &, int KS, int LS,
		int lineA, int tA, bool activeA, int lineB, int tB, bool activeB, Ld &ld, St &st)
	{
		run<T, INV, HIN, HOUT, GIN, GOUT, LAY, 0, 1>(buf, tws, KS, LS, lineA, tA, activeA, lineB, tB, activeB, ld, st);
	}

	// Three-line transforms for the register D-product (stage3): the forward FFT takes the radices in plan order,
	// loads from global memory and leaves the last stage's outputs in registers (frequency t + u*TPL + q*NS_LAST);
	// the inverse takes them in reverse order, so its first stage consumes exactly those registers, and stores to
	// global memory.
	static constexpr bool POW2 = ((RMAX % Rs == 0) && ...);
	static constexpr int NS_LAST = N / RAD[NST - 1];
	template <typename T, int S, int NS, class Ld>
	__device__ static __forceinline__ void fwd3(cx<T> (&e)[3][RMAX], cx<T> *buf, const cx<T> *tws, int LS, int lb, int t,
		bool active, Ld &ld)
	{
		constexpr int R = RAD[S];
		constexpr int IN = S == 0 ? IO_GLOBAL : IO_LDS, OUT = S == NST - 1 ? IO_REG : IO_LDS;
		stage3<T, false, N, TPL, RMAX, NS, R, IN, OUT>(e, buf, tws, LS, lb, t, active, ld, ld);
		if constexpr (S < NST - 1) fwd3<T, S + 1, NS * R>(e, buf, tws, LS, lb, t, active, ld);
	}
	template <typename T, int S, int NS, class St>
	__device__ static __forceinline__ void inv3(cx<T> (&e)[3][RMAX], cx<T> *buf, const cx<T> *tws, int LS, int lb, int t,
		bool active, St &st)
	{
		constexpr int R = RAD[NST - 1 - S];
		constexpr int IN = S == 0 ? IO_REG : IO_LDS, OUT = S == NST - 1 ? IO_GLOBAL : IO_LDS;
		stage3<T, true, N, TPL, RMAX, NS, R, IN, OUT>(e, buf, tws, LS, lb, t, active, st, st);
		if constexpr (S < NST - 1) inv3<T, S + 1, NS * R>(e, buf, tws, LS, lb, t, active, st);
	}
};

// Host-side plan: the power of two as radix-16 stages with one radix-8/4/2 remainder (2^(4a+1) uses 8*4 instead of
// 16*2), then the odd primes 13, 11, 7, 5, 3 (7, 11 and 13 only in grids of ADDA's FFTW rule).
inline bool make_plan2(int n, FftPlan2 &p)
{
	p.n = n;
	p.nst = 0;
	int m = n, k2 = 0, r5 = 0, r3 = 0, r7 = 0, r11 = 0, r13 = 0;
	while (m % 13 == 0) { r13++; m /= 13; }
	while (m % 11 == 0) { r11++; m /= 11; }
	while (m % 7 == 0) { r7++; m /= 7; }
	while (m % 5 == 0) { r5++; m /= 5; }
	while (m % 3 == 0) { r3++; m /= 3; }
	while (m % 2 == 0) { k2++; m /= 2; }
	if (m != 1) return false;
	int rad[kMaxStages], ns = 0;
	int a = k2 / 4, r = k2 % 4;
	int extra[2] = {0, 0}, ne = 0;
	if (r == 1) {
		if (a > 0) { a--; extra[ne++] = 8; extra[ne++] = 4; }
		else extra[ne++] = 2;
	}
	else if (r == 2) extra[ne++] = 4;
	else if (r == 3) extra[ne++] = 8;
	if (a + ne + r3 + r5 + r7 + r11 + r13 > kMaxStages) return false;
	for (int i = 0; i < a; i++) rad[ns++] = 16;
	for (int i = 0; i < ne; i++) rad[ns++] = extra[i];
	for (int i = 0; i < r13; i++) rad[ns++] = 13;
	for (int i = 0; i < r11; i++) rad[ns++] = 11;
	for (int i = 0; i < r7; i++) rad[ns++] = 7;
	for (int i = 0; i < r5; i++) rad[ns++] = 5;
	for (int i = 0; i < r3; i++) rad[ns++] = 3;
	if (ns == 0) return false;  // n == 1 is not a valid grid
	p.rmax = 1;
	for (int i = 0; i < ns; i++) {
		p.rad[i] = rad[i];
		if (rad[i] > p.rmax) p.rmax = rad[i];
	}
	p.nst = ns;
	return true;
}

}  // namespace adda
