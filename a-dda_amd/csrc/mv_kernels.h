// mv_kernels.h -- the five matvec pass kernels (P1 scatter + FFT-x, P2/P4 FFT-z, P3 FFT-y.D.FFT-y^-1, P5 inverse
// FFT-x + gather), templated on precision and FFT plan policy (fft_core.h: Dyn<RM> run-time plan, Fix<...>
// compile-time plan). Their instantiations live in several translation units (mv_*.hip) that compile in parallel;
// each exposes the kernels of its plans through MvKernels.
#pragma once
#include "fft_core.h"
#include "krylov_dev.h"

#include <type_traits>

namespace adda {

// Per-RHS layout of a batched launch: K right-hand sides share the operator (D-matrix, geometry, plans); the RHS
// index is part of the block index. A single solve is K = 1.
struct Bq {
	int K;               // right-hand sides in the launch
	long vs;             // stride of the dipole vectors (arg, result) between RHS, complex elements
	long xs, xs2;        // strides of the work buffers the pass reads / writes
	int cs;              // stride of the cc_sqrt tables (complex elements)
	const double *gate;  // RHS k is skipped when gate[k*DK_NSLOTS + DK_STATUS] != 0 (null: never)
	int pair = 0;        // P5: launch-order remap of the tiles (pair_tile), the launch padded to whole groups of 16
	int ntile = 0;       // P5 with pair: the number of real tiles (the padding blocks write zero partial sums)
	int kl = 1;          // P3: right-hand sides per block (consecutive RHS sharing the block's D rows and twiddles)
	// launch list (K <= 64): only the RHS whose bit is set get blocks, nl = popcount(mask) launch slots per tile; 0 =
	// every RHS has its slot (gated ones exit at once, which still costs their dispatch)
	unsigned long long mask = 0;
	int nl = 0;
};
// The launch slots per tile and the RHS of slot s: the s-th set bit of the mask (a uniform binary search on popcounts)
__device__ __forceinline__ int bq_nk(const Bq &b) { return b.mask ? b.nl : b.K; }
__device__ __forceinline__ int bq_rhs(const Bq &b, int s)
{
	if (!b.mask) return s;
	unsigned long long m = b.mask;
	int pos = 0;
	for (int w = 32; w > 0; w >>= 1) {
		const int c = __popcll(m & ((1ull << w) - 1));
		if (s >= c) { s -= c; m >>= w; pos += w; }
	}
	return pos;
}
// Tile of the b-th block of a launch whose tiles 2p and 2p + 1 sit 8 blocks apart (groups of 16 blocks: the even
// member of 8 pairs, then the odd ones). Blocks are dealt to the 8 XCDs in turn, so a pair shares one XCD and its L2.
// A bijection of [0, 16*ceil(n/16)); tiles >= n are padding.
__device__ __forceinline__ int pair_tile(int b)
{
	return 2 * ((b >> 4) * 8 + (b & 7)) + ((b >> 3) & 1);
}
static __device__ __forceinline__ bool skip_rhs(const Bq &b, int k)
{
	return b.gate && b.gate[(size_t)k * DK_NSLOTS + DK_STATUS] != 0;
}

// Deterministic block reduction for any block size up to 512: every thread of the block must call it; the total is
// returned to thread 0. Blocks of whole waves reduce each wave with a fixed xor butterfly (no barrier), then thread 0
// adds the wave sums in wave order; other block sizes use a fixed-order tree over LDS. red_lds: blockDim doubles.
template <typename T>
__device__ __forceinline__ double block_sum(double s, double *red_lds)
{
	const int tid = threadIdx.x, n = blockDim.x;
	__syncthreads();  // red_lds may still be read by a previous call (or be the caller's buffer)
	if (n % 64 == 0) {
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
		if ((tid & 63) == 0) red_lds[tid >> 6] = s;
		__syncthreads();
		double t = 0;
		if (tid == 0)
			for (int w = 0; w < n / 64; w++) t += red_lds[w];
		return t;
	}
	red_lds[tid] = s;
	__syncthreads();
	for (int stride = 256; stride > 0; stride >>= 1) {
		if (tid < stride && tid + stride < n) red_lds[tid] += red_lds[tid + stride];
		__syncthreads();
	}
	return tid == 0 ? red_lds[0] : 0.0;
}

// Three block sums at once, each equal bit for bit to block_sum of that value, with one pair of barriers instead of
// three (blocks of whole waves; other block sizes take three block_sum calls). The totals land in a, b, c of thread 0.
template <typename T>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *red_lds)
{
	const int tid = threadIdx.x, n = blockDim.x;
	if (n % 64 != 0) {
		a = block_sum<T>(a, red_lds);
		b = block_sum<T>(b, red_lds);
		c = block_sum<T>(c, red_lds);
		return;
	}
	__syncthreads();  // red_lds may still be read by a previous call (or be the caller's buffer)
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		a += __shfl_xor(a, off, 64);
		b += __shfl_xor(b, off, 64);
		c += __shfl_xor(c, off, 64);
	}
	const int nw = n / 64;
	if ((tid & 63) == 0) {
		red_lds[tid >> 6] = a;
		red_lds[nw + (tid >> 6)] = b;
		red_lds[2 * nw + (tid >> 6)] = c;
	}
	__syncthreads();
	if (tid == 0) {
		a = b = c = 0;
		for (int w = 0; w < nw; w++) a += red_lds[w];
		for (int w = 0; w < nw; w++) b += red_lds[nw + w];
		for (int w = 0; w < nw; w++) c += red_lds[2 * nw + w];
	}
}

// Box row (z, y) of the particle: with dense = 1 its dipoles are the consecutive indices s .. s+n-1 at the consecutive
// cells x0 .. x0+n-1 (a row without holes of a particle whose dipoles are enumerated in z, y, x order, as ADDA's are);
// an empty row is dense with n = 0. Other rows (holes, or another dipole order) are read through the cell map.
struct alignas(16) RowDesc {
	int s, x0, n, dense;
};

// ==================================================================================================================
// kernels: matvec pipeline, layout Xa[z<boxZ][x<gridX][y<boxY][c], Xb[z<gridZ][x<gridX][y<boxY][c],
// D[zt<DsizeZ][x<DsizeX][yt<DsizeY][6]. Every kernel loads its twiddle table into LDS first.

template <typename T>
__device__ __forceinline__ void load_tw(cx<T> *tws, const cx<T> *__restrict__ twg, int N)
{
	for (int i = threadIdx.x; i < N; i += blockDim.x) tws[i] = twg[i];
	__syncthreads();
}

// P1: block = (z, y0..y0+LY, RHS), all three components of LY y-rows; the RHS is fastest in the block index (the RHS
// of a tile share its cell-map and material reads). Each thread gathers the three components of one dipole --
// contiguous in the AoS vector -- into the LDS lines yl*3 + c (x along the line); the last FFT stage stores
// Xa[z][x][y][c] with (y, c) fastest across threads, runs of 3*LY elements. TV: the type of the dipole vector (FP64
// vectors around an FP32 transform: TV = double, T = float; otherwise TV = T).
template <typename T, class PL, typename TV = T>
__global__ void __launch_bounds__(512) k_p1(const cx<TV> *__restrict__ arg, cx<T> *__restrict__ Xa,
	const int *__restrict__ cellmap, const RowDesc *__restrict__, const int *__restrict__ mat,
	const cx<T> *__restrict__ ccs, FftPlan2 p, const cx<T> *__restrict__ twg, int bX, int bY, int LY, int her, Bq bq)
{
	const int nk = bq_nk(bq), kq = bq_rhs(bq, blockIdx.x % nk);
	const int tile = bq.pair ? pair_tile(blockIdx.x / nk) : blockIdx.x / nk;
	if (skip_rhs(bq, kq) || (bq.pair && tile >= bq.ntile)) return;  // pair: as in P5 (the Xa runs of tiles 2p, 2p + 1
	                                                                  // share cache lines)
	arg += kq * bq.vs;
	Xa += kq * bq.xs2;
	ccs += (size_t)kq * bq.cs;
	extern __shared__ char smem_raw[];
	using LAY = std::conditional_t<PL::SWZ, LaySwz<false>, LayPad>;  // lines are also walked line-fastest
	const int N = PL::n(p), TPL = PL::tpl(p), LS = PL::SWZ ? N + 1 : rowstride_odd(N), NL = 3 * LY;
	cx<T> *tws = reinterpret_cast<cx<T> *>(smem_raw), *buf = tws + N;
	PL::template load_tw<T>(tws, twg, N);
	const int nyt = (bY + LY - 1) / LY;
	const int z = tile / nyt, y0 = (tile % nyt) * LY;
	const int ny = bY - y0 < LY ? bY - y0 : LY;  // rows of this tile
	const int tid = threadIdx.x;
	const int lineA = tid / TPL, tA = tid % TPL;  // x fastest (reads the LDS lines)
	const int lineB = tid % NL, tB = tid / NL;    // (y, c) fastest (writes Xa)
	const bool actA = lineA < 3 * ny, actB = tB < TPL && lineB < 3 * ny;
	// the box fills at most half of the line (gridX >= 2*boxX); with a compile-time plan the transform never reads the
	// upper half, so it is not zero-filled
	const int NF = PL::HALF_IN ? N / 2 : N, total = LY * NF;
	// GU elements per thread and round: the cell-map reads of all of them are issued together, then their argument and
	// material reads, then the coupling constants -- three dependent loads in flight for GU elements at a time.
	constexpr int GU = 4;
	const int *cm = cellmap + ((size_t)z * bY + y0) * bX;
	const bool onemat = bq.cs == 3;  // one material: no per-dipole material lookup
	for (int b0 = tid; b0 < total; b0 += GU * blockDim.x) {
		int dd[GU], mm[GU] = {};
		cx<T> av[GU][3];
#pragma unroll
		for (int u = 0; u < GU; u++) {
			const int idx = b0 + u * blockDim.x, yl = idx / NF, x = idx - yl * NF;
			dd[u] = (idx < total && x < bX && yl < ny) ? cm[yl * bX + x] : -1;
		}
#pragma unroll
		for (int u = 0; u < GU; u++)
			if (dd[u] >= 0) {
				const cx<TV> *a = arg + 3 * (size_t)dd[u];
				mm[u] = onemat ? 0 : 3 * mat[dd[u]];
#pragma unroll
				for (int c = 0; c < 3; c++) av[u][c] = cvt<T>(her ? conjc(a[c]) : a[c]);
			}
#pragma unroll
		for (int u = 0; u < GU; u++) {
			const int idx = b0 + u * blockDim.x, yl = idx / NF, x = idx - yl * NF;
			if (idx < total) {
#pragma unroll
				for (int c = 0; c < 3; c++)
					buf[LAY::pos(3 * yl + c, x, 1, LS)] = dd[u] >= 0 ? mul(ccs[mm[u] + c], av[u][c]) : mk<T>(0, 0);
			}
		}
	}
	__syncthreads();
	auto ld = [&](int line, int x) -> cx<T> { return buf[LAY::pos(line, x, 1, LS)]; };
	cx<T> *xo = Xa + (size_t)z * N * 3 * bY + 3 * y0 + lineB;
	const size_t xs = (size_t)3 * bY;
	auto st = [&](int, int x, cx<T> v) { xo[x * xs] = v; };
	PL::template fft<T, false, true, false, false, true, LAY>(buf, tws, p, 1, LS, lineA, tA, actA, lineB, tB, actB, ld,
		st);
}

// P1 for compile-time plans, with the gather fused into the first FFT stage: that stage loads its inputs -- element x of
// line (yl, c) is cc_sqrt[c]*arg[3d + c] of the dipole d at cell x of row y0 + yl, zero for an empty cell -- straight
// from the dipole vector, so the gathered lines never make an LDS round trip of their own. Thread t of a line loads the
// cells t + 16q (radix 16): lanes of the three component lines of a row read the 48 (24) contiguous bytes of each
// dipole together. Dense rows (RowDesc) give the dipole index arithmetically, so the argument loads depend on one
// descriptor load per thread and never on a cell-map read. Same arithmetic as k_p1, so the results are the same.
// NT: the Xa stores with the nontemporal hint (a kernel of its own: the same instructions otherwise, so the same
// bits; the launcher picks it for configurations with XN bit 0).
template <typename T, class PL, typename TV = T, bool NT = false>
__global__ void __launch_bounds__(512) k_p1s(const cx<TV> *__restrict__ arg, cx<T> *__restrict__ Xa,
	const int *__restrict__ cellmap, const RowDesc *__restrict__ rows, const int *__restrict__ mat,
	const cx<T> *__restrict__ ccs, FftPlan2 p, const cx<T> *__restrict__ twg, int bX, int bY, int LY, int her, Bq bq)
{
	const int nk = bq_nk(bq), kq = bq_rhs(bq, blockIdx.x % nk);
	const int tile = bq.pair ? pair_tile(blockIdx.x / nk) : blockIdx.x / nk;
	if (skip_rhs(bq, kq) || (bq.pair && tile >= bq.ntile)) return;  // pair: as in P5 (the Xa runs of tiles 2p, 2p + 1
	                                                                  // share cache lines)
	arg += kq * bq.vs;
	Xa += kq * bq.xs2;
	ccs += (size_t)kq * bq.cs;
	extern __shared__ char smem_raw[];
	using LAY = std::conditional_t<PL::SWZ, LaySwz<false>, LayPad>;  // lines are also walked line-fastest
	constexpr int N = PL::N, TPL = PL::TPL;
	const int LS = PL::SWZ ? N + 1 : rowstride_odd(N), NL = 3 * LY;
	cx<T> *tws = reinterpret_cast<cx<T> *>(smem_raw), *buf = tws + N;
	PL::template load_tw<T>(tws, twg, N);
	const int nyt = (bY + LY - 1) / LY;
	const int z = tile / nyt, y0 = (tile % nyt) * LY;
	const int ny = bY - y0 < LY ? bY - y0 : LY;
	const int tid = threadIdx.x;
	const int lineA = tid / TPL, tA = tid % TPL;  // x fastest (the gather of the first stage)
	const int lineB = tid % NL, tB = tid / NL;    // (y, c) fastest (writes Xa)
	const bool actA = lineA < 3 * ny, actB = tB < TPL && lineB < 3 * ny;
	const int yl = lineA / 3, c = lineA - 3 * yl;
	const size_t row = (size_t)z * bY + y0 + yl;
	RowDesc rd = {0, 0, 0, 1};
	if (actA) rd = rows[row];
	const int *cm = cellmap + row * bX;
	const bool onemat = bq.cs == 3;  // one material: no per-dipole material lookup
	// The gather runs in phases over all inputs of the thread's first stage -- dipole indices, then material indices,
	// then arguments and coupling constants -- so each phase's loads are in flight together (one element at a time
	// would wait on each load in turn). The loads of empty cells read dipole 0 and are discarded: a load under a
	// condition gets a basic block of its own, and the compiler waits for it before the next.
	constexpr int NSF = PL::NSLOTF;
	// Indices are 32-bit (a dipole vector holds fewer than 2^32 elements). ai: element index 3d + c of the argument
	// (dipole 0 for an empty cell); on a dense row it is a0 + 3x, linear in the cell, with no product per element.
	int dd[NSF], mm[NSF];
	unsigned ai[NSF];
	cx<TV> av[NSF];
	cx<T> gv[NSF];
	if (rd.dense) {
		const unsigned a0 = 3u * (unsigned)(rd.s - rd.x0) + (unsigned)c;  // modulo 2^32, exact where used
		PL::template for_first<true>(tA, [&](int sl, int x) {
			const unsigned o = (unsigned)(x - rd.x0);
			const bool in = actA && x < bX && o < (unsigned)rd.n;
			dd[sl] = in ? rd.s + (int)o : -1;
			ai[sl] = in ? a0 + 3u * (unsigned)x : (unsigned)c;
		});
	}
	else {
		PL::template for_first<true>(tA, [&](int sl, int x) { dd[sl] = actA && x < bX ? cm[x] : -1; });
		PL::template for_first<true>(tA, [&](int sl, int) { ai[sl] = 3u * (unsigned)max(dd[sl], 0) + c; });
	}
	if (!onemat) PL::template for_first<true>(tA, [&](int sl, int) { mm[sl] = 3 * mat[(unsigned)max(dd[sl], 0)]; });
	else PL::template for_first<true>(tA, [&](int sl, int) { mm[sl] = 0; });
	PL::template for_first<true>(tA, [&](int sl, int) { gv[sl] = ccs[(unsigned)(mm[sl] + c)]; });
	PL::template for_first<true>(tA, [&](int sl, int) { av[sl] = arg[ai[sl]]; });
	PL::template for_first<true>(tA, [&](int sl, int) {
		const cx<T> a = cvt<T>(her ? conjc(av[sl]) : av[sl]);
		gv[sl] = dd[sl] >= 0 ? mul(gv[sl], a) : mk<T>(0, 0);
	});
	auto ld = [&](int, int, int sl) -> cx<T> { return gv[sl]; };
	// the tile's base is the block's; a thread's elements lie within one z-plane of Xa (byte offsets far below 2^32)
	cx<T> *xo = Xa + (size_t)z * N * 3 * bY + 3 * y0;
	const unsigned xsb = 3u * bY * (unsigned)sizeof(cx<T>), lb = (unsigned)lineB * (unsigned)sizeof(cx<T>);
	auto st = [&](int, int x, cx<T> v) {
		if constexpr (NT) st_nt(&at_b(xo, lb + (unsigned)x * xsb), v);
		else at_b(xo, lb + (unsigned)x * xsb) = v;
	};
	PL::template fft<T, false, true, false, true, true, LAY>(buf, tws, p, 1, LS, lineA, tA, actA, lineB, tB, actB, ld,
		st);
}

// P2 / P4: FFT along z of Q = gridX*3*boxY contiguous columns; rows k < n_in read (zero above), rows k < n_out written.
template <typename T, bool INV, class PL>
__global__ void __launch_bounds__(512) k_zfft(const cx<T> *__restrict__ in, cx<T> *__restrict__ out, FftPlan2 p,
	const cx<T> *__restrict__ twg, long Q, int TC, int n_in, int n_out, Bq bq)
{
	const int nk = bq_nk(bq), kq = bq_rhs(bq, blockIdx.x % nk);
	if (skip_rhs(bq, kq)) return;
	// pair: tiles narrower than a cache line -- tiles 2p and 2p + 1 share lines -- go 8 blocks apart, on one XCD
	// (pair_tile; the launch is padded to whole groups of 16 tiles)
	const int tile = bq.pair ? pair_tile(blockIdx.x / nk) : blockIdx.x / nk;
	if (bq.pair && tile >= bq.ntile) return;
	in += kq * bq.xs;
	out += kq * bq.xs2;
	extern __shared__ char smem_raw[];
	const int N = PL::n(p), TPL = PL::tpl(p);
	cx<T> *tws = reinterpret_cast<cx<T> *>(smem_raw), *buf = tws + N;
	PL::template load_tw<T>(tws, twg, N);
	__builtin_assume(TC >= 2);  // column tiles (element k of column l at buf[k*TC + l]); the launcher guarantees TC >= 2
	const long q0 = (long)tile * TC;
	const int line = threadIdx.x % TC, t = threadIdx.x / TC;
	const bool act = t < TPL && q0 + line < Q;
	const cx<T> *ip = in + q0 + line;
	cx<T> *op = out + q0 + line;
	// The forward transform reads rows k < boxZ <= N/2, the inverse writes rows k < boxZ: the other half of the rows is
	// zero / not needed. With boxZ == N/2 (the usual grid) no per-row check remains.
	if ((INV ? PL::HALF_OUT : PL::HALF_IN) && (INV ? 2 * n_out == N : 2 * n_in == N)) {
		auto ld = [&](int, int k) -> cx<T> { return ip[(size_t)k * Q]; };
		auto st = [&](int, int k, cx<T> v) { op[(size_t)k * Q] = v; };
		PL::template fft<T, INV, !INV, INV, true, true>(buf, tws, p, TC, 1, line, t, act, line, t, act, ld, st);
	}
	else {
		auto ld = [&](int, int k) -> cx<T> { return k < n_in ? ip[(size_t)k * Q] : mk<T>(0, 0); };
		auto st = [&](int, int k, cx<T> v) {
			if (k < n_out) op[(size_t)k * Q] = v;
		};
		PL::template fft<T, INV, !INV, INV, true, true>(buf, tws, p, TC, 1, line, t, act, line, t, act, ld, st);
	}
}

// Sign flip of both parts of a complex number by an integer mask (0 or the sign bit): one XOR per part.
template <typename T> __device__ __forceinline__ cx<T> flip(cx<T> a, unsigned m);
template <> __device__ __forceinline__ cx<float> flip<float>(cx<float> a, unsigned m)
{
	return mk<float>(__uint_as_float(__float_as_uint(a.x) ^ m), __uint_as_float(__float_as_uint(a.y) ^ m));
}
template <> __device__ __forceinline__ cx<double> flip<double>(cx<double> a, unsigned m)
{
	const unsigned long long M = (unsigned long long)m << 32;
	return mk<double>(__longlong_as_double(__double_as_longlong(a.x) ^ M),
		__longlong_as_double(__double_as_longlong(a.y) ^ M));
}

// acc + s*(d*a) for a compile-time sign s (the compiler folds the sign into the FMA operands)
template <typename T, int S> __device__ __forceinline__ cx<T> cmac(cx<T> acc, cx<T> d, cx<T> a)
{
	return S > 0 ? mk<T>(acc.x + d.x * a.x - d.y * a.y, acc.y + d.x * a.y + d.y * a.x)
	             : mk<T>(acc.x - d.x * a.x + d.y * a.y, acc.y - d.x * a.y - d.y * a.x);
}

// Symmetric 3x3 product (cSymMatrVec, src/cmplx.h:369-376) at one frequency, in place in LDS (the x, y, z components
// at q0, q1, q2). D = (d0 f1 f2; f1 d3 f4; f2 f4 d5) with f1, f4 already flipped for the y mirror. The mirror planes
// of the stored octant flip the elements that are odd in the mirrored coordinate (src/matvec.c:270-305 for y and z;
// x alike, G being odd in x for xy and xz): SX = -1 flips xy and xz (column gridX - x), SZ = -1 flips xz and yz
// (plane gridZ - z). The compile-time signs fold into the FMA operands.
template <typename T, int SX, int SZ>
__device__ __forceinline__ void sym3_lds(cx<T> *q0, cx<T> *q1, cx<T> *q2, cx<T> d0, cx<T> f1, cx<T> d2, cx<T> d3,
	cx<T> f4, cx<T> d5)
{
	const cx<T> a0 = *q0, a1 = *q1, a2 = *q2;
	*q0 = cmac<T, SX * SZ>(cmac<T, SX>(mul(d0, a0), f1, a1), d2, a2);
	*q1 = cmac<T, SZ>(cmac<T, SX>(mul(d3, a1), f1, a0), f4, a2);
	*q2 = cmac<T, SX * SZ>(cmac<T, SZ>(mul(d5, a2), f4, a1), d2, a0);
}

// P3. With the reduced D-matrix the stored part is an octant: z < DsizeZ, y < DsizeY (IndexDmatrix_mv,
// src/matvec.c:87-103) and also x < DsizeX = gridX/2 + 1, since G is even or odd in x like in y and z (ADDA keeps x
// whole for its x-slab distribution). A block takes LP consecutive column units u at one plane z <= gridZ/2. With
// column pairs (xpair) unit u is column x = u and its mirror gridX - u, at plane z and at the mirror plane gridZ - z --
// up to four (plane, column) groups that share every stored D element, so each is read from HBM once per matvec.
// Without them (a block too large for LDS otherwise) unit u is column u alone, which reads the stored column
// gridX - u when u >= DsizeX. Group g = s*LP + pr (bit 0 of s: plane gridZ - z, bit 1: column gridX - u) holds lines
// 3g + c; bit g of `pm` says whether the group exists (not for self-mirrored planes or columns, nor beyond the last
// unit). The RHS index is fastest in the block index: the K
// blocks that share a D tile run together and the tile is served from cache after its first read. Lines are
// contiguous boxY runs of Xb (stride 3, components interleaved). Forward FFT-y into LDS (natural order), the D-matrix
// product (src/matvec.c:270-305: mirror indices and sign flips, cSymMatrVec src/cmplx.h:369-376), inverse FFT-y,
// keep y < boxY. The box occupies at most half of every grid line (gridY >= 2*boxY), so with a compile-time plan the
// forward transform skips the zero upper half and the inverse computes only the lower half; FULL (boxY == gridY/2,
// a plan with both half options) drops the y < boxY checks, which then hold for every index the transforms touch.
// Without the reduced D-matrix every group is one column at one plane (s = 0).
// Reflected product (cReflMatrVec, src/cmplx.h:391-401: M21 = M12, M31 = -M13, M32 = -M23) accumulated into o0..o2,
// with the compile-time sign SX of a mirrored column (flips xy and xz, as for D).
template <typename T, int SX>
__device__ __forceinline__ void refl3_acc(cx<T> &o0, cx<T> &o1, cx<T> &o2, cx<T> r0, cx<T> r1, cx<T> r2, cx<T> r3,
	cx<T> r4, cx<T> r5, cx<T> b0, cx<T> b1, cx<T> b2)
{
	o0 = cmac<T, SX>(cmac<T, SX>(cmac<T, 1>(o0, r0, b0), r1, b1), r2, b2);
	o1 = cmac<T, 1>(cmac<T, 1>(cmac<T, SX>(o1, r1, b0), r3, b1), r4, b2);
	o2 = cmac<T, 1>(cmac<T, -1>(cmac<T, -SX>(o2, r2, b0), r4, b1), r5, b2);
}

// Symmetric product of sym3_lds returning its result (the inputs stay unchanged)
template <typename T, int SX, int SZ>
__device__ __forceinline__ void sym3_val(cx<T> &o0, cx<T> &o1, cx<T> &o2, cx<T> d0, cx<T> f1, cx<T> d2, cx<T> d3,
	cx<T> f4, cx<T> d5, cx<T> a0, cx<T> a1, cx<T> a2)
{
	o0 = cmac<T, SX * SZ>(cmac<T, SX>(mul(d0, a0), f1, a1), d2, a2);
	o1 = cmac<T, SZ>(cmac<T, SX>(mul(d3, a1), f1, a0), f4, a2);
	o2 = cmac<T, SX * SZ>(cmac<T, SZ>(mul(d5, a2), f4, a1), d2, a0);
}

// D-matrix and reflected products of one column side at frequency y (SURF): lines A (plane z) and, when the block
// holds it (hz), B (plane gZ - z). The reflected correlation along z is a product with the z-backward transform of the
// source (src/matvec.c:249,288-302); that transform at plane kz is the forward one at plane -kz, so plane z takes
// R(z).X(gZ - z) and plane gZ - z takes R(gZ - z).X(z) -- both lines are in this block. Self-mirrored planes (z = 0,
// 2z = gZ) take R(z).X(z).
template <typename T, int SX>
__device__ __forceinline__ void surf_side(cx<T> *qa0, cx<T> *qa1, cx<T> *qa2, cx<T> *qb0, cx<T> *qb1, cx<T> *qb2,
	bool hz, cx<T> d0, cx<T> f1, cx<T> d2, cx<T> d3, cx<T> f4, cx<T> d5, const cx<T> *ra, const cx<T> *rb,
	unsigned my, unsigned mx)
{
	const cx<T> a0 = *qa0, a1 = *qa1, a2 = *qa2;
	const cx<T> b0 = hz ? *qb0 : a0, b1 = hz ? *qb1 : a1, b2 = hz ? *qb2 : a2;
	cx<T> o0, o1, o2;
	sym3_val<T, SX, 1>(o0, o1, o2, d0, f1, d2, d3, f4, d5, a0, a1, a2);
	refl3_acc<T, SX>(o0, o1, o2, ra[0], flip<T>(ra[1], my ^ mx), flip<T>(ra[2], mx), ra[3], flip<T>(ra[4], my), ra[5],
		b0, b1, b2);
	*qa0 = o0;
	*qa1 = o1;
	*qa2 = o2;
	if (hz) {
		sym3_val<T, SX, -1>(o0, o1, o2, d0, f1, d2, d3, f4, d5, b0, b1, b2);
		refl3_acc<T, SX>(o0, o1, o2, rb[0], flip<T>(rb[1], my ^ mx), flip<T>(rb[2], mx), rb[3], flip<T>(rb[4], my),
			rb[5], a0, a1, a2);
		*qb0 = o0;
		*qb1 = o1;
		*qb2 = o2;
	}
}

// SURF: the particle is above a substrate and R (the transformed image-dipole term, [z < gridZ][xt < DsizeX][yt <
// DsizeY][6], even or odd in x and y like D, full in z) is added to the product (requires the reduced D-matrix, whose
// plane pairs put every line's reflected partner into its block).
// SWZ: the lines in unpadded swizzled rows (LaySwz<true>, line stride N; compile-time plans with N % 16 == 0) instead
// of padded ones: N/16 fewer elements per line, at the cost of an address computation per LDS access.
// MULTI: the block runs bq.kl right-hand sides in turn (batched launches); otherwise one (a loop the compiler cannot
// see through keeps loop-invariant addresses live across it, which costs registers and occupancy).
template <typename T, class PL, bool FULL, bool SURF = false, bool SWZ = false, bool MULTI = false>
__device__ __forceinline__ void p3_body(cx<T> *buf, const cx<T> *tws, const FftPlan2 &p, cx<T> *lp0,
	const cx<T> *__restrict__ D, int u0, int z, int gX, int bY, int gZ, int DsX, int DsY, int LP, int NL1,
	unsigned pm, int reduced, int transposed, int line, int t, bool act, const Bq &bq, int k0,
	const cx<T> *__restrict__ R = nullptr)
{
	// padded rows by default: their offsets stay immediates (a swizzle costs an address computation per access, which
	// matters where P3 is close to VALU-bound, in FP32)
	using LAY = std::conditional_t<SWZ, LaySwz<true>, LayPad>;
	const int N = PL::n(p), LS = SWZ ? N : rowstride_lm(N);
	// D-matrix row of element idx = (pr, y) of the block: the stored row index (mirrored for the reduced form and for
	// the transposed product, src/matvec.c:181 and 270-305) and the sign masks of the y mirror (xy and yz elements,
	// low half of m) and of a single column beyond DsizeX (xy and xz, high half); the block's own plane never mirrors
	auto drow = [&](int idx, unsigned &m) -> const cx<T> * {
		const int pr = idx / N, y = idx - pr * N;
		int xi = u0 + pr, yi = y, zi = z;
		if (transposed) {  // A^T with the full D-matrix: D(-r) (never with the reduced form)
			xi = xi > 0 ? gX - xi : 0;
			yi = yi > 0 ? N - yi : 0;
			zi = zi > 0 ? gZ - zi : 0;
		}
		else if (yi >= DsY) yi = N - yi;
		const bool xm = reduced && xi >= DsX;
		if (xm) xi = gX - xi;
		m = ((reduced && y >= DsY) ? 0x8000u : 0u) | (xm ? 0x80000000u : 0u);
		return D + (((size_t)zi * DsX + xi) * DsY + yi) * 6;
	};
	auto dprod = [&](int idx, const cx<T> *d, unsigned m) {
		const int pr = idx / N, y = idx - pr * N;
		if (!((pm >> pr) & 1)) return;  // beyond the last unit
		const unsigned my = m << 16, mx = m & 0x80000000u;
		const cx<T> f1 = flip<T>(d[1], my ^ mx), f4 = flip<T>(d[4], my), d2 = flip<T>(d[2], mx);
		auto q = [&](int l) { return buf + LAY::pos(l, y, 1, LS); };
		const int l0 = 3 * pr;
		if constexpr (SURF) {
			// the R rows of planes z and gZ - z at the stored (column, y) of D's row (drow without the transposed
			// case: SURF runs with the reduced D-matrix only); R has the same row extents (DsX, DsY) per plane
			const int xs = u0 + pr < DsX ? u0 + pr : gX - (u0 + pr), ys = y < DsY ? y : N - y;
			const bool hz = (pm >> (LP + pr)) & 1;
			const cx<T> *ra = R + (((size_t)z * DsX + xs) * DsY + ys) * 6;
			const cx<T> *rb = hz ? R + (((size_t)(gZ - z) * DsX + xs) * DsY + ys) * 6 : ra;
			surf_side<T, 1>(q(l0), q(l0 + 1), q(l0 + 2), q(NL1 + l0), q(NL1 + l0 + 1), q(NL1 + l0 + 2), hz, d[0], f1,
				d2, d[3], f4, d[5], ra, rb, my, mx);
			if ((pm >> (2 * LP + pr)) & 1)
				surf_side<T, -1>(q(2 * NL1 + l0), q(2 * NL1 + l0 + 1), q(2 * NL1 + l0 + 2), q(3 * NL1 + l0),
					q(3 * NL1 + l0 + 1), q(3 * NL1 + l0 + 2), hz, d[0], f1, d2, d[3], f4, d[5], ra, rb, my, mx);
			return;
		}
		sym3_lds<T, 1, 1>(q(l0), q(l0 + 1), q(l0 + 2), d[0], f1, d2, d[3], f4, d[5]);
		if ((pm >> (LP + pr)) & 1)
			sym3_lds<T, 1, -1>(q(NL1 + l0), q(NL1 + l0 + 1), q(NL1 + l0 + 2), d[0], f1, d2, d[3], f4, d[5]);
		if ((pm >> (2 * LP + pr)) & 1)
			sym3_lds<T, -1, 1>(q(2 * NL1 + l0), q(2 * NL1 + l0 + 1), q(2 * NL1 + l0 + 2), d[0], f1, d2, d[3], f4, d[5]);
		if ((pm >> (3 * LP + pr)) & 1)
			sym3_lds<T, -1, -1>(q(3 * NL1 + l0), q(3 * NL1 + l0 + 1), q(3 * NL1 + l0 + 2), d[0], f1, d2, d[3], f4,
				d[5]);
	};
	// The D rows of this thread's first KP elements are loaded before the forward transform, which hides their latency
	// (in FP32 KP covers all of them with the reduced D-matrix: blockDim = 12*LP*N/RMAX threads for LP*N elements; in
	// FP64 one element, all of them with one thread per element, PT=1).
	constexpr int KP = SURF ? 0 : sizeof(T) == 4 ? (PL::RMAXC + 11) / 12 : 1;
	// the block's RHS k0 .. k0 + kl - 1 (those that exist and are not gated off) in turn; the skip test is uniform over
	// the block, so every thread meets the same barriers. The D rows are loaded again for each RHS (from the cache after
	// the first): held across the loop they would stay live through the inverse transform, at the cost of occupancy.
	for (int kk = 0; kk < (MULTI ? bq.kl : 1); kk++) {
		if (k0 + kk >= bq_nk(bq)) continue;
		const int kq = bq_rhs(bq, k0 + kk);
		if (skip_rhs(bq, kq)) continue;
		if (kk > 0) __syncthreads();  // the previous RHS's last LDS reads precede this forward transform's LDS writes
		cx<T> dpre[KP > 0 ? KP : 1][6];
		unsigned mpre[KP > 0 ? KP : 1];
#pragma unroll
		for (int k = 0; k < KP; k++) {
			const int idx = threadIdx.x + k * blockDim.x;
			if (idx < LP * N && ((pm >> (idx / N)) & 1)) {
				const cx<T> *dm = drow(idx, mpre[k]);
#pragma unroll
				for (int e = 0; e < 6; e++) dpre[k][e] = dm[e];
			}
		}
		cx<T> *lp = lp0 + (size_t)kq * bq.xs;
		auto ldg = [&](int, int y) -> cx<T> { return (FULL || y < bY) ? lp[3 * y] : mk<T>(0, 0); };
		auto stl = [&](int l, int y, cx<T> v) { buf[LAY::pos(l, y, 1, LS)] = v; };
		PL::template fft<T, false, true, false, true, false, LAY>(buf, tws, p, 1, LS, line, t, act, line, t, act, ldg, stl);
		__syncthreads();
#pragma unroll
		for (int k = 0; k < KP; k++) {
			const int idx = threadIdx.x + k * blockDim.x;
			if (idx < LP * N) dprod(idx, dpre[k], mpre[k]);
		}
		for (int idx = threadIdx.x + KP * blockDim.x; idx < LP * N; idx += blockDim.x) {
			if (!((pm >> (idx / N)) & 1)) continue;
			unsigned m;
			const cx<T> *dm = drow(idx, m);
			cx<T> d[6];
#pragma unroll
			for (int e = 0; e < 6; e++) d[e] = dm[e];
			dprod(idx, d, m);
		}
		__syncthreads();
		auto ldl = [&](int l, int y) -> cx<T> { return buf[LAY::pos(l, y, 1, LS)]; };
		auto stg = [&](int, int y, cx<T> v) {
			if (FULL || y < bY) lp[3 * y] = v;
		};
		PL::template fft<T, true, false, true, false, true, LAY>(buf, tws, p, 1, LS, line, t, act, line, t, act, ldl, stg);
	}
}

// Column units of P3: x < DsizeX with column pairs (each with its mirror), else every x; and the (plane, column)
// groups of three lines a unit can hold (z, gZ - z with the reduced D-matrix, times x, gX - x with column pairs)
__host__ __device__ __forceinline__ int p3_units(int gX, int DsX, int reduced, int xpair)
{
	return reduced && xpair ? DsX : gX;
}
__host__ __device__ __forceinline__ int p3_groups(int reduced, int xpair) { return reduced ? (xpair ? 4 : 2) : 1; }

// TWG (compile-time plans, kernel k_p3g): twg is the plan's staged twiddle table in global memory (k_stage_tw), read
// through the cache instead of being copied into LDS. That frees N elements of LDS per block, which at 256-point lines
// lets one more block share a CU.
// LPC > 0: the number of column units per block fixed at compile time (LP = 1 is what the autotuner picks at the
// sizes measured), which turns the group offsets and the group tests of the D-matrix product into constants.
template <typename T, class PL, bool TWG, int LPC, bool SURF = false, bool SWZ = false, bool MULTI = false>
__device__ __forceinline__ void p3_kernel(cx<T> *Xb, const cx<T> *__restrict__ D, FftPlan2 p,
	const cx<T> *__restrict__ twg, int gX, int bY, int gZ, int DsX, int DsY, int DsZ, int LPr, int reduced, int xpair,
	int transposed, Bq bq, const cx<T> *__restrict__ R = nullptr)
{
	const int LP = LPC > 0 ? LPC : LPr;
	// blocks of kl consecutive RHS (the group index fastest, so the groups that share a D tile run together)
	// (k0: the first launch slot of the block's group; its RHS through the launch list)
	const int nk = bq_nk(bq), kl = MULTI ? bq.kl : 1, nkg = (nk + kl - 1) / kl, k0 = (blockIdx.x % nkg) * kl;
	const int tile = blockIdx.x / nkg;
	{
		bool any = false;
		for (int kk = 0; kk < kl && k0 + kk < nk; kk++) any |= !skip_rhs(bq, bq_rhs(bq, k0 + kk));
		if (!any) return;
	}
	extern __shared__ char smem_raw[];
	const int N = PL::n(p), TPL = PL::tpl(p), NL1 = 3 * LP;
	const cx<T> *tws;
	cx<T> *buf;
	if constexpr (TWG) {
		tws = twg;
		buf = reinterpret_cast<cx<T> *>(smem_raw);
	}
	else {
		cx<T> *tl = reinterpret_cast<cx<T> *>(smem_raw);
		PL::template load_tw<T>(tl, twg, N);
		tws = tl;
		buf = tl + N;
	}
	const int nu = p3_units(gX, DsX, reduced, xpair), nut = (nu + LP - 1) / LP;
	int z = tile / nut, u0 = (tile % nut) * LP;
	if (bq.pair) {
		// single columns of the reduced D-matrix (LP = 1, even gridX): column x and its mirror gridX - x read the same
		// stored D row, so they go 8 blocks apart -- on one XCD at about the same time -- and the second read is served
		// by that XCD's L2. Pair p of plane z is (p, gridX - p) for 0 < p < gridX/2 and (0, gridX/2) for p = 0; the
		// launch is padded to whole groups of 16 blocks (8 pairs).
		const int hp = gX / 2, pi = (tile >> 4) * 8 + (tile & 7), mem = (tile >> 3) & 1;
		if (pi >= bq.ntile) return;
		z = pi / hp;
		const int pp = pi - z * hp;
		u0 = mem ? (pp ? gX - pp : hp) : pp;
	}
	// blocks run over z <= gridZ/2 (< DsizeZ) with the reduced D-matrix; mirror planes and columns are handled alongside
	const bool hz = reduced && z != 0 && 2 * z != gZ;
	unsigned pm = 0;
	for (int pr = 0; pr < LP && u0 + pr < nu; pr++) {
		const int u = u0 + pr;
		const bool hx = reduced && xpair && u != 0 && 2 * u != gX;
		pm |= (1u | (hz ? 1u << LP : 0u) | (hx ? 1u << (2 * LP) : 0u) | (hz && hx ? 1u << (3 * LP) : 0u)) << pr;
	}
	const int tid = threadIdx.x, line = tid / TPL, t = tid % TPL;
	const int g = line / 3, s = g / LP, pr = g - s * LP;
	const bool act = (pm >> g) & 1;
	// line (s, pr, c): element y of it at lp[3*y], lp = &Xb[plane][column][0][c] of RHS 0 (plus k*xs for RHS k;
	// inactive lines never access it)
	const int zz = (s & 1) ? gZ - z : z, xx = (s & 2) ? gX - (u0 + pr) : u0 + pr;
	cx<T> *lp = act ? Xb + ((size_t)zz * gX + xx) * 3 * bY + line % 3 : Xb;
	if (PL::HALF_IN && PL::HALF_OUT && 2 * bY == N)
		p3_body<T, PL, true, SURF, SWZ, MULTI>(buf, tws, p, lp, D, u0, z, gX, bY, gZ, DsX, DsY, LP, NL1, pm, reduced,
			transposed, line, t, act, bq, k0, R);
	else
		p3_body<T, PL, false, SURF, SWZ, MULTI>(buf, tws, p, lp, D, u0, z, gX, bY, gZ, DsX, DsY, LP, NL1, pm, reduced,
			transposed, line, t, act, bq, k0, R);
}

template <typename T, class PL, int LPC = 0, bool SWZ = false, bool MULTI = false>
__global__ void __launch_bounds__(512) k_p3(cx<T> *Xb, const cx<T> *__restrict__ D, FftPlan2 p,
	const cx<T> *__restrict__ twg, int gX, int bY, int gZ, int DsX, int DsY, int DsZ, int LP, int reduced, int xpair,
	int transposed, Bq bq)
{
	p3_kernel<T, PL, false, LPC, false, SWZ, MULTI>(Xb, D, p, twg, gX, bY, gZ, DsX, DsY, DsZ, LP, reduced, xpair,
		transposed, bq);
}
// P3 for a particle above a substrate: the D-matrix product plus the reflected term R (twiddles in LDS, any LP)
template <typename T, class PL>
__global__ void __launch_bounds__(512) k_p3sub(cx<T> *Xb, const cx<T> *__restrict__ D, const cx<T> *__restrict__ R,
	FftPlan2 p, const cx<T> *__restrict__ twg, int gX, int bY, int gZ, int DsX, int DsY, int DsZ, int LP, int reduced,
	int xpair, int transposed, Bq bq)
{
	p3_kernel<T, PL, false, 0, true>(Xb, D, p, twg, gX, bY, gZ, DsX, DsY, DsZ, LP, reduced, xpair, transposed, bq, R);
}
// Register budget of five waves per SIMD: the twiddle loads from global memory would otherwise be hoisted into
// registers, costing the occupancy that the LDS saving buys.
template <typename T, class PL, int LPC = 0>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5))) k_p3g(cx<T> *Xb,
	const cx<T> *__restrict__ D, FftPlan2 p, const cx<T> *__restrict__ twg, int gX, int bY, int gZ, int DsX, int DsY,
	int DsZ, int LP, int reduced, int xpair, int transposed, Bq bq)
{
	p3_kernel<T, PL, true, LPC>(Xb, D, p, twg, gX, bY, gZ, DsX, DsY, DsZ, LP, reduced, xpair, transposed, bq);
}

// D-matrix product at frequency y of one grid line of plane zz, for a block at plane z (src/matvec.c:270-305:
// reduced-FFT mirror indices and sign flips; cSymMatrVec src/cmplx.h:369-376), applied to the three components a0..a2
// in place. Dl: the line's stored D row [yt < DsY][6] (already mirrored in x and z, see k_p3r); xm: the line's column
// is the mirror of its stored one (flips xy and xz).
template <typename T>
__device__ __forceinline__ void dmul3(const cx<T> *Dl, int y, int zz, int N, int DsY, int DsZ, int reduced,
	int transposed, bool xm, cx<T> &a0, cx<T> &a1, cx<T> &a2)
{
	int yi = y;
	if (transposed) yi = yi > 0 ? N - yi : 0;
	else if (yi >= DsY) yi = N - yi;
	const cx<T> *dm = Dl + yi * 6;
	const cx<T> d0 = dm[0], d3 = dm[3], d5 = dm[5];
	cx<T> f1 = dm[1], f2 = dm[2], f4 = dm[4];
	if (reduced) {
		if (y >= DsY) {
			f1 = scl(f1, (T)-1);
			if (zz >= DsZ) f2 = scl(f2, (T)-1);
			else f4 = scl(f4, (T)-1);
		}
		else if (zz >= DsZ) {
			f2 = scl(f2, (T)-1);
			f4 = scl(f4, (T)-1);
		}
		if (xm) {
			f1 = scl(f1, (T)-1);
			f2 = scl(f2, (T)-1);
		}
	}
	const cx<T> b0 = a0, b1 = a1, b2 = a2;
	a0 = add(add(mul(d0, b0), mul(f1, b1)), mul(f2, b2));
	a1 = add(add(mul(f1, b0), mul(d3, b1)), mul(f4, b2));
	a2 = add(add(mul(f2, b0), mul(f4, b1)), mul(d5, b2));
}

// P3 with the D-matrix product in registers (compile-time power-of-two plans): thread t of group g = (plane, pair)
// carries the three components of its butterflies through the forward FFT-y; the last stage leaves frequencies
// t + u*TPL + q*NS_LAST in registers, where the product with D is applied, and the inverse FFT-y of the reversed
// radix order starts from exactly those registers -- no LDS round trip and no separate pass over the spectrum.
// Blocks take LP consecutive columns x at one plane z (and at the mirror plane z' = gZ - z, sharing the D tile);
// groups g < LP are plane z, the others z'. Columns x >= DsizeX read the stored row of gridX - x.
template <typename T, class PL>
__global__ void __launch_bounds__(256) k_p3r(cx<T> *Xb, const cx<T> *__restrict__ D, FftPlan2 p,
	const cx<T> *__restrict__ twg, int gX, int bY, int gZ, int DsX, int DsY, int DsZ, int LP, int reduced, int,
	int transposed, Bq bq)
{
	constexpr int N = PL::N, TPL = PL::TPL, RMAX = PL::RMAX, RL = PL::RAD[PL::NST - 1], NSL = PL::NS_LAST;
	const int nk = bq_nk(bq), kq = bq_rhs(bq, blockIdx.x % nk), tile = blockIdx.x / nk;
	if (skip_rhs(bq, kq)) return;
	Xb += kq * bq.xs;
	extern __shared__ char smem_raw[];
	const int LS = rowstride_lm(N);
	cx<T> *tws = reinterpret_cast<cx<T> *>(smem_raw), *buf = tws + N, *Dt = buf + 6 * LP * LS;
	const int nxt = gX / LP;
	const int z = tile / nxt, x0 = (tile % nxt) * LP;
	// the block's D rows [pr][yt][6] into LDS with coalesced loads (each row is contiguous in D)
	{
		int zi = z;
		if (transposed) zi = zi > 0 ? gZ - zi : 0;
		else if (zi >= DsZ) zi = gZ - zi;
		const int rl = DsY * 6;
		for (int i = threadIdx.x; i < LP * rl; i += blockDim.x) {
			const int pr = i / rl, x = x0 + pr;
			const int xi = transposed ? (x > 0 ? gX - x : 0) : (x < DsX ? x : gX - x);
			Dt[i] = D[((size_t)zi * DsX + xi) * rl + i % rl];
		}
	}
	load_tw(tws, twg, N);  // ends with a barrier, which also publishes the D rows
	const bool has2 = reduced && z != 0 && 2 * z != gZ;
	const int zp = gZ - z;
	const int tid = threadIdx.x, g = tid / TPL, t = tid % TPL;
	const int s = g / LP, pr = g % LP, zz = s ? zp : z;
	const bool act = g < (has2 ? 2 : 1) * LP;
	cx<T> *base = Xb + ((size_t)zz * gX + x0 + pr) * 3 * bY;  // element (y, c) at base[3y + c]
	cx<T> e[3][RMAX];
	auto ld = [&](int c, int y) -> cx<T> { return y < bY ? base[3 * y + c] : mk<T>(0, 0); };
	PL::template fwd3<T, 0, 1>(e, buf, tws, LS, 3 * g, t, act, ld);
	if (act) {
		const cx<T> *Dl = Dt + pr * DsY * 6;
		const bool xm = reduced && x0 + pr >= DsX;
#pragma unroll
		for (int u = 0; u < RMAX / RL; u++)
#pragma unroll
			for (int q = 0; q < RL; q++) {
				const int i = u * RL + q;
				dmul3<T>(Dl, t + u * TPL + q * NSL, zz, N, DsY, DsZ, reduced, transposed, xm, e[0][i], e[1][i],
					e[2][i]);
				__builtin_amdgcn_sched_barrier(0);  // one frequency at a time: bounds the registers held by D
			}
	}
	auto st = [&](int c, int y, cx<T> v) {
		if (y < bY) base[3 * y + c] = v;
	};
	PL::template inv3<T, 0, 1>(e, buf, tws, LS, 3 * g, t, act, st);
}

// P5: block = (z, y0..y0+LY, RHS), all three components; stage 0 loads Xa[z][x][y][c] with (y, c) fastest (runs of
// 3*LY elements), the inverse FFT lands in the LDS lines yl*3 + c, and each thread then gathers one dipole:
// result = arg + S.X (her: conj), three contiguous components. Per block, at partials[(k*rows + row)*nb + tile] (nb
// blocks per RHS): row 0 ||result||^2; with dmode 1 (w = dotw) or 2 (w = arg) also rows 1, 2 = Re, Im of
// sum result.conj(w) -- an inner product of the Krylov iteration fused into the pass that writes its operand.
// TV: the type of the dipole vectors (arg, res, dotw); the transform and the coupling constants are in T. With
// TV = double, T = float the correction S.X is rounded to FP32 and added to the FP64 argument in FP64.
template <typename T, class PL, bool DOT, typename TV = T>
__global__ void __launch_bounds__(512) k_p5(const cx<T> *__restrict__ Xa, const cx<TV> *__restrict__ arg,
	cx<TV> *__restrict__ res, const int *__restrict__ cellmap, const RowDesc *__restrict__, const int *__restrict__ mat,
	const cx<T> *__restrict__ ccs, FftPlan2 p, const cx<T> *__restrict__ twg, int bX, int bY, int LY, int her,
	double *__restrict__ partials, const cx<TV> *__restrict__ dotw, int dmode, Bq bq)
{
	const int nk = bq_nk(bq), kq = bq_rhs(bq, blockIdx.x % nk);
	const int tile = bq.pair ? pair_tile(blockIdx.x / nk) : blockIdx.x / nk;
	if (skip_rhs(bq, kq)) return;
	if (bq.pair && tile >= bq.ntile) {  // padding of the paired launch: zero partial sums, nothing else
		if (partials && threadIdx.x == 0) {
			const int nr = DOT ? 3 : 1;
			const size_t nb = gridDim.x / nk;
			for (int r = 0; r < nr; r++) partials[(size_t)kq * nr * nb + r * nb + tile] = 0;
		}
		return;
	}
	Xa += kq * bq.xs;
	arg += kq * bq.vs;
	res += kq * bq.vs;
	if (dotw) dotw += kq * bq.vs;
	ccs += (size_t)kq * bq.cs;
	extern __shared__ char smem_raw[];
	using LAY = std::conditional_t<PL::SWZ, LaySwz<false>, LayPad>;  // lines are also walked line-fastest
	const int N = PL::n(p), TPL = PL::tpl(p), LS = PL::SWZ ? N + 1 : rowstride_odd(N), NL = 3 * LY;
	cx<T> *tws = reinterpret_cast<cx<T> *>(smem_raw), *buf = tws + N;
	double *red_lds = reinterpret_cast<double *>(buf);  // the lines' space, reused after the gather (>= 48*LY doubles)
	PL::template load_tw<T>(tws, twg, N);
	const int nyt = (bY + LY - 1) / LY;
	const int z = tile / nyt, y0 = (tile % nyt) * LY;
	const int ny = bY - y0 < LY ? bY - y0 : LY;
	const int tid = threadIdx.x;
	const int lineA = tid % NL, tA = tid / NL;    // (y, c) fastest (reads Xa)
	const int lineB = tid / TPL, tB = tid % TPL;  // x fastest (writes the LDS lines)
	const bool actA = tA < TPL && lineA < 3 * ny, actB = lineB < 3 * ny;
	double s = 0, dr = 0, di = 0;
	// x < boxX <= N/2: with a half-output plan the rows are walked at the constant stride N/2 (no division)
	const int NF = PL::HALF_OUT ? N / 2 : bX, total = LY * NF;
	// Elements per thread and round: with a compile-time plan one round covers the block (LY*NF <= 3*LY*TPL*RMAX/6).
	constexpr int GU = PL::RMAXC > 0 ? (PL::RMAXC + 5) / 6 : 4;
	const int *cm = cellmap + ((size_t)z * bY + y0) * bX;
	const bool onemat = bq.cs == 3;  // one material: the coupling constants are the same for every dipole
	// The fused-dot variant loads the argument of its dipoles before the transform as well (it waits on more loads
	// after it); the plain pass loads it afterwards, which keeps its registers for one more wave per SIMD (measured:
	// plain FP64 P5 86 -> 78 us, FP32 48 -> 46 us, while the FP32 fused-dot pass would slow from 59 to 75 us).
	constexpr bool PRE = DOT;
	struct Round {
		int dd[GU], mm[GU];
		cx<TV> av[PRE ? GU : 1][3];
	};
	// Cell map, material and argument of a round are independent of the transform: the first round's loads are issued
	// before it, which hides their latency. The coupling constants (a table of a few entries) and the dot weights are
	// read after it, so that they do not hold registers through the transform.
	auto fetch = [&](int b0, Round &g) {
#pragma unroll
		for (int u = 0; u < GU; u++) {
			const int idx = b0 + u * blockDim.x, yl = idx / NF, x = idx - yl * NF;
			g.dd[u] = (idx < total && x < bX && yl < ny) ? cm[yl * bX + x] : -1;
		}
#pragma unroll
		for (int u = 0; u < GU; u++) {
			g.mm[u] = 0;
			if (g.dd[u] >= 0) {
				if (!onemat) g.mm[u] = 3 * mat[g.dd[u]];
				if constexpr (PRE)
#pragma unroll
					for (int c = 0; c < 3; c++) g.av[u][c] = arg[3 * (size_t)g.dd[u] + c];
			}
		}
	};
	auto finish = [&](int b0, const Round &g) {
		cx<T> cs[GU][3];
		cx<TV> wv[GU][3], al[PRE ? 1 : GU][3];
		auto av = [&](int u, int c) -> cx<TV> {
			if constexpr (PRE) return g.av[u][c];
			else return al[u][c];
		};
#pragma unroll
		for (int u = 0; u < GU; u++)
			if (g.dd[u] >= 0)
#pragma unroll
				for (int c = 0; c < 3; c++) {
					if constexpr (!PRE) al[u][c] = arg[3 * (size_t)g.dd[u] + c];
					cs[u][c] = ccs[g.mm[u] + c];
					if (DOT) wv[u][c] = dmode == 1 ? dotw[3 * (size_t)g.dd[u] + c] : av(u, c);
				}
#pragma unroll
		for (int u = 0; u < GU; u++)
			if (g.dd[u] >= 0) {
				const int idx = b0 + u * blockDim.x, yl = idx / NF, x = idx - yl * NF;
				cx<TV> *r = res + 3 * (size_t)g.dd[u];
#pragma unroll
				for (int c = 0; c < 3; c++) {
					const cx<TV> a = her ? conjc(av(u, c)) : av(u, c);
					cx<TV> rv = add(a, cvt<TV>(mul(cs[u][c], buf[LAY::pos(3 * yl + c, x, 1, LS)])));
					if (her) rv = conjc(rv);
					r[c] = rv;
					s += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
					if (DOT) {
						const cx<TV> w = wv[u][c];
						dr += (double)rv.x * w.x + (double)rv.y * w.y;
						di += (double)rv.y * w.x - (double)rv.x * w.y;
					}
				}
			}
	};
	Round g0;
	fetch(tid, g0);
	const cx<T> *xi = Xa + (size_t)z * N * 3 * bY + 3 * y0 + lineA;
	const size_t xs = (size_t)3 * bY;
	auto ld = [&](int, int x) -> cx<T> { return xi[x * xs]; };
	auto st = [&](int line, int x, cx<T> v) { buf[LAY::pos(line, x, 1, LS)] = v; };
	// only x < boxX <= N/2 is gathered: the inverse transform computes the lower half of each line
	PL::template fft<T, true, false, true, true, false, LAY>(buf, tws, p, 1, LS, lineA, tA, actA, lineB, tB, actB, ld,
		st);
	__syncthreads();
	finish(tid, g0);
	for (int b0 = tid + GU * blockDim.x; b0 < total; b0 += GU * blockDim.x) {  // run-time plans only
		Round g;
		fetch(b0, g);
		finish(b0, g);
	}
	if (partials) {
		const int rows = DOT ? 3 : 1;
		const size_t nb = gridDim.x / nk, o = (size_t)kq * rows * nb + tile;
		if (DOT) {
			block_sum3<T>(s, dr, di, red_lds);
			if (threadIdx.x == 0) {
				partials[o] = s;
				partials[o + nb] = dr;
				partials[o + 2 * nb] = di;
			}
		}
		else {
			const double tt = block_sum<T>(s, red_lds);
			if (threadIdx.x == 0) partials[o] = tt;
		}
	}
}

// P5 for compile-time plans, with the gather fused into the last stage of the inverse FFT-x: each output x < boxX of
// line (yl, c) goes straight to result[3d + c] = arg[3d + c] + cc_sqrt*X for the dipole d at that cell (her: conj),
// without passing through LDS. The lanes of the three component lines of a row write the 48 (24) contiguous bytes of
// each dipole in one store instruction, so every written cache line is complete. Dense rows (RowDesc) give the dipole
// index arithmetically. The partial sums are those of k_p5 (rows 0..2 at partials[(k*rows + row)*nb + tile]),
// accumulated in another order. The last stage leaves its outputs in registers; then the dipole indices, materials,
// arguments, coupling constants (and dot weights) of all of them are found in phases whose loads are in flight together
// -- one round trip per thread for dense rows of one material -- and nothing is held in registers through the
// transform. (Measured: loading the arguments before the transform, beside its own loads, took FP32 P5 from 44 to
// 47 us -- 99 registers, or a spill under a budget of five waves per SIMD.)
// NT: the Xa loads with the nontemporal hint (a kernel of its own, as in P1; XN bit 1).
template <typename T, class PL, int DM, typename TV = T, bool NT = false>
__global__ void __launch_bounds__(512) k_p5s(const cx<T> *__restrict__ Xa, const cx<TV> *__restrict__ arg,
	cx<TV> *__restrict__ res, const int *__restrict__ cellmap, const RowDesc *__restrict__ rows,
	const int *__restrict__ mat, const cx<T> *__restrict__ ccs, FftPlan2 p, const cx<T> *__restrict__ twg, int bX,
	int bY, int LY, int her, double *__restrict__ partials, const cx<TV> *__restrict__ dotw, int, Bq bq)
{
	constexpr bool DOT = DM != 0;
	const int nk = bq_nk(bq), kq = bq_rhs(bq, blockIdx.x % nk);
	const int tile = bq.pair ? pair_tile(blockIdx.x / nk) : blockIdx.x / nk;
	if (skip_rhs(bq, kq)) return;
	if (bq.pair && tile >= bq.ntile) {  // padding of the paired launch: zero partial sums, nothing else
		if (partials && threadIdx.x == 0) {
			const int nr = DOT ? 3 : 1;
			const size_t nb = gridDim.x / nk;
			for (int r = 0; r < nr; r++) partials[(size_t)kq * nr * nb + r * nb + tile] = 0;
		}
		return;
	}
	Xa += kq * bq.xs;
	arg += kq * bq.vs;
	res += kq * bq.vs;
	if (dotw) dotw += kq * bq.vs;
	ccs += (size_t)kq * bq.cs;
	extern __shared__ char smem_raw[];
	using LAY = std::conditional_t<PL::SWZ, LaySwz<false>, LayPad>;  // lines are also walked line-fastest
	constexpr int N = PL::N, TPL = PL::TPL;
	const int LS = PL::SWZ ? N + 1 : rowstride_odd(N), NL = 3 * LY;
	cx<T> *tws = reinterpret_cast<cx<T> *>(smem_raw), *buf = tws + N;
	double *red_lds = reinterpret_cast<double *>(buf);  // the lines' space, reused after the transform (>= 48*LY doubles)
	PL::template load_tw<T>(tws, twg, N);
	const int nyt = (bY + LY - 1) / LY;
	const int z = tile / nyt, y0 = (tile % nyt) * LY;
	const int ny = bY - y0 < LY ? bY - y0 : LY;
	const int tid = threadIdx.x;
	const int lineA = tid % NL, tA = tid / NL;    // (y, c) fastest (reads Xa)
	const int lineB = tid / TPL, tB = tid % TPL;  // x fastest (the gather of the last stage)
	const bool actA = tA < TPL && lineA < 3 * ny, actB = lineB < 3 * ny;
	const int yl = lineB / 3, c = lineB - 3 * yl;
	const size_t row = (size_t)z * bY + y0 + yl;
	RowDesc rd = {0, 0, 0, 1};
	if (actB) rd = rows[row];
	const int *cm = cellmap + row * bX;
	const bool onemat = bq.cs == 3;
	constexpr int NS = PL::NSLOT;
	// phases whose loads are in flight together (loaded one output at a time, each would wait on the last); empty cells
	// load dipole 0 and discard it (a conditional load would get a basic block and a wait of its own)
	// Indices are 32-bit (a dipole vector holds fewer than 2^32 elements). ai: element index 3d + c of the argument and
	// result (dipole 0 for an empty cell); on a dense row it is a0 + 3x, linear in the cell, with no product per element.
	int dd[NS];
	unsigned ai[NS];
	cx<TV> av[NS];
	auto find_dipoles = [&] {
		if (rd.dense) {
			const unsigned a0 = 3u * (unsigned)(rd.s - rd.x0) + (unsigned)c;  // modulo 2^32, exact where used
			PL::template for_last<true>(tB, [&](int sl, int x) {
				const unsigned o = (unsigned)(x - rd.x0);
				const bool in = x < bX && o < (unsigned)rd.n;
				dd[sl] = in ? rd.s + (int)o : -1;
				ai[sl] = in ? a0 + 3u * (unsigned)x : (unsigned)c;
			});
		}
		else {
			PL::template for_last<true>(tB, [&](int sl, int x) { dd[sl] = x < bX ? cm[x] : -1; });
			PL::template for_last<true>(tB, [&](int sl, int) { ai[sl] = 3u * (unsigned)max(dd[sl], 0) + c; });
		}
	};
	auto load_args = [&] { PL::template for_last<true>(tB, [&](int sl, int) { av[sl] = arg[ai[sl]]; }); };
	double s = 0, dr = 0, di = 0;
	// the tile's base is the block's; a thread's elements lie within one z-plane of Xa (byte offsets far below 2^32)
	const cx<T> *xi = Xa + (size_t)z * N * 3 * bY + 3 * y0;
	const unsigned xsb = 3u * bY * (unsigned)sizeof(cx<T>), la = (unsigned)lineA * (unsigned)sizeof(cx<T>);
	auto ld = [&](int, int x) -> cx<T> {
		if constexpr (NT) return ld_nt(&at_b(xi, la + (unsigned)x * xsb));
		else return at_b(xi, la + (unsigned)x * xsb);
	};
	cx<T> ov[NS];
	auto st = [&](int, int, cx<T> v, int sl) { ov[sl] = v; };
	// only x < boxX <= N/2 is gathered: the inverse transform computes the lower half of each line
	PL::template fft<T, true, false, true, true, true, LAY>(buf, tws, p, 1, LS, lineA, tA, actA, lineB, tB, actB, ld,
		st);
	if (actB) {
		int mm[NS];
		cx<TV> wv[NS];
		cx<T> cv[NS];
		find_dipoles();
		if (!onemat) PL::template for_last<true>(tB, [&](int sl, int) { mm[sl] = 3 * mat[(unsigned)max(dd[sl], 0)]; });
		else PL::template for_last<true>(tB, [&](int sl, int) { mm[sl] = 0; });
		load_args();
		PL::template for_last<true>(tB, [&](int sl, int) {
			cv[sl] = ccs[(unsigned)(mm[sl] + c)];
			// the weights: another vector (DM 1), or the argument itself (DM 2), already in registers
			if constexpr (DM == 1) wv[sl] = dotw[ai[sl]];
			else if constexpr (DM == 2) wv[sl] = av[sl];
		});
		PL::template for_last<true>(tB, [&](int sl, int) {
			const int d = dd[sl];
			if (d < 0) return;
			const cx<TV> a = av[sl];
			cx<TV> rv = add(her ? conjc(a) : a, cvt<TV>(mul(cv[sl], ov[sl])));
			if (her) rv = conjc(rv);
			res[ai[sl]] = rv;
			s += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
			if (DOT) {
				const cx<TV> w = wv[sl];
				dr += (double)rv.x * w.x + (double)rv.y * w.y;
				di += (double)rv.y * w.x - (double)rv.x * w.y;
			}
		});
	}
	if (partials) {  // block_sum starts with a barrier: the transform's last LDS reads are done before red_lds is reused
		const int nr = DOT ? 3 : 1;
		const size_t nb = gridDim.x / nk, o = (size_t)kq * nr * nb + tile;
		if (DOT) {
			block_sum3<T>(s, dr, di, red_lds);
			if (threadIdx.x == 0) {
				partials[o] = s;
				partials[o + nb] = dr;
				partials[o + 2 * nb] = di;
			}
		}
		else {
			const double tt = block_sum<T>(s, red_lds);
			if (threadIdx.x == 0) partials[o] = tt;
		}
	}
}

// The staged twiddle table of a compile-time plan (Fix::load_tw's layout) written to global memory, for k_p3<..., true>.
// One block.
template <typename T, class PL>
__global__ void __launch_bounds__(256) k_stage_tw(cx<T> *out, const cx<T> *__restrict__ twg)
{
	PL::template load_tw<T>(out, twg, PL::N);
}

// The register D-product variant exists for compile-time power-of-two plans whose three-component register file
// stays within 48 complex values per thread (FP32 up to radix 16, FP64 up to radix 8).
template <class PL> struct is_fix : std::false_type {};
template <int... Rs> struct is_fix<Fix<Rs...>> : std::true_type {};
template <typename T, class PL> constexpr bool p3r_ok()
{
	if constexpr (is_fix<PL>::value) return PL::POW2 && 3 * PL::RMAX * sizeof(cx<T>) <= 48 * 8;
	else return false;
}

// The kernels of one plan: k[kind][0] FP64, k[kind][1] FP32 (entry addresses of the instantiations).
// MV_P1W / MV_P5W / MV_P5DW: FP64 dipole vectors around the FP32 transform (FP32 slot only).
// MV_P5D / MV_P5DW: P5 with the inner product against dotw (dmode 1); MV_P5D2 / MV_P5DW2 against the argument
// (dmode 2) -- the same kernel for run-time plans, which take dmode as an argument.
// MV_P3G: P3 with its twiddles read from global memory; MV_TWST: the kernel that stages that table (compile-time
// plans only).
// MV_P31 / MV_P3G1: the same two P3 kernels compiled for one column unit per block (LP = 1).
// MV_P3S: P3 with the substrate's reflected term (k_p3sub).
// MV_P3W1: P3 with one column unit per block and the lines in unpadded swizzled rows (compile-time plans, N % 16 == 0).
// MV_P3M: P3 whose blocks run several right-hand sides in turn (batched launches; padded rows, twiddles in LDS).
enum {
	MV_P1 = 0, MV_ZFWD = 1, MV_ZINV = 2, MV_P3 = 3, MV_P5 = 4, MV_P3R = 5, MV_P5D = 6, MV_P1W = 7, MV_P5W = 8,
	MV_P5DW = 9, MV_P3G = 10, MV_TWST = 11, MV_P31 = 12, MV_P3G1 = 13, MV_P3S = 14, MV_P3W1 = 15, MV_P3M = 16,
	MV_P5D2 = 17, MV_P5DW2 = 18, MV_KINDS = 19
};
struct MvKernels {
	const void *k[MV_KINDS][2];
	// the P1 / P5 kinds with the nontemporal hint on Xa (compile-time plans; null otherwise): nt[kind][fp32]
	const void *nt[MV_KINDS][2] = {};
	bool swz = false;  // the plan's P1/P5 kernels use the swizzled LDS rows (line stride N + 1)
};

template <typename T, class PL> void mv_fill_t(MvKernels &m)
{
	const int f = sizeof(T) == 4;
	// compile-time plans: P1/P5 with the gather fused into the first / last FFT stage
	if constexpr (is_fix<PL>::value) {
		m.k[MV_P1][f] = (const void *)&k_p1s<T, PL>;
		m.k[MV_P5][f] = (const void *)&k_p5s<T, PL, 0>;
		m.k[MV_P5D][f] = (const void *)&k_p5s<T, PL, 1>;
		m.k[MV_P5D2][f] = (const void *)&k_p5s<T, PL, 2>;
		m.nt[MV_P1][f] = (const void *)&k_p1s<T, PL, T, true>;
		m.nt[MV_P5][f] = (const void *)&k_p5s<T, PL, 0, T, true>;
		m.nt[MV_P5D][f] = (const void *)&k_p5s<T, PL, 1, T, true>;
		m.nt[MV_P5D2][f] = (const void *)&k_p5s<T, PL, 2, T, true>;
	}
	else {
		m.k[MV_P1][f] = (const void *)&k_p1<T, PL>;
		m.k[MV_P5][f] = (const void *)&k_p5<T, PL, false>;
		m.k[MV_P5D][f] = m.k[MV_P5D2][f] = (const void *)&k_p5<T, PL, true>;
	}
	m.k[MV_ZFWD][f] = (const void *)&k_zfft<T, false, PL>;
	m.k[MV_ZINV][f] = (const void *)&k_zfft<T, true, PL>;
	m.k[MV_P3][f] = (const void *)&k_p3<T, PL>;
	m.k[MV_P31][f] = (const void *)&k_p3<T, PL, 1>;
	m.k[MV_P3M][f] = (const void *)&k_p3<T, PL, 0, false, true>;
	m.k[MV_P3S][f] = (const void *)&k_p3sub<T, PL>;
	m.k[MV_P3R][f] = nullptr;
	if constexpr (p3r_ok<T, PL>()) m.k[MV_P3R][f] = (const void *)&k_p3r<T, PL>;
	m.k[MV_P3G][f] = m.k[MV_TWST][f] = m.k[MV_P3G1][f] = m.k[MV_P3W1][f] = nullptr;
	if constexpr (is_fix<PL>::value) {
		if constexpr (PL::SWZ) m.k[MV_P3W1][f] = (const void *)&k_p3<T, PL, 1, true>;
		m.k[MV_P3G][f] = (const void *)&k_p3g<T, PL>;
		m.k[MV_P3G1][f] = (const void *)&k_p3g<T, PL, 1>;
		m.k[MV_TWST][f] = (const void *)&k_stage_tw<T, PL>;
	}
	m.k[MV_P1W][f] = m.k[MV_P5W][f] = m.k[MV_P5DW][f] = m.k[MV_P5DW2][f] = nullptr;
	if constexpr (sizeof(T) == 4) {
		if constexpr (is_fix<PL>::value) {
			m.k[MV_P1W][f] = (const void *)&k_p1s<T, PL, double>;
			m.k[MV_P5W][f] = (const void *)&k_p5s<T, PL, 0, double>;
			m.k[MV_P5DW][f] = (const void *)&k_p5s<T, PL, 1, double>;
			m.k[MV_P5DW2][f] = (const void *)&k_p5s<T, PL, 2, double>;
			m.nt[MV_P1W][f] = (const void *)&k_p1s<T, PL, double, true>;
			m.nt[MV_P5W][f] = (const void *)&k_p5s<T, PL, 0, double, true>;
			m.nt[MV_P5DW][f] = (const void *)&k_p5s<T, PL, 1, double, true>;
			m.nt[MV_P5DW2][f] = (const void *)&k_p5s<T, PL, 2, double, true>;
		}
		else {
			m.k[MV_P1W][f] = (const void *)&k_p1<T, PL, double>;
			m.k[MV_P5W][f] = (const void *)&k_p5<T, PL, false, double>;
			m.k[MV_P5DW][f] = m.k[MV_P5DW2][f] = (const void *)&k_p5<T, PL, true, double>;
		}
	}
}
template <class PL> void mv_fill(MvKernels &m)
{
	m.swz = PL::SWZ;
	mv_fill_t<double, PL>(m);
	mv_fill_t<float, PL>(m);
}

// Largest dynamic LDS a pass uses (set as the kernels' limit).
constexpr int kMvLdsMax = 96 * 1024;

template <typename T, class PL> hipError_t mv_lds_t()
{
	hipError_t e;
	auto lim = [&](const void *k) { return e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kMvLdsMax); };
	// P1/P5: the stage-fused gathers for compile-time plans, the cell-map gathers for run-time plans
	if constexpr (is_fix<PL>::value) {
		if (lim((const void *)k_p1s<T, PL>) || lim((const void *)k_p5s<T, PL, 0>) ||
			lim((const void *)k_p5s<T, PL, 1>) || lim((const void *)k_p5s<T, PL, 2>)) return e;
		if (lim((const void *)k_p1s<T, PL, T, true>) || lim((const void *)k_p5s<T, PL, 0, T, true>) ||
			lim((const void *)k_p5s<T, PL, 1, T, true>) || lim((const void *)k_p5s<T, PL, 2, T, true>)) return e;
		if constexpr (sizeof(T) == 4) {
			if (lim((const void *)k_p1s<T, PL, double>) || lim((const void *)k_p5s<T, PL, 0, double>) ||
				lim((const void *)k_p5s<T, PL, 1, double>) || lim((const void *)k_p5s<T, PL, 2, double>)) return e;
			if (lim((const void *)k_p1s<T, PL, double, true>) || lim((const void *)k_p5s<T, PL, 0, double, true>) ||
				lim((const void *)k_p5s<T, PL, 1, double, true>) || lim((const void *)k_p5s<T, PL, 2, double, true>))
				return e;
		}
		if (lim((const void *)k_p3g<T, PL>) || lim((const void *)k_p3g<T, PL, 1>)) return e;
		if constexpr (PL::SWZ)
			if (lim((const void *)k_p3<T, PL, 1, true>)) return e;
	}
	else {
		if (lim((const void *)k_p1<T, PL>) || lim((const void *)k_p5<T, PL, false>) ||
			lim((const void *)k_p5<T, PL, true>)) return e;
		if constexpr (sizeof(T) == 4)
			if (lim((const void *)k_p1<T, PL, double>) || lim((const void *)k_p5<T, PL, false, double>) ||
				lim((const void *)k_p5<T, PL, true, double>)) return e;
	}
	if (lim((const void *)k_p3<T, PL>) || lim((const void *)k_p3<T, PL, 1>) || lim((const void *)k_p3sub<T, PL>) ||
		lim((const void *)k_p3<T, PL, 0, false, true>) ||
		lim((const void *)k_zfft<T, false, PL>) ||
		lim((const void *)k_zfft<T, true, PL>)) return e;
	if constexpr (p3r_ok<T, PL>())
		if (lim((const void *)k_p3r<T, PL>)) return e;
	return hipSuccess;
}

template <class... Ps> struct PlanList {};

// Fills m with the kernels of the first plan of the list equal to p; false if none.
template <class... Ps> bool mv_find(const FftPlan2 &p, MvKernels &m, PlanList<Ps...>)
{
	return (... || (Ps::match(p) ? (mv_fill<Ps>(m), true) : false));
}

template <class... Ps> hipError_t mv_lds_all(PlanList<Ps...>)
{
	hipError_t e = hipSuccess;
	((e = e ? e : mv_lds_t<double, Ps>()), ...);
	((e = e ? e : mv_lds_t<float, Ps>()), ...);
	return e;
}

// Instantiation groups (one translation unit each).
bool mv_find_dyn(const FftPlan2 &p, MvKernels &m);  // always succeeds (run-time plan)
bool mv_find_pow2_large(const FftPlan2 &p, MvKernels &m);
bool mv_find_pow2_small(const FftPlan2 &p, MvKernels &m);
bool mv_find_mixed(const FftPlan2 &p, MvKernels &m);
hipError_t mv_lds_dyn();
hipError_t mv_lds_pow2_large();
hipError_t mv_lds_pow2_small();
hipError_t mv_lds_mixed();

}  // namespace adda
