// integral.hip -- A.A^dagger normalisation integrals on FP64 MFMA (gfx950).
//
// ampIntegralMatrix::integrate (decayAmplitude/ampIntegralMatrix.cc:408-483)
// accumulates I_ij = sum_e wt_e A_i(e) conj(A_j(e)).  With the amplitude row of
// an event read as the real vector x_e = (Re A_0, Im A_0, Re A_1, ...) of length
// D = 2W (the interleaved complex128 row, no copy), the real symmetric matrix
//   C = sum_e wt_e x_e x_e^T                      (D x D)
// holds every product:
//   Re I_ij = C[2i][2j] + C[2i+1][2j+1],   Im I_ij = C[2i+1][2j] - C[2i][2j+1].
// C is computed as a rank-k update on v_mfma_f64_16x16x4_f64: 16x16 output
// tiles of the upper triangle (tile row <= tile column), k = events.
// A persistent grid of one 512-thread workgroup (8 waves) per CU walks event chunks
// (grid-stride); each chunk is staged once through LDS and every wave issues
// the MFMAs of its tiles from it.  Rows of up to 192 doubles (96 waves) take the
// grouped kernel (tile map fixed at compile time, operand blocks read once per
// step, pairs of waves splitting the steps); wider rows take windowed passes of
// the same scheme over column windows of 8 blocks.  Each workgroup writes its
// tile sums to a partial row; partial rows are summed in fixed order
// (column_sum_kernel), and the host assembles the complex matrix.
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace rpwa_amd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kIntThreads = 512;

struct IntArgs {
	const double* x;       // [M][D] row-major (interleaved complex rows)
	const double* w;       // importance weights [M] or null (weight = 1/w)
	uint64_t M;
	uint32_t D;            // 2W
	uint32_t Dp;           // D padded to a multiple of 16
	uint32_t ldsStride;    // doubles per LDS row
	uint32_t nT;           // Dp / 16
	uint32_t nTiles;       // nT (nT + 1) / 2
	uint32_t tileBase;     // windowed kernel: the pass index (the first pass writes the weight sums)
	uint32_t winCol0;      // windowed kernel: first column (double) of the row window
	uint32_t winCol1;      // windowed kernel: first column of the column window (off-diagonal pass)
	uint32_t winI0;        // windowed kernel: first column block of the row window
	uint32_t winJ0;        // windowed kernel: first column block of the column window
	double* partials;      // [nWg][nTiles][256]
	double* wsum;          // [nWg] sum of weights (1/w), or null
	int accumulate;        // add to the existing partials
	int directWeights;     // weight = w[e] instead of 1/w[e]
};

// ---------------------------------------------------------------------------
// Grouped kernel for rows of up to 16 kGrpMaxBlocks doubles (NTB column blocks
// of 16, row padded to Dp = 16 NTB).  The tile -> column-block map is fixed at
// compile time, so a wave reads each 16-column operand block its tiles need once
// per 4-event step and keeps it in registers (at most NTB LDS reads for up to
// 20 MFMAs, instead of two reads per MFMA).  The upper-triangle tiles are split
// into four contiguous groups of near-equal size, one per pair of waves
// {g, g + 4}; the two waves of a pair take alternate 4-event steps of every
// chunk, so no wave computes a dummy tile.  At the end the second wave of a
// pair hands its sums to the first through LDS, in rounds that every wave runs
// (same barriers), and the first writes the pair's sum (fixed order).
constexpr int kGrpMaxBlocks = 12;
constexpr int kLdsBytes = 160 * 1024;

constexpr int tri_row(int nT, int t)
{
	int i = 0;
	while (t >= nT - i) {
		t -= nT - i;
		++i;
	}
	return i;
}
constexpr int tri_col(int nT, int t) { return t - (tri_row(nT, t) * nT - tri_row(nT, t) * (tri_row(nT, t) - 1) / 2) + tri_row(nT, t); }
constexpr int grp_begin(int nT, int g) { return g * (nT * (nT + 1) / 2) / 4; }
constexpr int grp_size(int nT, int g) { return grp_begin(nT, g + 1) - grp_begin(nT, g); }
constexpr int grp_max(int nT) { return (nT * (nT + 1) / 2 + 3) / 4; }
// column blocks read by the tiles of group g: as A (rows) or as either operand
constexpr uint32_t grp_blocks(int nT, int g, bool rowsOnly)
{
	uint32_t m = 0;
	for (int t = grp_begin(nT, g); t < grp_begin(nT, g + 1); ++t) {
		m |= 1u << tri_row(nT, t);
		if (!rowsOnly)
			m |= 1u << tri_col(nT, t);
	}
	return m;
}
// events per chunk: the largest multiple of 8 (up to 64) whose two LDS buffers,
// [CH][Dp + 1] rows plus CH weights each, fit beside the 64 weight partials
constexpr int grp_chunk(int nT)
{
	const int ch = (kLdsBytes - 64 * 8) / (2 * 8 * (16 * nT + 2)) / 8 * 8;
	return ch < 64 ? ch : 64;
}
constexpr int grp_lds_doubles(int nT) { return 2 * grp_chunk(nT) * (16 * nT + 2); }
// tiles per hand-over round: four groups' worth must fit in the chunk buffers
constexpr int grp_round(int nT)
{
	const int r = grp_lds_doubles(nT) / (4 * 256);
	return r < grp_max(nT) ? r : grp_max(nT);
}

template <int NT, int T>
struct TileRC {
	static constexpr int r = tri_row(NT, T);
	static constexpr int c = tri_col(NT, T);
};

// one 4-event step of group G: read the group's operand blocks from the LDS row
// xr (already offset by the lane's column), then one MFMA per tile
template <int NTB, int G, bool WT, int... Q>
__device__ __forceinline__ void grp_step(const double* __restrict__ xr, double wt, d4* acc,
                                         std::integer_sequence<int, Q...>)
{
	constexpr int t0 = grp_begin(NTB, G);
	constexpr uint32_t used = grp_blocks(NTB, G, false);
	constexpr uint32_t rows = grp_blocks(NTB, G, true);
	double v[NTB], va[NTB];
#pragma unroll
	for (int b = 0; b < NTB; ++b)
		v[b] = (used >> b & 1) ? xr[16 * b] : 0.;
#pragma unroll
	for (int b = 0; b < NTB; ++b)
		va[b] = (WT && (rows >> b & 1)) ? v[b] * wt : v[b];
	((acc[Q] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[TileRC<NTB, t0 + Q>::r], v[TileRC<NTB, t0 + Q>::c], acc[Q],
	                                                 0, 0, 0)),
	 ...);
}

template <int NTB, bool WEIGHTED, int G>
__device__ __forceinline__ void grp_body(const IntArgs& a, double* lds, double* wsh)
{
	constexpr int CH = grp_chunk(NTB);
	// doubles per LDS row (16 NTB + 1) and per chunk buffer (rows, then weights); as
	// run-time values the kernel measured 3% faster at 11 blocks than with them folded
	const uint32_t STRIDE = a.ldsStride;
	const uint32_t BUF = CH * STRIDE + CH;
	constexpr int KL = (CH * 16 * NTB / 2 + kIntThreads - 1) / kIntThreads;  // double2 loads per thread
	constexpr int NG = grp_size(NTB, G);
	constexpr int kMaxG = grp_max(NTB);
	constexpr int R = grp_round(NTB);
	static_assert(CH % 8 == 0 && R >= 1, "chunk of whole step pairs, hand-over round of a tile at least");
	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int h = tid >> 8;     // which wave of the pair (waves 0-3 or 4-7)
	const int li = lane & 15;
	const int lk = lane >> 4;

	d4 acc[NG > 0 ? NG : 1];
#pragma unroll
	for (int q = 0; q < NG; ++q)
		acc[q] = d4{0., 0., 0., 0.};

	const uint32_t chunkD2 = CH * a.D / 2;
	uint32_t ldsOff[KL];
#pragma unroll
	for (int l = 0; l < KL; ++l) {
		const uint32_t idx = 2 * ((uint32_t)tid + (uint32_t)l * kIntThreads);
		ldsOff[l] = (idx / a.D) * STRIDE + idx % a.D;
	}
	// padding columns (D..Dp) of both buffers stay zero
	for (uint32_t i = tid; i < 2 * BUF; i += kIntThreads)
		lds[i] = 0.;
	__syncthreads();

	double wacc = 0.;
	const uint64_t nChunks = (a.M + CH - 1) / CH;
	double2 stage[KL];
	double wstage = 0.;
	auto load_chunk = [&](uint64_t c) {
		const uint64_t e0 = c * CH;
		const uint64_t nEv = a.M - e0 < (uint64_t)CH ? a.M - e0 : (uint64_t)CH;
		const uint32_t lim = (uint32_t)nEv * a.D / 2;
		const double2* src = reinterpret_cast<const double2*>(a.x + e0 * a.D);
#pragma unroll
		for (int l = 0; l < KL; ++l) {
			const uint32_t i2 = (uint32_t)tid + (uint32_t)l * kIntThreads;
			stage[l] = (i2 < lim) ? src[i2] : make_double2(0., 0.);
		}
		if (tid < CH) {
			const uint64_t e = e0 + tid;
			wstage = e < a.M ? (a.w ? (a.directWeights ? a.w[e] : 1. / a.w[e]) : 1.) : 0.;
		}
	};
	auto store_chunk = [&](int buf) {
		double* xs = lds + buf * BUF;
#pragma unroll
		for (int l = 0; l < KL; ++l) {
			const uint32_t i2 = (uint32_t)tid + (uint32_t)l * kIntThreads;
			if (i2 < chunkD2) {
				xs[ldsOff[l]] = stage[l].x;
				xs[ldsOff[l] + 1] = stage[l].y;
			}
		}
		if (tid < CH) {
			xs[CH * STRIDE + tid] = wstage;
			wacc += wstage;
		}
	};

	uint64_t c = blockIdx.x;
	int buf = 0;
	if (c < nChunks) {
		load_chunk(c);
		store_chunk(0);
	}
	__syncthreads();
	for (; c < nChunks; c += gridDim.x) {
		const uint64_t next = c + gridDim.x;
		if (next < nChunks)
			load_chunk(next);  // in flight while the MFMAs below run
		const double* xs = lds + buf * BUF;
		const double* wts = xs + CH * STRIDE;
		// not unrolled: unrolled, the steps' operand reads are hoisted together (more
		// registers, 4% slower at 11 blocks)
#pragma unroll 1
		for (int kk = 0; kk < CH / 8; ++kk) {
			const int row = (2 * kk + h) * 4 + lk;
			const double wt = WEIGHTED ? wts[row] : 1.;
			grp_step<NTB, G, WEIGHTED>(xs + row * STRIDE + li, wt, acc, std::make_integer_sequence<int, NG>{});
		}
		if (next < nChunks)
			store_chunk(buf ^ 1);
		__syncthreads();
		buf ^= 1;
	}

	// hand-over: [group][tile of the round][register][lane] in the chunk buffers;
	// v_mfma_f64_16x16x4_f64 leaves D[i][j] in lane j + 16 (i % 4), register i / 4
	double* red = lds + G * R * 256;
	double* out = a.partials + (size_t)blockIdx.x * a.nTiles * 256;
#pragma unroll
	for (int r0 = 0; r0 < kMaxG; r0 += R) {
		if (h == 1) {
#pragma unroll
			for (int q = r0; q < NG && q < r0 + R; ++q)
#pragma unroll
				for (int v = 0; v < 4; ++v)
					red[((q - r0) * 4 + v) * 64 + lane] = acc[q][v];
		}
		__syncthreads();
		if (h == 0) {
#pragma unroll
			for (int q = r0; q < NG && q < r0 + R; ++q) {
				double* o = out + (size_t)(grp_begin(NTB, G) + q) * 256;
#pragma unroll
				for (int v = 0; v < 4; ++v) {
					const int i = lk + 4 * v;
					double val = acc[q][v] + red[((q - r0) * 4 + v) * 64 + lane];
					if (a.accumulate)
						val += o[i * 16 + li];
					o[i * 16 + li] = val;
				}
			}
		}
		__syncthreads();
	}
	if (tid < CH)
		wsh[tid] = wacc;
	__syncthreads();
	if (a.wsum && tid == 0) {
		double s = 0.;
		for (int r = 0; r < CH; ++r)
			s += wsh[r];
		a.wsum[blockIdx.x] = a.accumulate ? a.wsum[blockIdx.x] + s : s;
	}
}

template <int NTB, bool WEIGHTED>
__global__ __launch_bounds__(kIntThreads) void integral_grp_kernel(IntArgs a)
{
	extern __shared__ double lds[];
	__shared__ double wsh[64];
	// wave-uniform branch: waves g and g + 4 run group g
	switch ((threadIdx.x >> 6) & 3) {
	case 0: grp_body<NTB, WEIGHTED, 0>(a, lds, wsh); break;
	case 1: grp_body<NTB, WEIGHTED, 1>(a, lds, wsh); break;
	case 2: grp_body<NTB, WEIGHTED, 2>(a, lds, wsh); break;
	default: grp_body<NTB, WEIGHTED, 3>(a, lds, wsh); break;
	}
}

// ---------------------------------------------------------------------------
// Windowed passes for rows wider than 16 kGrpMaxBlocks doubles (the Hessian's
// derivative rows): the column blocks are cut into windows of kWin; a pass
// stages only its windows' columns of every row (one window for a diagonal
// pass, the kWin (kWin + 1) / 2 tiles of its triangle; two windows for an
// off-diagonal pass, the kWin x kWin tiles between them) and computes its tiles
// with the grouped kernel's scheme (compile-time tile map, operand blocks read
// once per step, pairs of waves splitting the steps).  Blocks past the row
// (the last window's padding) are staged as zeros and their tiles not written.
constexpr int kWin = 8;

constexpr int win_blocks_of(bool rect) { return rect ? 2 * kWin : kWin; }
constexpr int win_tiles(bool rect) { return rect ? kWin * kWin : kWin * (kWin + 1) / 2; }
constexpr int win_row(bool rect, int t) { return rect ? t / kWin : tri_row(kWin, t); }
constexpr int win_col(bool rect, int t) { return rect ? kWin + t % kWin : tri_col(kWin, t); }
constexpr int win_begin(bool rect, int g) { return g * win_tiles(rect) / 4; }
constexpr int win_size(bool rect, int g) { return win_begin(rect, g + 1) - win_begin(rect, g); }
constexpr int win_max(bool rect) { return (win_tiles(rect) + 3) / 4; }
constexpr uint32_t win_blocks(bool rect, int g, bool rowsOnly)
{
	uint32_t m = 0;
	for (int t = win_begin(rect, g); t < win_begin(rect, g + 1); ++t) {
		m |= 1u << win_row(rect, t);
		if (!rowsOnly)
			m |= 1u << win_col(rect, t);
	}
	return m;
}
constexpr int win_chunk(bool rect)
{
	const int ch = (kLdsBytes - 64 * 8) / (2 * 8 * (16 * win_blocks_of(rect) + 2)) / 8 * 8;
	return ch < 64 ? ch : 64;
}
constexpr int win_lds_doubles(bool rect) { return 2 * win_chunk(rect) * (16 * win_blocks_of(rect) + 2); }

template <bool RECT, int T>
struct WinRC {
	static constexpr int r = win_row(RECT, T);
	static constexpr int c = win_col(RECT, T);
};

template <bool RECT, int G, bool WT, int... Q>
__device__ __forceinline__ void win_step(const double* __restrict__ xr, double wt, d4* acc,
                                         std::integer_sequence<int, Q...>)
{
	constexpr int NB = win_blocks_of(RECT);
	constexpr int t0 = win_begin(RECT, G);
	constexpr uint32_t used = win_blocks(RECT, G, false);
	constexpr uint32_t rows = win_blocks(RECT, G, true);
	double v[NB], va[NB];
#pragma unroll
	for (int b = 0; b < NB; ++b)
		v[b] = (used >> b & 1) ? xr[16 * b] : 0.;
#pragma unroll
	for (int b = 0; b < NB; ++b)
		va[b] = (WT && (rows >> b & 1)) ? v[b] * wt : v[b];
	((acc[Q] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[WinRC<RECT, t0 + Q>::r], v[WinRC<RECT, t0 + Q>::c], acc[Q],
	                                                 0, 0, 0)),
	 ...);
}

template <bool RECT, bool WEIGHTED, int G>
__device__ __forceinline__ void win_body(const IntArgs& a, double* lds, double* wsh)
{
	constexpr int CH = win_chunk(RECT);
	constexpr int WS = 16 * win_blocks_of(RECT);       // staged doubles per row
	constexpr int W0 = 16 * kWin;                      // doubles of the row window
	constexpr int KL = CH * WS / 2 / kIntThreads;      // double2 loads per thread (exact)
	constexpr int NG = win_size(RECT, G);
	constexpr int kMaxG = win_max(RECT);
	static_assert(CH * WS / 2 == KL * kIntThreads && CH % 8 == 0, "chunk of whole loads and step pairs");
	static_assert(win_lds_doubles(RECT) >= 4 * kMaxG * 256, "hand-over in one round");
	// LDS row stride (WS + 1) as a run-time value, as in the grouped kernel
	const uint32_t STRIDE = a.ldsStride;
	const uint32_t BUF = CH * STRIDE + CH;
	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int h = tid >> 8;
	const int li = lane & 15;
	const int lk = lane >> 4;

	d4 acc[NG];
#pragma unroll
	for (int q = 0; q < NG; ++q)
		acc[q] = d4{0., 0., 0., 0.};

	// staging geometry: double2 number tid + 512 l of the chunk's staged block is
	// column k of event e of the block; its source column g in the row
	uint32_t ldsOff[KL], srcOff[KL], srcEv[KL];
	bool colOk[KL];
#pragma unroll
	for (int l = 0; l < KL; ++l) {
		const uint32_t i2 = (uint32_t)tid + (uint32_t)l * kIntThreads;
		const uint32_t e = i2 / (WS / 2), k = 2 * (i2 % (WS / 2));
		const uint32_t g = k < (uint32_t)W0 ? a.winCol0 + k : a.winCol1 + (k - W0);
		ldsOff[l] = e * STRIDE + k;
		srcOff[l] = e * a.D + g;
		srcEv[l] = e;
		colOk[l] = g < a.D;
	}
	for (uint32_t i = tid; i < 2 * BUF; i += kIntThreads)
		lds[i] = 0.;
	__syncthreads();

	double wacc = 0.;
	const uint64_t nChunks = (a.M + CH - 1) / CH;
	double2 stage[KL];
	double wstage = 0.;
	auto load_chunk = [&](uint64_t c) {
		const uint64_t e0 = c * CH;
		const uint64_t nEv = a.M - e0 < (uint64_t)CH ? a.M - e0 : (uint64_t)CH;
		const double* src = a.x + e0 * a.D;
#pragma unroll
		for (int l = 0; l < KL; ++l)
			stage[l] = (colOk[l] && srcEv[l] < (uint32_t)nEv) ? *reinterpret_cast<const double2*>(src + srcOff[l])
			                                                  : make_double2(0., 0.);
		if (tid < CH) {
			const uint64_t e = e0 + tid;
			wstage = e < a.M ? (a.w ? (a.directWeights ? a.w[e] : 1. / a.w[e]) : 1.) : 0.;
		}
	};
	auto store_chunk = [&](int buf) {
		double* xs = lds + buf * BUF;
#pragma unroll
		for (int l = 0; l < KL; ++l) {
			xs[ldsOff[l]] = stage[l].x;
			xs[ldsOff[l] + 1] = stage[l].y;
		}
		if (tid < CH) {
			xs[CH * STRIDE + tid] = wstage;
			wacc += wstage;
		}
	};

	uint64_t c = blockIdx.x;
	int buf = 0;
	if (c < nChunks) {
		load_chunk(c);
		store_chunk(0);
	}
	__syncthreads();
	for (; c < nChunks; c += gridDim.x) {
		const uint64_t next = c + gridDim.x;
		if (next < nChunks)
			load_chunk(next);
		const double* xs = lds + buf * BUF;
		const double* wts = xs + CH * STRIDE;
#pragma unroll 1
		for (int kk = 0; kk < CH / 8; ++kk) {
			const int row = (2 * kk + h) * 4 + lk;
			const double wt = WEIGHTED ? wts[row] : 1.;
			win_step<RECT, G, WEIGHTED>(xs + row * STRIDE + li, wt, acc, std::make_integer_sequence<int, NG>{});
		}
		if (next < nChunks)
			store_chunk(buf ^ 1);
		__syncthreads();
		buf ^= 1;
	}

	// hand-over as in the grouped kernel (one round), then the pair's sums go to
	// their tiles of the full triangle (padding tiles past the row are dropped)
	double* red = lds + G * kMaxG * 256;
	double* out = a.partials + (size_t)blockIdx.x * a.nTiles * 256;
	if (h == 1) {
#pragma unroll
		for (int q = 0; q < NG; ++q)
#pragma unroll
			for (int v = 0; v < 4; ++v)
				red[(q * 4 + v) * 64 + lane] = acc[q][v];
	}
	__syncthreads();
	if (h == 0) {
#pragma unroll
		for (int q = 0; q < NG; ++q) {
			const int t = win_begin(RECT, G) + q;
			const uint32_t I = a.winI0 + win_row(RECT, t);
			const uint32_t J = RECT ? a.winJ0 + (win_col(RECT, t) - kWin) : a.winI0 + win_col(RECT, t);
			if (J >= a.nT)
				continue;
			double* o = out + (size_t)(I * a.nT - I * (I - 1) / 2 + (J - I)) * 256;
#pragma unroll
			for (int v = 0; v < 4; ++v) {
				const int i = lk + 4 * v;
				double val = acc[q][v] + red[(q * 4 + v) * 64 + lane];
				if (a.accumulate)
					val += o[i * 16 + li];
				o[i * 16 + li] = val;
			}
		}
	}
	__syncthreads();
	if (tid < CH)
		wsh[tid] = wacc;
	__syncthreads();
	if (a.wsum && a.tileBase == 0 && tid == 0) {
		double s = 0.;
		for (int r = 0; r < CH; ++r)
			s += wsh[r];
		a.wsum[blockIdx.x] = a.accumulate ? a.wsum[blockIdx.x] + s : s;
	}
}

template <bool RECT, bool WEIGHTED>
__global__ __launch_bounds__(kIntThreads) void integral_win_kernel(IntArgs a)
{
	extern __shared__ double lds[];
	__shared__ double wsh[64];
	switch ((threadIdx.x >> 6) & 3) {
	case 0: win_body<RECT, WEIGHTED, 0>(a, lds, wsh); break;
	case 1: win_body<RECT, WEIGHTED, 1>(a, lds, wsh); break;
	case 2: win_body<RECT, WEIGHTED, 2>(a, lds, wsh); break;
	default: win_body<RECT, WEIGHTED, 3>(a, lds, wsh); break;
	}
}

// fixed-order column sums of a [nRows][nCols] array
constexpr int kColElems = 64;
constexpr int kColChunks = 16;
__global__ __launch_bounds__(kColElems* kColChunks) void column_sum_kernel(
    const double* __restrict__ in, int nRows, uint32_t nCols, double* __restrict__ out)
{
	__shared__ double sh[kColChunks][kColElems];
	const int el = threadIdx.x % kColElems;
	const int ch = threadIdx.x / kColElems;
	const uint32_t col = blockIdx.x * kColElems + el;
	const int r0 = (int)((int64_t)nRows * ch / kColChunks);
	const int r1 = (int)((int64_t)nRows * (ch + 1) / kColChunks);
	double s = 0.;
	if (col < nCols)
		for (int r = r0; r < r1; ++r)
			s += in[(size_t)r * nCols + col];
	sh[ch][el] = s;
	__syncthreads();
	if (ch == 0 && col < nCols) {
		double t = 0.;
		for (int k = 0; k < kColChunks; ++k)
			t += sh[k][el];
		out[col] = t;
	}
}

}  // namespace

typedef void (*IntKernel)(IntArgs);
template <bool WT>
static IntKernel pick_grp(uint32_t nT)
{
	switch (nT) {
	case 1: return integral_grp_kernel<1, WT>;
	case 2: return integral_grp_kernel<2, WT>;
	case 3: return integral_grp_kernel<3, WT>;
	case 4: return integral_grp_kernel<4, WT>;
	case 5: return integral_grp_kernel<5, WT>;
	case 6: return integral_grp_kernel<6, WT>;
	case 7: return integral_grp_kernel<7, WT>;
	case 8: return integral_grp_kernel<8, WT>;
	case 9: return integral_grp_kernel<9, WT>;
	case 10: return integral_grp_kernel<10, WT>;
	case 11: return integral_grp_kernel<11, WT>;
	case 12: return integral_grp_kernel<12, WT>;
	default: return nullptr;
	}
}
static int grp_chunk_of(uint32_t nT) { return grp_chunk((int)nT); }

struct IntegralPlan {
	uint32_t D, Dp, nT, nTiles, ldsStride;
	int chunk;
	int nWg;
	size_t lds;
	IntKernel kern = nullptr;  // grouped kernel (one pass over every tile)
	// windowed passes (rows of more than 16 kGrpMaxBlocks doubles): diagonal and
	// off-diagonal window kernels, their LDS row strides and dynamic LDS
	bool windowed = false;
	IntKernel kernDiag = nullptr, kernRect = nullptr;
	uint32_t strideDiag = 0, strideRect = 0;
	size_t ldsDiag = 0, ldsRect = 0;
};

// D = real row length (2 x complex values per row)
static int integral_plan(int device, uint32_t D, bool weighted, IntegralPlan& p)
{
	p.D = D;
	p.Dp = (p.D + 15) / 16 * 16;
	p.nT = p.Dp / 16;
	p.nTiles = p.nT * (p.nT + 1) / 2;
	p.ldsStride = p.Dp + 1;  // odd stride spreads the 4 rows of a read over the banks
	p.kern = weighted ? pick_grp<true>(p.nT) : pick_grp<false>(p.nT);
	if (p.kern) {
		p.chunk = grp_chunk_of(p.nT);
		p.lds = 2 * ((size_t)p.chunk * p.ldsStride + p.chunk) * sizeof(double);
	} else {
		// wider rows (the Hessian's derivative rows): windowed passes
		p.windowed = true;
		p.kernDiag = weighted ? integral_win_kernel<false, true> : integral_win_kernel<false, false>;
		p.kernRect = weighted ? integral_win_kernel<true, true> : integral_win_kernel<true, false>;
		p.strideDiag = 16 * win_blocks_of(false) + 1;
		p.strideRect = 16 * win_blocks_of(true) + 1;
		p.ldsDiag = (size_t)win_lds_doubles(false) * sizeof(double);
		p.ldsRect = (size_t)win_lds_doubles(true) * sizeof(double);
		p.chunk = win_chunk(true);
		p.lds = std::max(p.ldsDiag, p.ldsRect);
	}
	int nCu = 0;
	RPWA_HIP_CHECK(hipDeviceGetAttribute(&nCu, hipDeviceAttributeMultiprocessorCount, device));
	p.nWg = nCu > 0 ? nCu : 1;
	return RPWA_OK;
}

// Launch the rank-k update's passes on stream s: one grouped launch, or the
// windowed diagonal and off-diagonal passes (every tile of the triangle written
// by exactly one pass).
static hipError_t launch_plan(const IntegralPlan& p, IntArgs a, hipStream_t s)
{
	auto allow = [](IntKernel k, size_t lds) -> hipError_t {
		// dynamic LDS beyond 64 KiB is requested explicitly
		if (lds + 64 * sizeof(double) > 64u * 1024u)
			return hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
			                           (int)lds);
		return hipSuccess;
	};
	if (p.windowed) {
		if (hipError_t e = allow(p.kernDiag, p.ldsDiag))
			return e;
		if (hipError_t e = allow(p.kernRect, p.ldsRect))
			return e;
		const uint32_t nWin = (p.nT + kWin - 1) / kWin;
		uint32_t pass = 0;
		for (uint32_t u = 0; u < nWin; ++u)
			for (uint32_t v = u; v < nWin; ++v, ++pass) {
				const bool rect = v > u;
				a.tileBase = pass;
				a.winI0 = u * kWin;
				a.winJ0 = v * kWin;
				a.winCol0 = 16 * u * kWin;
				a.winCol1 = 16 * v * kWin;
				a.ldsStride = rect ? p.strideRect : p.strideDiag;
				hipLaunchKernelGGL(rect ? p.kernRect : p.kernDiag, dim3(p.nWg), dim3(kIntThreads),
				                   rect ? p.ldsRect : p.ldsDiag, s, a);
				if (hipError_t e = hipGetLastError())
					return e;
			}
		return hipSuccess;
	}
	if (hipError_t e = allow(p.kern, p.lds))
		return e;
	a.tileBase = 0;
	hipLaunchKernelGGL(p.kern, dim3(p.nWg), dim3(kIntThreads), p.lds, s, a);
	return hipGetLastError();
}

// Per-device work area of the rank-k update (workgroup partial tiles, weight
// sums, reduced tiles) and its stream, kept between calls and grown on demand:
// a call then allocates nothing of its own on the device (an allocation and a
// free of the ~35 MB partials, with the free's device synchronisation, cost
// about as much as a tenth of the update at 5M x 88).  Calls on one device are
// serialised by the mutex; the area lives until the process ends.
struct SyrkWork {
	std::mutex m;
	hipStream_t s = nullptr;
	double* part = nullptr;
	double* wsum = nullptr;
	double* red = nullptr;
	size_t partCap = 0, wsumCap = 0, redCap = 0;
};
constexpr int kSyrkDevices = 64;

static SyrkWork& syrk_work(int device)
{
	static SyrkWork work[kSyrkDevices];
	return work[device];
}

static hipError_t syrk_grow(double*& buf, size_t& cap, size_t need)
{
	if (need <= cap)
		return hipSuccess;
	if (buf)
		(void)hipFree(buf);
	buf = nullptr;
	cap = 0;
	hipError_t e = hipMalloc(&buf, need * sizeof(double));
	if (e == hipSuccess)
		cap = need;
	return e;
}

// Weighted real rank-k update C = sum_e wt_e x_e x_e^T of M rows of D doubles
// (device or host memory); wt_e = w[e] (directWeights) or 1/w[e] or 1.
// Returns the full symmetric D x D matrix and the sum of the weights.
int syrk_run(int device, const double* d_x, uint64_t M, uint32_t D, const double* d_w, bool directWeights,
             const double* h_x, const double* h_w, std::vector<double>& full, double* wtot)
{
	if (device < 0 || device >= kSyrkDevices) {
		set_error("rank-k update: device index out of range");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(device));
	const bool weighted = (d_w != nullptr) || (h_w != nullptr);
	IntegralPlan p;
	int st = integral_plan(device, D, weighted, p);
	if (st != RPWA_OK)
		return st;
	SyrkWork& work = syrk_work(device);
	std::lock_guard<std::mutex> lock(work.m);
	if (!work.s)
		RPWA_HIP_CHECK(hipStreamCreateWithFlags(&work.s, hipStreamNonBlocking));
	hipStream_t s = work.s;
	double* d_stage[2] = {nullptr, nullptr};
	double* d_wstage[2] = {nullptr, nullptr};
	const size_t partN = (size_t)p.nWg * p.nTiles * 256;
	std::vector<double> C((size_t)p.nTiles * 256);
	std::vector<double> wsums(p.nWg);
	auto fail = [&](int code) {
		if (code != RPWA_OK)
			(void)hipStreamSynchronize(s);  // nothing of this call may still run on the work area
		(void)hipFree(d_stage[0]); (void)hipFree(d_stage[1]); (void)hipFree(d_wstage[0]); (void)hipFree(d_wstage[1]);
		return code;
	};
#define INT_CHECK(expr)                                                               \
	do {                                                                              \
		hipError_t e_ = (expr);                                                       \
		if (e_ != hipSuccess) {                                                       \
			set_error(std::string(#expr) + ": " + hipGetErrorString(e_));             \
			return fail(RPWA_ERR_HIP);                                                \
		}                                                                             \
	} while (0)
	INT_CHECK(syrk_grow(work.part, work.partCap, partN));
	INT_CHECK(syrk_grow(work.wsum, work.wsumCap, p.nWg));
	INT_CHECK(syrk_grow(work.red, work.redCap, (size_t)p.nTiles * 256));
	double* d_part = work.part;
	double* d_wsum = work.wsum;
	double* d_red = work.red;
	// every pass re-reads the rows for its own set of output tiles
	auto launch_passes = [&](IntArgs aa) -> hipError_t { return launch_plan(p, aa, s); };
	IntArgs a;
	a.D = p.D; a.Dp = p.Dp; a.ldsStride = p.ldsStride; a.nT = p.nT; a.nTiles = p.nTiles;
	a.partials = d_part; a.wsum = d_wsum; a.directWeights = directWeights ? 1 : 0;
	if (d_x) {
		a.x = d_x; a.w = d_w; a.M = M; a.accumulate = 0;
		INT_CHECK(launch_passes(a));
	} else {
		// stream the host array through two device buffers
		const uint64_t chunkEv = std::max<uint64_t>(1, (256ull << 20) / (8ull * D));
		for (int b = 0; b < 2; ++b) {
			INT_CHECK(hipMalloc(&d_stage[b], chunkEv * 8ull * D));
			if (weighted)
				INT_CHECK(hipMalloc(&d_wstage[b], chunkEv * sizeof(double)));
		}
		hipEvent_t done[2];
		INT_CHECK(hipEventCreate(&done[0]));
		INT_CHECK(hipEventCreate(&done[1]));
		bool first = true;
		int b = 0;
		for (uint64_t off = 0; off < M; off += chunkEv, b ^= 1) {
			const uint64_t n = std::min<uint64_t>(chunkEv, M - off);
			if (!first)
				INT_CHECK(hipEventSynchronize(done[b]));
			INT_CHECK(hipMemcpyAsync(d_stage[b], h_x + off * D, n * 8ull * D, hipMemcpyHostToDevice, s));
			if (weighted)
				INT_CHECK(hipMemcpyAsync(d_wstage[b], h_w + off, n * sizeof(double), hipMemcpyHostToDevice, s));
			a.x = d_stage[b]; a.w = weighted ? d_wstage[b] : nullptr; a.M = n;
			a.accumulate = first ? 0 : 1;
			INT_CHECK(launch_passes(a));
			INT_CHECK(hipEventRecord(done[b], s));
			first = false;
		}
		INT_CHECK(hipStreamSynchronize(s));
		(void)hipEventDestroy(done[0]);
		(void)hipEventDestroy(done[1]);
	}
	const int nb = (int)((p.nTiles * 256 + kColElems - 1) / kColElems);
	hipLaunchKernelGGL(column_sum_kernel, dim3(nb), dim3(kColElems * kColChunks), 0, s, d_part, p.nWg,
	                   p.nTiles * 256, d_red);
	INT_CHECK(hipGetLastError());
	INT_CHECK(hipMemcpyAsync(C.data(), d_red, C.size() * sizeof(double), hipMemcpyDeviceToHost, s));
	INT_CHECK(hipMemcpyAsync(wsums.data(), d_wsum, wsums.size() * sizeof(double), hipMemcpyDeviceToHost, s));
	INT_CHECK(hipStreamSynchronize(s));
#undef INT_CHECK
	// the symmetric matrix from the upper-triangle tiles, tile by tile: a
	// diagonal tile as computed (with weights its two halves can differ in the
	// last bit), an off-diagonal tile at (I, J) and transposed at (J, I)
	full.assign((size_t)D * D, 0.);
	for (uint32_t I = 0, t = 0; I < p.nT; ++I)
		for (uint32_t J = I; J < p.nT; ++J, ++t) {
			const double* tile = C.data() + (size_t)t * 256;
			const uint32_t r1 = std::min<uint32_t>(16, D - 16 * I), c1 = std::min<uint32_t>(16, D - 16 * J);
			for (uint32_t r = 0; r < r1; ++r)
				for (uint32_t c = 0; c < c1; ++c) {
					const uint32_t row = 16 * I + r, col = 16 * J + c;
					full[(size_t)row * D + col] = tile[r * 16 + c];
					if (I != J)
						full[(size_t)col * D + row] = tile[r * 16 + c];
				}
		}
	if (wtot) {
		double t = 0.;
		for (double v : wsums)
			t += v;
		*wtot = t;
	}
	fail(RPWA_OK);
	return RPWA_OK;
}

static int integral_run(int device, const double* d_amps, uint64_t M, uint32_t W, const double* d_w,
                        double* outSums, const double* h_amps, const double* h_w)
{
	std::vector<double> C;
	double wtot = 0.;
	const uint32_t D = 2 * W;
	int st = syrk_run(device, d_amps, M, D, d_w, false, h_amps, h_w, C, &wtot);
	if (st != RPWA_OK)
		return st;
	for (uint32_t i = 0; i < W; ++i)
		for (uint32_t j = 0; j < W; ++j) {
			const double re = C[(size_t)(2 * i) * D + 2 * j] + C[(size_t)(2 * i + 1) * D + 2 * j + 1];
			const double im = C[(size_t)(2 * i + 1) * D + 2 * j] - C[(size_t)(2 * i) * D + 2 * j + 1];
			outSums[2 * ((size_t)i * W + j)] = re;
			outSums[2 * ((size_t)i * W + j) + 1] = im;
		}
	if (d_w || h_w) {
		// renormalise to the weight integral (ampIntegralMatrix.cc:477-483)
		const double weightNorm = wtot / (double)M;
		const double inv = 1 / weightNorm;
		for (size_t i = 0; i < (size_t)2 * W * W; ++i)
			outSums[i] *= inv;
	}
	return RPWA_OK;
}

// per event w_e = 1 / sum_i |A_i(e)|^2 / I_ii (0 for an all-zero event)
__global__ void inverse_norm_kernel(const double2* __restrict__ amps, uint64_t M, uint32_t W,
                                    const double* __restrict__ invDiag, double* __restrict__ w)
{
	const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= M)
		return;
	double s = 0.;
	for (uint32_t i = 0; i < W; ++i) {
		const double2 a = amps[e * W + i];
		s += (a.x * a.x + a.y * a.y) * invDiag[i];
	}
	w[e] = s > 0. ? 1. / s : 0.;
}

static int moments_run(int device, const double* d_amps, uint64_t M, uint32_t W, const double* normDiag,
                       int perEventNorm, double* out)
{
	std::vector<double> C;
	double wtot = 0.;
	const uint32_t D = 2 * W;
	double* d_w = nullptr;
	double* d_inv = nullptr;
	std::vector<double> inv(W);
	for (uint32_t i = 0; i < W; ++i)
		inv[i] = normDiag[i] > 0. ? 1. / normDiag[i] : 0.;
	if (perEventNorm) {
		RPWA_HIP_CHECK(hipSetDevice(device));
		RPWA_HIP_CHECK(hipMalloc(&d_w, M * sizeof(double)));
		if (hipMalloc(&d_inv, W * sizeof(double)) != hipSuccess ||
		    hipMemcpy(d_inv, inv.data(), W * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
			(void)hipFree(d_w);
			if (d_inv)
				(void)hipFree(d_inv);
			set_error("rpwa_data_moments_device: allocation or copy failed");
			return RPWA_ERR_HIP;
		}
		hipLaunchKernelGGL(inverse_norm_kernel, dim3((uint32_t)((M + 255) / 256)), dim3(256), 0, 0,
		                   reinterpret_cast<const double2*>(d_amps), M, W, d_inv, d_w);
		const hipError_t e = hipDeviceSynchronize();
		(void)hipFree(d_inv);
		if (e != hipSuccess) {
			(void)hipFree(d_w);
			set_error(std::string("rpwa_data_moments_device: ") + hipGetErrorString(e));
			return RPWA_ERR_HIP;
		}
	}
	const int st = syrk_run(device, d_amps, M, D, d_w, true, nullptr, nullptr, C, &wtot);
	if (d_w)
		(void)hipFree(d_w);
	if (st != RPWA_OK)
		return st;
	for (uint32_t i = 0; i < W; ++i)
		for (uint32_t j = 0; j < W; ++j) {
			const double re = C[(size_t)(2 * i) * D + 2 * j] + C[(size_t)(2 * i + 1) * D + 2 * j + 1];
			const double im = C[(size_t)(2 * i + 1) * D + 2 * j] - C[(size_t)(2 * i) * D + 2 * j + 1];
			const double sc = std::sqrt(inv[i] * inv[j]) / (double)M;
			out[2 * ((size_t)i * W + j)] = re * sc;
			out[2 * ((size_t)i * W + j) + 1] = im * sc;
		}
	return RPWA_OK;
}

// ---------------------------------------------------------------------------
// The integral without allocations (rpwa_integral_device_noalloc): the same
// kernels and reduction order as integral_run on device input, with every
// intermediate in one static device array instead of hipMalloc'd buffers.
// Layout of the array (doubles): workgroup partial tiles
// [kNaWg][kNaTiles][256], workgroup weight sums [kNaWg], reduced tiles
// [kNaTiles][256], the complex result [W][W].  Every element a call reads is
// written earlier in the same call, so nothing needs clearing between calls.
constexpr uint32_t kNaMaxWaves = 128;                       // D = 256: 16 x 16 tiles of 16 x 16
constexpr uint32_t kNaTiles = (2 * kNaMaxWaves / 16) * (2 * kNaMaxWaves / 16 + 1) / 2;  // 136
constexpr int kNaWg = 256;                                  // one workgroup per CU of an MI355X
constexpr size_t kNaPartial = (size_t)kNaWg * kNaTiles * 256;
constexpr size_t kNaWsum = kNaPartial;
constexpr size_t kNaRed = kNaWsum + kNaWg;
constexpr size_t kNaOut = kNaRed + (size_t)kNaTiles * 256;
constexpr size_t kNaDoubles = kNaOut + 2ull * kNaMaxWaves * kNaMaxWaves;  // 8,980,736 doubles = 68.5 MiB
constexpr int kNaMaxDevices = 64;

static __device__ __align__(256) double scratch[kNaDoubles];

namespace {

// per-device address of `scratch`, looked up once per device with that device current
std::once_flag g_naOnce[kNaMaxDevices];
double* g_naPtr[kNaMaxDevices];
hipError_t g_naErr[kNaMaxDevices];

// the complex sums [W][W] from the reduced tiles (the host assembly of integral_run),
// renormalised by the mean weight when weights are given (ampIntegralMatrix.cc:477-483)
__global__ __launch_bounds__(256) void assemble_kernel(const double* __restrict__ red, uint32_t nT, uint32_t W,
                                                       const double* __restrict__ wsum, int nWg, uint64_t M,
                                                       int weighted, double* __restrict__ out)
{
	const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= W * W)
		return;
	const uint32_t i = idx / W, j = idx % W;
	auto c = [&](uint32_t r, uint32_t col) {
		uint32_t I = r / 16, J = col / 16;
		if (I > J) {
			const uint32_t t = r;
			r = col;
			col = t;
			const uint32_t u = I;
			I = J;
			J = u;
		}
		const uint32_t t = I * nT - I * (I - 1) / 2 + (J - I);
		return red[(size_t)t * 256 + (r % 16) * 16 + (col % 16)];
	};
	double re = c(2 * i, 2 * j) + c(2 * i + 1, 2 * j + 1);
	double im = c(2 * i + 1, 2 * j) - c(2 * i, 2 * j + 1);
	if (weighted) {
		double t = 0.;
		for (int r = 0; r < nWg; ++r)
			t += wsum[r];
		const double weightNorm = t / (double)M;
		const double inv = 1 / weightNorm;
		re *= inv;
		im *= inv;
	}
	out[2 * (size_t)idx] = re;
	out[2 * (size_t)idx + 1] = im;
}

}  // namespace

static int integral_noalloc(int device, const double* d_amps, uint64_t M, uint32_t W, const double* d_w,
                            double* outSums)
{
	if (W > kNaMaxWaves) {
		set_error("rpwa_integral_device_noalloc: " + std::to_string(W) + " waves exceed the static work area (at "
		          "most " + std::to_string(kNaMaxWaves) + "); use rpwa_integral_device");
		return RPWA_ERR_UNSUPPORTED;
	}
	int nDev = 0;
	RPWA_HIP_CHECK(hipGetDeviceCount(&nDev));
	if (device < 0 || device >= nDev || device >= kNaMaxDevices) {
		set_error("rpwa_integral_device_noalloc: no GPU with index " + std::to_string(device));
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(device));
	std::call_once(g_naOnce[device], [device]() {
		g_naErr[device] = hipGetSymbolAddress(reinterpret_cast<void**>(&g_naPtr[device]), HIP_SYMBOL(scratch));
	});
	if (g_naErr[device] != hipSuccess || !g_naPtr[device]) {
		set_error(std::string("rpwa_integral_device_noalloc: hipGetSymbolAddress(scratch): ") +
		          hipGetErrorString(g_naErr[device]));
		return RPWA_ERR_HIP;
	}
	double* const base = g_naPtr[device];
	const bool weighted = d_w != nullptr;
	IntegralPlan p;
	if (int st = integral_plan(device, 2 * W, weighted, p))
		return st;
	p.nWg = std::min(p.nWg, kNaWg);
	IntArgs a;
	a.x = d_amps;
	a.w = d_w;
	a.M = M;
	a.D = p.D;
	a.Dp = p.Dp;
	a.ldsStride = p.ldsStride;
	a.nT = p.nT;
	a.nTiles = p.nTiles;
	a.partials = base;
	a.wsum = base + kNaWsum;
	a.accumulate = 0;
	a.directWeights = 0;
	RPWA_HIP_CHECK(launch_plan(p, a, 0));
	const int nb = (int)((p.nTiles * 256 + kColElems - 1) / kColElems);
	hipLaunchKernelGGL(column_sum_kernel, dim3(nb), dim3(kColElems * kColChunks), 0, 0, base, p.nWg,
	                   p.nTiles * 256, base + kNaRed);
	RPWA_HIP_CHECK(hipGetLastError());
	hipLaunchKernelGGL(assemble_kernel, dim3((W * W + 255) / 256), dim3(256), 0, 0, base + kNaRed, p.nT, W,
	                   base + kNaWsum, p.nWg, M, weighted ? 1 : 0, base + kNaOut);
	RPWA_HIP_CHECK(hipGetLastError());
	RPWA_HIP_CHECK(hipMemcpy(outSums, base + kNaOut, (size_t)W * W * 2 * sizeof(double), hipMemcpyDeviceToHost));
	return RPWA_OK;
}

}  // namespace rpwa_amd

extern "C" int rpwa_integral_device_noalloc(int device, const double* d_amps, uint64_t nmbEvents,
                                            uint32_t nmbWaves, const double* d_weights, double* outSums)
{
	rpwa_amd::clear_error();
	if (!d_amps || !outSums || nmbEvents == 0 || nmbWaves == 0) {
		rpwa_amd::set_error("rpwa_integral_device_noalloc: null pointer or empty input");
		return RPWA_ERR_INVALID;
	}
	try {
		return rpwa_amd::integral_noalloc(device, d_amps, nmbEvents, nmbWaves, d_weights, outSums);
	} catch (const std::exception& ex) {
		rpwa_amd::set_error(ex.what());
		return RPWA_ERR_INVALID;
	}
}

extern "C" int rpwa_data_moments_device(int device, const double* d_amps, uint64_t nmbEvents, uint32_t nmbWaves,
                                        const double* normDiag, int perEventNorm, double* outMoments)
{
	rpwa_amd::clear_error();
	if (!d_amps || !normDiag || !outMoments || nmbEvents == 0 || nmbWaves == 0) {
		rpwa_amd::set_error("rpwa_data_moments_device: null pointer or empty input");
		return RPWA_ERR_INVALID;
	}
	try {
		return rpwa_amd::moments_run(device, d_amps, nmbEvents, nmbWaves, normDiag, perEventNorm, outMoments);
	} catch (const std::exception& ex) {
		rpwa_amd::set_error(ex.what());
		return RPWA_ERR_NOMEM;
	}
}

extern "C" int rpwa_integral(int device, const double* amps, uint64_t nmbEvents, uint32_t nmbWaves,
                             const double* weights, double* outSums)
{
	rpwa_amd::clear_error();
	if (!amps || !outSums || nmbEvents == 0 || nmbWaves == 0) {
		rpwa_amd::set_error("rpwa_integral: null pointer or empty input");
		return RPWA_ERR_INVALID;
	}
	try {
		return rpwa_amd::integral_run(device, nullptr, nmbEvents, nmbWaves, nullptr, outSums, amps,
		                              weights);
	} catch (const std::exception& ex) {
		rpwa_amd::set_error(ex.what());
		return RPWA_ERR_NOMEM;
	}
}

extern "C" int rpwa_integral_device(int device, const double* d_amps, uint64_t nmbEvents,
                                    uint32_t nmbWaves, const double* d_weights, double* outSums)
{
	rpwa_amd::clear_error();
	if (!d_amps || !outSums || nmbEvents == 0 || nmbWaves == 0) {
		rpwa_amd::set_error("rpwa_integral_device: null pointer or empty input");
		return RPWA_ERR_INVALID;
	}
	try {
		return rpwa_amd::integral_run(device, d_amps, nmbEvents, nmbWaves, d_weights, outSums,
		                              nullptr, nullptr);
	} catch (const std::exception& ex) {
		rpwa_amd::set_error(ex.what());
		return RPWA_ERR_NOMEM;
	}
}
