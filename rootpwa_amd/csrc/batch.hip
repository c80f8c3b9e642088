// batch.hip -- batched fits over many bins and starts (include/rootpwa_amd_batch.h).
//
// One evaluation of all fits is four kernels, captured once in a hipGraph and
// replayed for every L-BFGS iteration:
//   batch_event_kernel   persistent waves; each wave takes work units (bin,
//                        chunk of 32 tiles of 32 events, group of 16 starts)
//                        and computes, tile by tile, S = A T (FP32 MFMA),
//                        sigma, ln sigma (FP64) and G += A^H (S / sigma) (FP32
//                        MFMA), writing one partial per unit;
//   batch_reduce_kernel  per fit: the chunk partials in fixed order (FP64),
//                        the norm term T^H I T and its derivative (FP64),
//                        -lnL and gradient;
//   batch_lbfgs_kernel   per fit: Armijo line search, L-BFGS two-loop
//                        direction, stopping tests, next trial point;
//   batch_plan_kernel    per bin: the running starts packed into the leading
//                        columns, and the list of work units of the next pass
//                        (only column groups holding a running start).
// A partial depends only on its chunk and on the start's data, never on the
// wave that computed it or on which column the start occupies, and the chunk
// partials are summed in chunk order: results do not depend on timing, on the
// grid, or on how many other fits are still running.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "rootpwa_amd_batch.h"
#include "rpwa/fitDrivers.h"

namespace rpwa_amd {
namespace {

constexpr int kCols = 64;        // starts per bin: 4 column groups
constexpr int kGroup = 16;       // starts per column group (one wave: 32 MFMA columns)
constexpr int kEv = 32;          // events per tile (MFMA 32x32 rows)
constexpr int kChunkTiles = 32;  // tiles per work unit (1024 events): short units keep a batched fit's tail short
constexpr int kThreads = 256;    // 4 independent waves per block
constexpr int kMaxMem = 16;      // L-BFGS correction pairs
constexpr int kTStride = kGroup + 1;  // LDS row of one wave's production amplitudes (float2)

typedef float f32x16 __attribute__((ext_vector_type(16)));
// global-address-space views, so amplitude loads are global (not flat) loads
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f32x4 gfloat4;
typedef __attribute__((address_space(1))) const float gfloat;

struct BatchEventArgs {
	const float* const* binAmps;  // per bin: [tiles * 32][rowLen] float (complex64 pairs)
	const uint32_t* chunkBin;     // [nChunks]
	const uint32_t* chunkTile0;   // [nChunks] first tile of the chunk within its bin
	const uint32_t* nEvents;      // [nBins]
	const uint32_t* units;        // [nChunks * 4] chunk * 4 + column group, this pass
	const uint32_t* nUnits;       // [1]
	const int32_t* colMap;        // [nBins][64] start held by each column (-1: none)
	const float* T32;             // [nBins][64][rowLen] by start
	const double* flat2;          // [nBins][64] by start
	float* partG;                 // [nChunks * 4][16][Wp][2]: per unit, start slot-major
	double* partLn;               // [nChunks * 4][16]
	double* partInv;              // [nChunks * 4][16]
};

__device__ __forceinline__ float swap_pair(float v)
{
	// lanes 2k <-> 2k+1 (DPP quad_perm [1,0,3,2])
	return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

__device__ __forceinline__ double swap_pair_d(double v)
{
	const uint64_t u = (uint64_t)__double_as_longlong(v);
	const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0xB1, 0xF, 0xF, true);
	const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0xB1, 0xF, 0xF, true);
	return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// One wave per work unit.  Lane (h, j): MFMA column j = (start slot j / 2 of
// the group, Re / Im of S by j % 2); MFMA rows are events (S = A T) or wave
// components (G = A^T V).  KQ = Wp / 4.
template <int KQ>
__global__ __launch_bounds__(kThreads, 2) void batch_event_kernel(const BatchEventArgs a)
{
	constexpr int NT = (KQ + 3) / 4;            // tiles of 32 wave components
	constexpr uint32_t rowLen = 8 * KQ;         // 2 Wp floats per event
	constexpr uint32_t Wp = 4 * KQ;
	constexpr int lastCols = (int)rowLen - 32 * (NT - 1);  // valid columns of the last component tile

	__shared__ float2 Tlds[kThreads / 64][Wp * kTStride];
	const int lane = threadIdx.x & 63;
	const int h = lane >> 5, j = lane & 31;
	const int cout = j & 1;                     // this lane's column: Re (0) or Im (1) of S
	float2* Tw = Tlds[threadIdx.x >> 6];
	const uint32_t nWaves = gridDim.x * (kThreads / 64);
	const uint32_t nUnits = *a.nUnits;

	for (uint32_t u = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); u < nUnits; u += nWaves) {
		const uint32_t unit = a.units[u];
		const uint32_t cg = unit >> 2, q = unit & 3;
		const uint32_t b = a.chunkBin[cg];
		const uint32_t nEv = a.nEvents[b];
		const uint32_t tile0 = a.chunkTile0[cg];
		const uint32_t tile1 = min(tile0 + (uint32_t)kChunkTiles, (nEv + kEv - 1) / kEv);
		const int start = a.colMap[b * kCols + kGroup * q + (j >> 1)];

		// production amplitudes of the group's 16 starts in this wave's LDS
		// region, [wave i][start slot] (Re, Im); padding columns are zero
		{
			const int* cm = a.colMap + b * kCols + kGroup * q;
			for (int idx = lane; idx < (int)Wp * kGroup; idx += 64) {
				const int i = idx / kGroup, sl = idx % kGroup;
				const int st = cm[sl];
				float2 v = make_float2(0.f, 0.f);
				if (st >= 0) {
					const float* Tb = a.T32 + ((size_t)b * kCols + st) * rowLen;
					v = make_float2(Tb[2 * i], Tb[2 * i + 1]);
				}
				Tw[i * kTStride + sl] = v;
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
		const double f2 = start >= 0 ? a.flat2[(size_t)b * kCols + start] : 1.;
		const float2* Tj = Tw + (j >> 1);

		f32x16 X[NT];
#pragma unroll
		for (int x = 0; x < NT; ++x)
			X[x] = f32x16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
		double mant = 1.;  // running product of sigma, renormalised per tile
		int expo = 0;
		double invAcc = 0.;
		const float* amps = a.binAmps[b];

		// Operands are loaded kPd1 steps (GEMM1) and kPd2 steps (GEMM2) ahead; the
		// scheduling barriers keep the compiler from hoisting every load of a
		// tile (register pressure).
		constexpr int kPd2 = 2;
		constexpr int kPd1 = KQ < 4 ? KQ : 4;
		for (uint32_t t = tile0; t < tile1; ++t) {
			// S[32 events][32 columns] = A_tile T
			const gfloat4* rowA = (const gfloat4*)(amps + ((size_t)t * kEv + j) * rowLen);
			f32x16 S = f32x16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
			f32x4 avBuf[kPd1];
#pragma unroll
			for (int qq = 0; qq < kPd1; ++qq)
				avBuf[qq] = rowA[2 * qq + h];
			float2 t0n = Tj[(2 * h) * kTStride], t1n = Tj[(2 * h + 1) * kTStride];
#pragma unroll
			for (int q4 = 0; q4 < KQ; ++q4) {
				// K steps 4 q4 .. 4 q4 + 3 of lane half h are the wave components
				// 8 q4 + 4 h + 0..3, i.e. (Re, Im) of waves i0 = 4 q4 + 2 h and i0 + 1;
				// [Sre Sim] = [Are Aim] [[Tre Tim] [-Tim Tre]]
				const f32x4 av = avBuf[q4 % kPd1];
				if (q4 + kPd1 < KQ)
					avBuf[q4 % kPd1] = rowA[2 * (q4 + kPd1) + h];
				const float2 t0 = t0n, t1 = t1n;
				if (q4 + 1 < KQ) {
					t0n = Tj[(4 * (q4 + 1) + 2 * h) * kTStride];
					t1n = Tj[(4 * (q4 + 1) + 2 * h + 1) * kTStride];
				}
				S = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, cout ? t0.y : t0.x, S, 0, 0, 0);
				S = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, cout ? t0.x : -t0.y, S, 0, 0, 0);
				S = __builtin_amdgcn_mfma_f32_32x32x2f32(av.z, cout ? t1.y : t1.x, S, 0, 0, 0);
				S = __builtin_amdgcn_mfma_f32_32x32x2f32(av.w, cout ? t1.x : -t1.y, S, 0, 0, 0);
				__builtin_amdgcn_sched_barrier(0);
			}
			// per register p (event 8 (p / 4) + 4 h + p % 4 of the tile, column j):
			// sigma = |S|^2 + flat^2 (FP64), ln sigma via the running product,
			// V = S / sigma; then K step p of X[wave components][columns] +=
			// A_tile^T V, which pairs register p of both lane halves.  Columns of
			// the last component tile past 2 Wp read within the row's tile and are
			// zeroed.
			const uint32_t evBase = t * kEv;
			const gfloat* tileB = (const gfloat*)(amps + ((size_t)t * kEv + 4 * h) * rowLen);
			float a2Buf[kPd2][NT];
#pragma unroll
			for (int pp = 0; pp < kPd2; ++pp)
#pragma unroll
				for (int x = 0; x < NT; ++x)
					a2Buf[pp][x] = tileB[(8 * (pp >> 2) + (pp & 3)) * rowLen + 32 * x + (x < NT - 1 ? j : min(j, lastCols - 1))];
#pragma unroll
			for (int p = 0; p < 16; ++p) {
				float a2[NT];
#pragma unroll
				for (int x = 0; x < NT; ++x)
					a2[x] = (x < NT - 1 || lastCols >= 32 || j < lastCols) ? a2Buf[p % kPd2][x] : 0.f;
				if (p + kPd2 < 16) {
					const int pn = p + kPd2;
					const gfloat* rowB = tileB + (8 * (pn >> 2) + (pn & 3)) * rowLen;
#pragma unroll
					for (int x = 0; x < NT; ++x)
						a2Buf[p % kPd2][x] = rowB[32 * x + (x < NT - 1 ? j : min(j, lastCols - 1))];
				}
				const int row = 8 * (p >> 2) + 4 * h + (p & 3);
				const bool valid = evBase + (uint32_t)row < nEv;
				const float me = S[p];
				const float other = swap_pair(me);
				const float re = cout ? other : me, im = cout ? me : other;
				double sig = fma((double)re, (double)re, fma((double)im, (double)im, f2));
				sig = valid ? sig : 1.;
				const float inv = __builtin_amdgcn_rcpf((float)sig);
				const float vp = valid ? me * inv : 0.f;
				if ((p & 1) == cout) {  // the pair splits the event terms
					mant *= sig;
					invAcc += valid ? (double)inv : 0.;
				}
#pragma unroll
				for (int x = 0; x < NT; ++x)
					X[x] = __builtin_amdgcn_mfma_f32_32x32x2f32(a2[x], vp, X[x], 0, 0, 0);
				__builtin_amdgcn_sched_barrier(0);
			}
			{
				// mant = m 2^e with m in [1, 2): keep m, carry e (sigma > 0 is finite or inf/nan)
				const uint64_t ub = (uint64_t)__double_as_longlong(mant);
				const int e = (int)((ub >> 52) & 0x7FF);
				if (e != 0 && e != 0x7FF) {
					expo += e - 1023;
					mant = __longlong_as_double((long long)((ub & 0x800FFFFFFFFFFFFFull) | (1023ull << 52)));
				}
			}
		}

		// partial of this unit: G_i = sum conj(A_i) V per start slot.  Rows 2i,
		// 2i+1 of X are registers v, v+1; columns 2s, 2s+1 are lanes j, j^1:
		// Re G = X[2i][2s] + X[2i+1][2s+1], Im G = X[2i][2s+1] - X[2i+1][2s]
		const uint32_t slot = cg * 4 + q;
		float* pg = a.partG + (size_t)slot * Wp * (2 * kGroup);
#pragma unroll
		for (int x = 0; x < NT; ++x)
#pragma unroll
			for (int v = 0; v < 16; v += 2) {
				const int row0 = 32 * x + 8 * (v >> 2) + 4 * h + (v & 3);
				const float own0 = X[x][v], own1 = X[x][v + 1];
				const float p0 = swap_pair(own0), p1 = swap_pair(own1);
				if (cout == 0 && row0 < (int)rowLen) {
					const int i = row0 >> 1;
					pg[((j >> 1) * Wp + i) * 2 + 0] = own0 + p1;
					pg[((j >> 1) * Wp + i) * 2 + 1] = p0 - own1;
				}
			}
		double lnv = log(mant) + (double)expo * 0.69314718055994530942;
		lnv += swap_pair_d(lnv);
		lnv += __shfl_xor(lnv, 32);
		double inv = invAcc + swap_pair_d(invAcc);
		inv += __shfl_xor(inv, 32);
		if (h == 0 && cout == 0) {
			a.partLn[(size_t)slot * kGroup + (j >> 1)] = lnv;
			a.partInv[(size_t)slot * kGroup + (j >> 1)] = inv;
		}
	}
}

typedef void (*BatchEventKernel)(const BatchEventArgs);

// KQ = Wp / 4, Wp = waves padded to a multiple of 8
BatchEventKernel pick_event_kernel(int KQ)
{
	switch (KQ) {
	case 2: return batch_event_kernel<2>;
	case 4: return batch_event_kernel<4>;
	case 6: return batch_event_kernel<6>;
	case 8: return batch_event_kernel<8>;
	case 10: return batch_event_kernel<10>;
	case 12: return batch_event_kernel<12>;
	case 14: return batch_event_kernel<14>;
	case 16: return batch_event_kernel<16>;
	case 18: return batch_event_kernel<18>;
	case 20: return batch_event_kernel<20>;
	case 22: return batch_event_kernel<22>;
	default: return batch_event_kernel<24>;
	}
}

// minimiser state of all fits (FP64), fit f = bin * 64 + start
struct FitState {
	double* x;      // [nFits][P] accepted point
	double* g;      // gradient at x
	double* d;      // search direction
	double* xt;     // trial point (evaluated next)
	double* gt;     // gradient at xt (from the reduction)
	double* S;      // [nFits][m][P]
	double* Y;
	double* rho;    // [nFits][m]
	double* f;      // value at x
	double* ft;     // value at xt
	double* alpha;
	double* dphi0;
	double* gamma;
	int32_t* status;
	int32_t* nEval;
	int32_t* backtracks;
	int32_t* mCount;
	int32_t* mHead;
};

// status codes of a batched fit (> 0 finished, as NLopt's result codes)
constexpr int ST_INIT = 0;
constexpr int ST_SEARCH = 1;
constexpr int ST_FTOL = 3;
constexpr int ST_XTOL = 4;
constexpr int ST_MAXEVAL = 5;
constexpr int ST_NOPROGRESS = -4;
constexpr int ST_BADSTART = -1;
constexpr int ST_PAD = -100;  // padding fit (start >= nStarts): never runs

__device__ __forceinline__ bool running(int st) { return st == ST_INIT || st == ST_SEARCH; }

struct BatchReduceArgs {
	const float* partG;
	const double* partLn;
	const double* partInv;
	const uint32_t* binChunk0;  // [nBins + 1]
	const int32_t* startCol;    // [nBins][64] column of each start this pass (-1: not evaluated)
	const float* T32;
	const double* acc;          // [nBins][W][W] complex, normalised
	const double* totAcc;       // [nBins]
	const double* xt;
	double* ft;
	double* gt;
	uint32_t W, Wp, P, rowLen;
};

// per fit: gradient partials in chunk order, norm term, value and gradient.
// kRedParts x 128 threads: thread (h, i) adds part h of the bin's chunks for
// wave component i (the parts are then added in order 0, 1, ...), and the norm
// term's row sums read T from LDS with the loads of eight columns issued
// together.
constexpr int kRedParts = 4;
constexpr int kRedHalf = 128;
constexpr int kRedThreads = kRedParts * kRedHalf;
static_assert(RPWA_BATCH_MAX_WAVES <= kRedHalf, "one thread per wave component and part");

__global__ __launch_bounds__(kRedThreads) void batch_reduce_kernel(const BatchReduceArgs r)
{
	const uint32_t b = blockIdx.x / kCols;
	const int col = r.startCol[blockIdx.x];
	if (col < 0)
		return;
	const uint32_t q = (uint32_t)col / kGroup, jc = (uint32_t)col % kGroup;
	const uint32_t c0 = r.binChunk0[b], c1 = r.binChunk0[b + 1];
	const size_t fit = blockIdx.x;
	const uint32_t tid = threadIdx.x;
	const uint32_t i = tid % kRedHalf, h = tid / kRedHalf;
	__shared__ double red[3][kRedThreads];
	__shared__ double2 gh[kRedParts - 1][kRedHalf];   // the later parts' gradient sums
	__shared__ double2 Ts[kRedHalf];   // T of the fit
	const float* Tf = r.T32 + fit * r.rowLen;
	if (tid < r.W)
		Ts[tid] = make_double2((double)Tf[2 * tid], (double)Tf[2 * tid + 1]);
	double gre = 0., gim = 0.;
	if (i < r.W) {
		// chunk partials of wave component i, added in chunk order; the loads of
		// eight chunks are issued before their adds
		const size_t cstride = (size_t)4 * r.Wp * 2 * kGroup;  // floats between consecutive chunks
		const uint32_t nc = c1 - c0;
		const uint32_t ca = c0 + (uint32_t)(((uint64_t)nc * h) / kRedParts);
		const uint32_t cb = c0 + (uint32_t)(((uint64_t)nc * (h + 1)) / kRedParts);
		// fit-major partials: a fit's wave components are contiguous, so its reads
		// use whole cache lines (with [wave][start] they took 8 bytes of each)
		const float* pg = r.partG + (size_t)(ca * 4 + q) * r.Wp * (2 * kGroup) + ((size_t)jc * r.Wp + i) * 2;
		uint32_t c = ca;
		for (; c + 8 <= cb; c += 8, pg += 8 * cstride) {
			float2 v[8];
#pragma unroll
			for (int k = 0; k < 8; ++k)
				v[k] = *reinterpret_cast<const float2*>(pg + k * cstride);
#pragma unroll
			for (int k = 0; k < 8; ++k) {
				gre += (double)v[k].x;
				gim += (double)v[k].y;
			}
		}
		for (; c < cb; ++c, pg += cstride) {
			gre += (double)pg[0];
			gim += (double)pg[1];
		}
		if (h)
			gh[h - 1][i] = make_double2(gre, gim);
	}
	// the per-chunk log sums and weight sums: thread t takes chunks c0+t, c0+t+kRedThreads, ...
	double lnPart = 0., invPart = 0.;
	for (uint32_t c = c0 + tid; c < c1; c += kRedThreads) {
		lnPart += r.partLn[(size_t)(c * 4 + q) * kGroup + jc];
		invPart += r.partInv[(size_t)(c * 4 + q) * kGroup + jc];
	}
	__syncthreads();
	double normPart = 0.;
	if (h == 0 && i < r.W) {
#pragma unroll
		for (int part = 0; part < kRedParts - 1; ++part) {
			gre += gh[part][i].x;
			gim += gh[part][i].y;
		}
		// u_i = sum_j T_j conj(I_ij); norm = sum_i Re(T_i conj(u_i)); d norm / d T_i = 2 u_i
		const double2* I = reinterpret_cast<const double2*>(r.acc) + ((size_t)b * r.W + i) * r.W;
		double ure = 0., uim = 0.;
		uint32_t jw = 0;
		for (; jw + 8 <= r.W; jw += 8) {
			double2 iv[8];
#pragma unroll
			for (int k = 0; k < 8; ++k)
				iv[k] = I[jw + k];
#pragma unroll
			for (int k = 0; k < 8; ++k) {
				const double2 t = Ts[jw + k];
				ure += t.x * iv[k].x + t.y * iv[k].y;
				uim += t.y * iv[k].x - t.x * iv[k].y;
			}
		}
		for (; jw < r.W; ++jw) {
			const double2 t = Ts[jw], iv = I[jw];
			ure += t.x * iv.x + t.y * iv.y;
			uim += t.y * iv.x - t.x * iv.y;
		}
		const double2 ti = Ts[i];
		normPart = ti.x * ure + ti.y * uim;
		const double dre = -2. * gre + 2. * ure, dim = -2. * gim + 2. * uim;
		double* gout = r.gt + fit * r.P;
		if (i == 0)
			gout[0] = dre;
		else {
			gout[2 * i - 1] = dre;
			gout[2 * i] = dim;
		}
	}
	red[0][tid] = normPart;
	red[1][tid] = lnPart;
	red[2][tid] = invPart;
	__syncthreads();
	for (uint32_t w = kRedThreads / 2; w > 0; w >>= 1) {  // fixed-order tree over the threads
		if (tid < w) {
			red[0][tid] += red[0][tid + w];
			red[1][tid] += red[1][tid + w];
			red[2][tid] += red[2][tid + w];
		}
		__syncthreads();
	}
	if (tid == 0) {
		const double lnSum = red[1][0], invSum = red[2][0];
		const double flat = r.xt[fit * r.P + r.P - 1];
		const double eta = r.totAcc[b];
		r.ft[fit] = -lnSum + (red[0][0] + flat * flat * eta);
		r.gt[fit * r.P + r.P - 1] = -2. * flat * invSum + 2. * flat * eta;
	}
}

// trial point -> FP32 production amplitudes and flat^2 for the event kernel
__device__ void write_trial(const double* xt, uint32_t W, uint32_t P, float* T32, double* flat2)
{
	for (uint32_t i = threadIdx.x; i < W; i += blockDim.x) {
		const double re = i == 0 ? xt[0] : xt[2 * i - 1];
		const double im = i == 0 ? 0. : xt[2 * i];
		T32[2 * i] = (float)re;
		T32[2 * i + 1] = (float)im;
	}
	if (threadIdx.x == 0) {
		const double f = xt[P - 1];
		*flat2 = f * f;
	}
}

struct BatchPrepArgs {
	const double* xt;
	float* T32;
	double* flat2;
	uint32_t W, P, rowLen, nStarts;
};

// evaluation at given points (no minimiser step)
__global__ __launch_bounds__(64) void batch_prepare_kernel(const BatchPrepArgs p)
{
	const uint32_t fit = blockIdx.x, s = fit % kCols;
	if (s >= p.nStarts)
		return;
	write_trial(p.xt + (size_t)fit * p.P, p.W, p.P, p.T32 + (size_t)fit * p.rowLen, p.flat2 + fit);
}

struct BatchLbfgsArgs {
	FitState st;
	float* T32;
	double* flat2;
	uint32_t W, P, rowLen, m, maxEval;
	double xtolRel, ftolAbs;
};

__device__ __forceinline__ double wave_sum(double v)
{
	for (int k = 32; k >= 1; k >>= 1)
		v += __shfl_xor(v, k);
	return v;
}

// one wave per fit; each lane holds parameters lane, lane + 64, lane + 128
__global__ __launch_bounds__(64) void batch_lbfgs_kernel(const BatchLbfgsArgs a)
{
	const size_t fit = blockIdx.x;
	const FitState& st = a.st;
	const int status = st.status[fit];
	if (!running(status))
		return;
	const int lane = threadIdx.x;
	const uint32_t P = a.P, m = a.m;
	double* x = st.x + fit * P;
	double* g = st.g + fit * P;
	double* d = st.d + fit * P;
	double* xt = st.xt + fit * P;
	const double* gt = st.gt + fit * P;
	double* Sm = st.S + fit * m * P;
	double* Ym = st.Y + fit * m * P;
	double* rho = st.rho + fit * m;
	constexpr int kPer = 3;  // P <= 192
	const double ft = st.ft[fit];
	int nEval = st.nEval[fit] + 1;
	int newStatus = status;
	double alpha = st.alpha[fit];
	bool newDirection = false;

	if (status == ST_INIT) {
		if (!isfinite(ft)) {
			newStatus = ST_BADSTART;
		} else {
			for (int q = 0; q < kPer; ++q) {
				const uint32_t k = lane + 64 * q;
				if (k < P) {
					x[k] = xt[k];
					g[k] = gt[k];
				}
			}
			if (lane == 0)
				st.f[fit] = ft;
			newStatus = ST_SEARCH;
			newDirection = true;
		}
	} else {
		const double f = st.f[fit];
		const double dphi0 = st.dphi0[fit];
		if (isfinite(ft) && ft <= f + 1e-4 * alpha * dphi0) {
			// accept the trial point
			bool xConv = a.xtolRel > 0.;
			double sy = 0., ss = 0., yy = 0.;
			const int head = st.mHead[fit];
			double* sNew = Sm + (size_t)head * P;
			double* yNew = Ym + (size_t)head * P;
			for (int q = 0; q < kPer; ++q) {
				const uint32_t k = lane + 64 * q;
				if (k < P) {
					const double sk = xt[k] - x[k], yk = gt[k] - g[k];
					if (!(fabs(sk) < a.xtolRel * 0.5 * (fabs(xt[k]) + fabs(x[k])) || xt[k] == x[k]))
						xConv = false;
					sNew[k] = sk;
					yNew[k] = yk;
					sy += sk * yk;
					ss += sk * sk;
					yy += yk * yk;
					x[k] = xt[k];
					g[k] = gt[k];
				}
			}
			sy = wave_sum(sy);
			ss = wave_sum(ss);
			yy = wave_sum(yy);
			xConv = __all(xConv);
			if (lane == 0)
				st.f[fit] = ft;
			st.backtracks[fit] = 0;
			if (a.ftolAbs > 0. && fabs(ft - f) < a.ftolAbs)
				newStatus = ST_FTOL;
			else if (xConv)
				newStatus = ST_XTOL;
			else {
				if (sy > 1e-12 * sqrt(ss * yy)) {
					if (lane == 0) {
						rho[head] = 1. / sy;
						st.gamma[fit] = sy / yy;
						st.mHead[fit] = (head + 1) % (int)m;
						st.mCount[fit] = min(st.mCount[fit] + 1, (int)m);
					}
				}
				newDirection = true;
			}
		} else {
			const int bt = st.backtracks[fit] + 1;
			if (bt > 60) {
				newStatus = ST_NOPROGRESS;
			} else {
				double an = isfinite(ft) ? -dphi0 * alpha * alpha / (2. * (ft - f - dphi0 * alpha)) : 0.5 * alpha;
				if (!isfinite(an) || an < 0.1 * alpha || an > 0.5 * alpha)
					an = (an < 0.1 * alpha || !isfinite(an)) ? 0.1 * alpha : 0.5 * alpha;
				alpha = an;
				if (lane == 0)
					st.backtracks[fit] = bt;
				if (alpha == 0.)
					newStatus = ST_NOPROGRESS;
			}
		}
	}
	__syncthreads();  // history scalars written by lane 0 above

	if (newDirection) {
		// d = -H g (two-loop recursion, newest pair first)
		const int count = st.mCount[fit], head = st.mHead[fit];
		const double gamma = st.gamma[fit];
		double q[kPer];
		for (int r = 0; r < kPer; ++r) {
			const uint32_t k = lane + 64 * r;
			q[r] = k < P ? g[k] : 0.;
		}
		// the stored pairs in registers, loaded together (the recursion below is a
		// chain of dependent dot products; no load waits inside it)
		double sreg[kMaxMem][kPer], yreg[kMaxMem][kPer], rreg[kMaxMem];
#pragma unroll
		for (int c = 0; c < kMaxMem; ++c) {
			const int slot = (head - 1 - c + 2 * (int)m) % (int)m;
			const bool use = c < count;
			rreg[c] = use ? rho[slot] : 0.;
#pragma unroll
			for (int r = 0; r < kPer; ++r) {
				const uint32_t k = lane + 64 * r;
				const bool in = use && k < P;
				sreg[c][r] = in ? Sm[(size_t)slot * P + k] : 0.;
				yreg[c][r] = in ? Ym[(size_t)slot * P + k] : 0.;
			}
		}
		double al[kMaxMem];
#pragma unroll
		for (int c = 0; c < kMaxMem; ++c) {
			if (c < count) {
				double s = 0.;
				for (int r = 0; r < kPer; ++r) {
					const uint32_t k = lane + 64 * r;
					if (k < P)
						s += sreg[c][r] * q[r];
				}
				al[c] = rreg[c] * wave_sum(s);
				for (int r = 0; r < kPer; ++r) {
					const uint32_t k = lane + 64 * r;
					if (k < P)
						q[r] -= al[c] * yreg[c][r];
				}
			}
		}
		if (count > 0)
			for (int r = 0; r < kPer; ++r)
				q[r] *= gamma;
#pragma unroll
		for (int c = kMaxMem - 1; c >= 0; --c) {
			if (c < count) {
				double s = 0.;
				for (int r = 0; r < kPer; ++r) {
					const uint32_t k = lane + 64 * r;
					if (k < P)
						s += yreg[c][r] * q[r];
				}
				const double beta = rreg[c] * wave_sum(s);
				for (int r = 0; r < kPer; ++r) {
					const uint32_t k = lane + 64 * r;
					if (k < P)
						q[r] += sreg[c][r] * (al[c] - beta);
				}
			}
		}
		double gd = 0., gg = 0.;
		for (int r = 0; r < kPer; ++r) {
			const uint32_t k = lane + 64 * r;
			if (k < P) {
				gd -= g[k] * q[r];
				gg += g[k] * g[k];
			}
		}
		gd = wave_sum(gd);
		gg = wave_sum(gg);
		if (count == 0 || !(gd < 0.)) {
			// steepest descent with a step of unit length; history restarts
			for (int r = 0; r < kPer; ++r) {
				const uint32_t k = lane + 64 * r;
				if (k < P)
					d[k] = -g[k];
			}
			if (lane == 0 && count > 0) {
				st.mCount[fit] = 0;
				st.mHead[fit] = 0;
			}
			gd = -gg;
			alpha = gg > 0. ? 1. / sqrt(gg) : 0.;
		} else {
			for (int r = 0; r < kPer; ++r) {
				const uint32_t k = lane + 64 * r;
				if (k < P)
					d[k] = -q[r];
			}
			alpha = 1.;
		}
		if (lane == 0)
			st.dphi0[fit] = gd;
		if (gg == 0.)
			newStatus = ST_FTOL;  // stationary point
	}
	if (newStatus == ST_SEARCH && nEval >= (int)a.maxEval)
		newStatus = ST_MAXEVAL;
	if (lane == 0) {
		st.status[fit] = newStatus;
		st.nEval[fit] = nEval;
		st.alpha[fit] = alpha;
	}
	if (newStatus == ST_SEARCH) {
		for (int r = 0; r < kPer; ++r) {
			const uint32_t k = lane + 64 * r;
			if (k < P)
				xt[k] = x[k] + alpha * d[k];
		}
		__syncthreads();
		write_trial(xt, a.W, P, a.T32 + fit * a.rowLen, a.flat2 + fit);
	}
}

struct BatchPlanArgs {
	const int32_t* status;      // [nBins][64]
	const uint32_t* binChunk0;  // [nBins + 1]
	int32_t* colMap;            // [nBins][64]
	int32_t* startCol;          // [nBins][64]
	uint32_t* unitCount;        // [nBins] scratch
	uint32_t* units;            // [nChunks * 4]
	uint32_t* nUnits;
	int32_t* nRunning;
	uint32_t nBins, nStarts;
	int evalAll;                // 1: every start < nStarts is evaluated (no minimiser state)
};

// one block: per bin the evaluated starts packed into the leading columns (in
// start order), the column groups that hold them, and the work units of the
// next pass in (bin, chunk, group) order; the number of running fits.  A wave
// handles a bin with one lane per start (the packing is a ballot and a
// popcount), the unit offsets are a block-wide scan over the bins.
__global__ __launch_bounds__(256) void batch_plan_kernel(const BatchPlanArgs a)
{
	constexpr int kPlanWaves = 256 / 64;
	__shared__ int32_t red[256];
	__shared__ uint32_t scan[256];
	const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	int32_t count = 0;
	for (uint32_t b = wv; b < a.nBins; b += kPlanWaves) {
		const uint32_t s = lane;
		const bool run = a.evalAll ? s < a.nStarts : running(a.status[(size_t)b * kCols + s]);
		const uint64_t mask = __ballot(run);
		const int k = __popcll(mask & ((1ull << lane) - 1ull));  // runs among the starts before s
		const int nRun = __popcll(mask);
		a.startCol[(size_t)b * kCols + s] = run ? k : -1;
		if (run)
			a.colMap[(size_t)b * kCols + k] = (int32_t)s;
		if ((int)lane >= nRun)
			a.colMap[(size_t)b * kCols + lane] = -1;
		if (lane == 0) {
			const uint32_t groups = (uint32_t)(nRun + kGroup - 1) / kGroup;
			a.unitCount[b] = groups * (a.binChunk0[b + 1] - a.binChunk0[b]);
			count += a.evalAll ? 0 : nRun;
		}
	}
	red[threadIdx.x] = count;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w)
			red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		*a.nRunning = red[0];
	// exclusive scan of the unit counts over the bins, 256 bins at a time
	uint32_t carry = 0;
	for (uint32_t b0 = 0; b0 < a.nBins; b0 += 256) {
		const uint32_t b = b0 + threadIdx.x;
		const uint32_t n = b < a.nBins ? a.unitCount[b] : 0u;
		scan[threadIdx.x] = n;
		__syncthreads();
		for (uint32_t d = 1; d < 256; d <<= 1) {  // inclusive Hillis-Steele scan
			const uint32_t v = threadIdx.x >= d ? scan[threadIdx.x - d] : 0u;
			__syncthreads();
			scan[threadIdx.x] += v;
			__syncthreads();
		}
		if (b < a.nBins)
			a.unitCount[b] = carry + scan[threadIdx.x] - n;  // offset of bin b
		carry += scan[255];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		*a.nUnits = carry;
	__syncthreads();
	// the units of bin b (a wave per bin): chunk-major, column group minor
	for (uint32_t b = wv; b < a.nBins; b += kPlanWaves) {
		const uint32_t off = a.unitCount[b];
		const uint32_t n = (b + 1 < a.nBins ? a.unitCount[b + 1] : carry) - off;
		const uint32_t chunks = a.binChunk0[b + 1] - a.binChunk0[b];
		const uint32_t groups = chunks ? n / chunks : 0;
		for (uint32_t k = lane; k < n; k += 64) {
			const uint32_t c = k / groups, g = k % groups;
			a.units[off + k] = (a.binChunk0[b] + c) * 4 + g;
		}
	}
}

struct BatchInitArgs {
	FitState st;
	float* T32;
	double* flat2;
	uint32_t W, P, rowLen, nStarts;
};

// fit state at the start values xt (set by the host)
__global__ __launch_bounds__(64) void batch_init_kernel(const BatchInitArgs a)
{
	const size_t fit = blockIdx.x;
	const uint32_t s = fit % kCols;
	if (threadIdx.x == 0) {
		a.st.status[fit] = s < a.nStarts ? ST_INIT : ST_PAD;
		a.st.nEval[fit] = 0;
		a.st.backtracks[fit] = 0;
		a.st.mCount[fit] = 0;
		a.st.mHead[fit] = 0;
		a.st.alpha[fit] = 0.;
		a.st.gamma[fit] = 1.;
		a.st.f[fit] = 0.;
	}
	if (s < a.nStarts)
		write_trial(a.st.xt + fit * a.P, a.W, a.P, a.T32 + fit * a.rowLen, a.flat2 + fit);
}

// complex128 [N][W] -> scaled complex64 [tiles * 32][rowLen] (zero padding)
__global__ void batch_transform_kernel(const double2* __restrict__ src, uint64_t nEvents, uint32_t W,
                                       uint32_t Wp, uint64_t nRows, const double* __restrict__ scale,
                                       float2* __restrict__ dst)
{
	const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= nRows * Wp)
		return;
	const uint64_t e = idx / Wp;
	const uint32_t w = (uint32_t)(idx % Wp);
	float2 out = make_float2(0.f, 0.f);
	if (e < nEvents && w < W) {
		const double2 v = src[e * W + w];
		const double f = scale[w];
		out = make_float2((float)(v.x * f), (float)(v.y * f));
	}
	dst[idx] = out;
}

uint64_t splitmix64(uint64_t x)
{
	x += 0x9E3779B97F4A7C15ull;
	x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
	x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
	return x ^ (x >> 31);
}

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

struct rpwa_batch {
	int device = 0;
	uint32_t W = 0, Wp = 0, rowLen = 0, NT = 0, nStarts = 0, P = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	int nCu = 0;
	struct Bin {
		uint64_t nEvents = 0;
		uint32_t nTiles = 0;
		float* d_amps = nullptr;
		std::vector<double> acc;  // [W][W] complex, normalised
		double totAcc = 1.;
	};
	std::vector<Bin> bins;
	bool finished = false;
	uint32_t nTiles = 0, nChunks = 0, nWg = 0, wgPerCu = 0, wgPerCuFixed = 0, maxWgPerCu = 1;
	uint64_t totalEvents = 0;
	uint64_t lastEvaluations = 0;
	// device
	float** d_binAmps = nullptr;
	uint32_t* d_chunkBin = nullptr;
	uint32_t* d_chunkTile0 = nullptr;
	uint32_t* d_binChunk0 = nullptr;
	uint32_t* d_nEvents = nullptr;
	uint32_t* d_units = nullptr;
	uint32_t* d_nUnits = nullptr;
	uint32_t* d_unitCount = nullptr;
	int32_t* d_colMap = nullptr;
	int32_t* d_startCol = nullptr;
	float* d_T32 = nullptr;
	double* d_flat2 = nullptr;
	float* d_partG = nullptr;
	double* d_partLn = nullptr;
	double* d_partInv = nullptr;
	double* d_acc = nullptr;
	double* d_totAcc = nullptr;
	int32_t* d_nRunning = nullptr;
	int32_t* d_noStatus = nullptr;  // all zero: status array of the evaluation path
	FitState st{};
	uint32_t memory = 0;  // history depth the state was allocated for
	std::vector<void*> allocations;
	hipGraphExec_t graph = nullptr;
	uint32_t graphMem = 0, graphMaxEval = 0;
	double graphXtol = -1., graphFtol = -1.;
};

namespace {

int check_batch(const rpwa_batch* b, const char* fn, bool mustBeFinished)
{
	if (!b) {
		set_error(std::string(fn) + ": null batch");
		return RPWA_ERR_INVALID;
	}
	if (mustBeFinished && !b->finished) {
		set_error(std::string(fn) + ": rpwa_batch_finish has not been called");
		return RPWA_ERR_STATE;
	}
	return RPWA_OK;
}

template <typename T>
int dalloc(rpwa_batch* b, T** p, size_t n)
{
	void* q = nullptr;
	RPWA_HIP_CHECK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
	b->allocations.push_back(q);
	*p = reinterpret_cast<T*>(q);
	return RPWA_OK;
}

#define RPWA_TRY(expr)              \
	do {                            \
		const int rc_ = (expr);     \
		if (rc_ != RPWA_OK)         \
			return rc_;             \
	} while (0)

size_t nFits(const rpwa_batch* b) { return b->bins.size() * kCols; }

BatchEventArgs event_args(const rpwa_batch* b)
{
	BatchEventArgs a;
	a.binAmps = b->d_binAmps;
	a.chunkBin = b->d_chunkBin;
	a.chunkTile0 = b->d_chunkTile0;
	a.nEvents = b->d_nEvents;
	a.units = b->d_units;
	a.nUnits = b->d_nUnits;
	a.colMap = b->d_colMap;
	a.T32 = b->d_T32;
	a.flat2 = b->d_flat2;
	a.partG = b->d_partG;
	a.partLn = b->d_partLn;
	a.partInv = b->d_partInv;
	return a;
}

BatchReduceArgs reduce_args(const rpwa_batch* b)
{
	BatchReduceArgs r;
	r.partG = b->d_partG;
	r.partLn = b->d_partLn;
	r.partInv = b->d_partInv;
	r.binChunk0 = b->d_binChunk0;
	r.startCol = b->d_startCol;
	r.T32 = b->d_T32;
	r.acc = b->d_acc;
	r.totAcc = b->d_totAcc;
	r.xt = b->st.xt;
	r.ft = b->st.ft;
	r.gt = b->st.gt;
	r.W = b->W;
	r.Wp = b->Wp;
	r.P = b->P;
	r.rowLen = b->rowLen;
	return r;
}

int launch_event(rpwa_batch* b)
{
	hipLaunchKernelGGL(pick_event_kernel((int)(b->Wp / 4)), dim3(b->nWg), dim3(kThreads), 0, b->stream,
	                   event_args(b));
	RPWA_HIP_CHECK(hipGetLastError());
	return RPWA_OK;
}

int launch_reduce(rpwa_batch* b)
{
	hipLaunchKernelGGL(batch_reduce_kernel, dim3((uint32_t)nFits(b)), dim3(kRedThreads), 0, b->stream,
	                   reduce_args(b));
	RPWA_HIP_CHECK(hipGetLastError());
	return RPWA_OK;
}

// evalAll: plan for an evaluation of every start (no minimiser state)
int launch_plan(rpwa_batch* b, bool evalAll)
{
	BatchPlanArgs a;
	a.status = evalAll ? b->d_noStatus : b->st.status;
	a.binChunk0 = b->d_binChunk0;
	a.colMap = b->d_colMap;
	a.startCol = b->d_startCol;
	a.unitCount = b->d_unitCount;
	a.units = b->d_units;
	a.nUnits = b->d_nUnits;
	a.nRunning = b->d_nRunning;
	a.nBins = (uint32_t)b->bins.size();
	a.nStarts = b->nStarts;
	a.evalAll = evalAll ? 1 : 0;
	hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(256), 0, b->stream, a);
	RPWA_HIP_CHECK(hipGetLastError());
	return RPWA_OK;
}

int launch_lbfgs(rpwa_batch* b, uint32_t maxEval, double xtolRel, double ftolAbs)
{
	BatchLbfgsArgs a;
	a.st = b->st;
	a.T32 = b->d_T32;
	a.flat2 = b->d_flat2;
	a.W = b->W;
	a.P = b->P;
	a.rowLen = b->rowLen;
	a.m = b->memory;
	a.maxEval = maxEval;
	a.xtolRel = xtolRel;
	a.ftolAbs = ftolAbs;
	hipLaunchKernelGGL(batch_lbfgs_kernel, dim3((uint32_t)nFits(b)), dim3(64), 0, b->stream, a);
	RPWA_HIP_CHECK(hipGetLastError());
	return RPWA_OK;
}

int alloc_state(rpwa_batch* b, uint32_t memory)
{
	if (b->memory == memory && b->st.x)
		return RPWA_OK;
	const size_t F = nFits(b), P = b->P;
	FitState& s = b->st;
	RPWA_TRY(dalloc(b, &s.x, F * P));
	RPWA_TRY(dalloc(b, &s.g, F * P));
	RPWA_TRY(dalloc(b, &s.d, F * P));
	RPWA_TRY(dalloc(b, &s.S, F * memory * P));
	RPWA_TRY(dalloc(b, &s.Y, F * memory * P));
	RPWA_TRY(dalloc(b, &s.rho, F * memory));
	RPWA_TRY(dalloc(b, &s.f, F));
	RPWA_TRY(dalloc(b, &s.alpha, F));
	RPWA_TRY(dalloc(b, &s.dphi0, F));
	RPWA_TRY(dalloc(b, &s.gamma, F));
	RPWA_TRY(dalloc(b, &s.status, F));
	RPWA_TRY(dalloc(b, &s.nEval, F));
	RPWA_TRY(dalloc(b, &s.backtracks, F));
	RPWA_TRY(dalloc(b, &s.mCount, F));
	RPWA_TRY(dalloc(b, &s.mHead, F));
	b->memory = memory;
	if (b->graph) {
		(void)hipGraphExecDestroy(b->graph);
		b->graph = nullptr;
	}
	return RPWA_OK;
}

// padded [nBins * 64][P] from the caller's [nBins][nStarts][P]
void pad_params(const rpwa_batch* b, const double* par, std::vector<double>& out)
{
	const size_t P = b->P;
	out.assign(nFits(b) * P, 0.);
	for (size_t bi = 0; bi < b->bins.size(); ++bi)
		for (uint32_t s = 0; s < b->nStarts; ++s)
			std::memcpy(&out[(bi * kCols + s) * P], par + (bi * b->nStarts + s) * P, P * sizeof(double));
}

int time_event(rpwa_batch* b, uint32_t reps, double* ms)
{
	RPWA_HIP_CHECK(hipEventRecord(b->ev0, b->stream));
	for (uint32_t r = 0; r < reps; ++r)
		RPWA_TRY(launch_event(b));
	RPWA_HIP_CHECK(hipEventRecord(b->ev1, b->stream));
	RPWA_HIP_CHECK(hipEventSynchronize(b->ev1));
	float t = 0.f;
	RPWA_HIP_CHECK(hipEventElapsedTime(&t, b->ev0, b->ev1));
	*ms = (double)t / std::max<uint32_t>(reps, 1);
	return RPWA_OK;
}

void host_start_values(const rpwa_batch* b, uint32_t seed, std::vector<double>& out)
{
	const uint32_t P = b->P;
	out.assign(b->bins.size() * b->nStarts * (size_t)P, 0.);
	for (size_t bi = 0; bi < b->bins.size(); ++bi) {
		const double sqrtN = std::sqrt((double)b->bins[bi].nEvents);
		for (uint32_t s = 0; s < b->nStarts; ++s) {
			const uint32_t fitSeed =
				(uint32_t)splitmix64(((uint64_t)seed << 32) + ((uint64_t)bi << 16) + (uint64_t)s);
			rpwa::mt19937Uniform random(fitSeed);
			double* p = &out[(bi * b->nStarts + s) * (size_t)P];
			for (uint32_t k = 0; k < P; ++k) {
				double v = random.uniform(0.01, sqrtN);
				if (random.rndm() > 0.5)
					v = -v;
				p[k] = v;
			}
		}
	}
}

// trial points (padded [nBins * 64][P]) -> T32 / flat^2, plan of an evaluation of every start
int prepare_eval(rpwa_batch* b, const std::vector<double>& padded)
{
	RPWA_HIP_CHECK(hipMemcpyAsync(b->st.xt, padded.data(), padded.size() * sizeof(double), hipMemcpyHostToDevice,
	                              b->stream));
	BatchPrepArgs pa{b->st.xt, b->d_T32, b->d_flat2, b->W, b->P, b->rowLen, b->nStarts};
	hipLaunchKernelGGL(batch_prepare_kernel, dim3((uint32_t)nFits(b)), dim3(64), 0, b->stream, pa);
	RPWA_HIP_CHECK(hipGetLastError());
	return launch_plan(b, true);
}

}  // namespace

extern "C" int rpwa_batch_create(rpwa_batch** out, int device, uint32_t nmbWaves, uint32_t nmbStarts)
{
	clear_error();
	if (!out) {
		set_error("rpwa_batch_create: null pointer");
		return RPWA_ERR_INVALID;
	}
	*out = nullptr;
	if (nmbWaves == 0 || nmbWaves > RPWA_BATCH_MAX_WAVES || nmbStarts == 0 || nmbStarts > RPWA_BATCH_MAX_STARTS) {
		set_error("rpwa_batch_create: need 1 <= nmbWaves <= 96 and 1 <= nmbStarts <= 64");
		return RPWA_ERR_INVALID;
	}
	int nDev = 0;
	if (hipGetDeviceCount(&nDev) != hipSuccess || device < 0 || device >= nDev) {
		set_error("rpwa_batch_create: no GPU with index " + std::to_string(device));
		return RPWA_ERR_INVALID;
	}
	std::unique_ptr<rpwa_batch> b(new rpwa_batch);
	b->device = device;
	b->W = nmbWaves;
	b->Wp = (nmbWaves + 7) / 8 * 8;
	b->rowLen = 2 * b->Wp;
	b->NT = (b->rowLen + 31) / 32;
	b->nStarts = nmbStarts;
	b->P = 2 * nmbWaves;
	RPWA_HIP_CHECK(hipSetDevice(device));
	RPWA_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
	RPWA_HIP_CHECK(hipEventCreate(&b->ev0));
	RPWA_HIP_CHECK(hipEventCreate(&b->ev1));
	RPWA_HIP_CHECK(hipDeviceGetAttribute(&b->nCu, hipDeviceAttributeMultiprocessorCount, device));
	*out = b.release();
	return RPWA_OK;
}

extern "C" int rpwa_batch_destroy(rpwa_batch* b)
{
	clear_error();
	if (!b) {
		set_error("rpwa_batch_destroy: null batch");
		return RPWA_ERR_INVALID;
	}
	(void)hipSetDevice(b->device);
	if (b->stream)
		(void)hipStreamSynchronize(b->stream);
	if (b->graph)
		(void)hipGraphExecDestroy(b->graph);
	for (auto& bin : b->bins)
		if (bin.d_amps)
			(void)hipFree(bin.d_amps);
	for (void* p : b->allocations)
		(void)hipFree(p);
	if (b->ev0)
		(void)hipEventDestroy(b->ev0);
	if (b->ev1)
		(void)hipEventDestroy(b->ev1);
	if (b->stream)
		(void)hipStreamDestroy(b->stream);
	delete b;
	return RPWA_OK;
}

extern "C" int rpwa_batch_add_bin(rpwa_batch* b, const double* amps, uint64_t nmbEvents, int onDevice,
                                  const double* normSums, uint64_t nmbNormEvents, const double* accSums,
                                  uint64_t nmbAccEvents, uint32_t* binIndex)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_add_bin", false));
	if (b->finished) {
		set_error("rpwa_batch_add_bin: rpwa_batch_finish was called already");
		return RPWA_ERR_STATE;
	}
	if (!amps || !normSums || nmbEvents == 0 || nmbNormEvents == 0 || (accSums && nmbAccEvents == 0)) {
		set_error("rpwa_batch_add_bin: null pointer or empty input");
		return RPWA_ERR_INVALID;
	}
	const uint32_t W = b->W;
	// pwaLikelihood::addNormIntegral / addAccIntegral (pwaLikelihood.cc:1068-1107):
	// element = sums / nmbEvents; eta = accEvents / normEvents
	std::vector<double> nrm(2 * (size_t)W * W), acc(2 * (size_t)W * W);
	const double* as = accSums ? accSums : normSums;
	const double nAcc = (double)nmbNormEvents;  // accMatrix.setNmbEvents(_numbAccEvents)
	for (size_t k = 0; k < (size_t)W * W; ++k) {
		nrm[2 * k] = normSums[2 * k] / (double)nmbNormEvents;
		nrm[2 * k + 1] = normSums[2 * k + 1] / (double)nmbNormEvents;
		acc[2 * k] = as[2 * k] / nAcc;
		acc[2 * k + 1] = as[2 * k + 1] / nAcc;
	}
	rpwa_batch::Bin bin;
	bin.nEvents = nmbEvents;
	bin.nTiles = (uint32_t)((nmbEvents + kEv - 1) / kEv);
	bin.totAcc = accSums ? (double)nmbAccEvents / (double)nmbNormEvents : 1.;
	// finishInit with normalised amplitudes (pwaLikelihood.cc:1311-1334)
	std::vector<double> scale(W);
	for (uint32_t i = 0; i < W; ++i) {
		const double ni = nrm[2 * ((size_t)i * W + i)];
		if (!(ni > 0.)) {
			set_error("rpwa_batch_add_bin: wave " + std::to_string(i) +
			          " has a zero normalisation integral (not supported in batched fits)");
			return RPWA_ERR_INVALID;
		}
		scale[i] = 1. / std::sqrt(ni);
	}
	bin.acc.resize(2 * (size_t)W * W);
	for (uint32_t i = 0; i < W; ++i)
		for (uint32_t j = 0; j < W; ++j) {
			const double d = std::sqrt(nrm[2 * ((size_t)i * W + i)]) * std::sqrt(nrm[2 * ((size_t)j * W + j)]);
			bin.acc[2 * ((size_t)i * W + j)] = acc[2 * ((size_t)i * W + j)] / d;
			bin.acc[2 * ((size_t)i * W + j) + 1] = acc[2 * ((size_t)i * W + j) + 1] / d;
		}
	RPWA_HIP_CHECK(hipSetDevice(b->device));
	const uint64_t nRows = (uint64_t)bin.nTiles * kEv;
	RPWA_HIP_CHECK(hipMalloc(&bin.d_amps, nRows * b->rowLen * sizeof(float)));
	double* d_scale = nullptr;
	double2* d_src = nullptr;
	int rc = RPWA_OK;
	do {
		if (hipMalloc(&d_scale, W * sizeof(double)) != hipSuccess ||
		    hipMemcpy(d_scale, scale.data(), W * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
			rc = RPWA_ERR_HIP;
			break;
		}
		const double2* src = reinterpret_cast<const double2*>(amps);
		if (!onDevice) {
			if (hipMalloc(&d_src, nmbEvents * W * sizeof(double2)) != hipSuccess ||
			    hipMemcpy(d_src, amps, nmbEvents * W * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
				rc = RPWA_ERR_NOMEM;
				break;
			}
			src = d_src;
		}
		const uint64_t n = nRows * b->Wp;
		hipLaunchKernelGGL(batch_transform_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, b->stream, src,
		                   nmbEvents, W, b->Wp, nRows, d_scale, reinterpret_cast<float2*>(bin.d_amps));
		if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
			rc = RPWA_ERR_HIP;
	} while (false);
	if (d_scale)
		(void)hipFree(d_scale);
	if (d_src)
		(void)hipFree(d_src);
	if (rc != RPWA_OK) {
		(void)hipFree(bin.d_amps);
		set_error("rpwa_batch_add_bin: device allocation, copy or transform failed");
		return rc;
	}
	if (binIndex)
		*binIndex = (uint32_t)b->bins.size();
	b->totalEvents += nmbEvents;
	b->bins.push_back(std::move(bin));
	return RPWA_OK;
}

extern "C" int rpwa_batch_finish(rpwa_batch* b)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_finish", false));
	if (b->finished)
		return RPWA_OK;
	if (b->bins.empty()) {
		set_error("rpwa_batch_finish: no bins");
		return RPWA_ERR_STATE;
	}
	RPWA_HIP_CHECK(hipSetDevice(b->device));
	const uint32_t nBins = (uint32_t)b->bins.size();
	// work chunks: each bin's tiles in runs of kChunkTiles
	std::vector<uint32_t> binChunk0(nBins + 1, 0), nEv(nBins), chunkBin, chunkTile0;
	std::vector<float*> ampPtr(nBins);
	std::vector<double> acc, totAcc(nBins);
	b->nTiles = 0;
	for (uint32_t i = 0; i < nBins; ++i) {
		const auto& bin = b->bins[i];
		const uint32_t chunks = (bin.nTiles + kChunkTiles - 1) / kChunkTiles;
		binChunk0[i + 1] = binChunk0[i] + chunks;
		for (uint32_t c = 0; c < chunks; ++c) {
			chunkBin.push_back(i);
			chunkTile0.push_back(c * kChunkTiles);
		}
		b->nTiles += bin.nTiles;
		nEv[i] = (uint32_t)bin.nEvents;
		ampPtr[i] = bin.d_amps;
		acc.insert(acc.end(), bin.acc.begin(), bin.acc.end());
		totAcc[i] = bin.totAcc;
	}
	b->nChunks = binChunk0[nBins];
	int occ = 0;
	RPWA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, pick_event_kernel((int)(b->Wp / 4)), kThreads,
	                                                            0));
	b->maxWgPerCu = (uint32_t)std::max(1, std::min(occ, 4));
	const size_t F = nFits(b), slots = (size_t)b->nChunks * 4;
	RPWA_TRY(dalloc(b, &b->d_binAmps, nBins));
	RPWA_TRY(dalloc(b, &b->d_chunkBin, b->nChunks));
	RPWA_TRY(dalloc(b, &b->d_chunkTile0, b->nChunks));
	RPWA_TRY(dalloc(b, &b->d_binChunk0, nBins + 1));
	RPWA_TRY(dalloc(b, &b->d_nEvents, nBins));
	RPWA_TRY(dalloc(b, &b->d_units, slots));
	RPWA_TRY(dalloc(b, &b->d_nUnits, 1));
	RPWA_TRY(dalloc(b, &b->d_unitCount, nBins));
	RPWA_TRY(dalloc(b, &b->d_colMap, F));
	RPWA_TRY(dalloc(b, &b->d_startCol, F));
	RPWA_TRY(dalloc(b, &b->d_T32, F * b->rowLen));
	RPWA_TRY(dalloc(b, &b->d_flat2, F));
	RPWA_TRY(dalloc(b, &b->d_partG, slots * b->Wp * 2 * kGroup));
	RPWA_TRY(dalloc(b, &b->d_partLn, slots * kGroup));
	RPWA_TRY(dalloc(b, &b->d_partInv, slots * kGroup));
	RPWA_TRY(dalloc(b, &b->d_acc, acc.size()));
	RPWA_TRY(dalloc(b, &b->d_totAcc, nBins));
	RPWA_TRY(dalloc(b, &b->d_nRunning, 1));
	RPWA_TRY(dalloc(b, &b->d_noStatus, F));
	RPWA_TRY(dalloc(b, &b->st.xt, F * b->P));
	RPWA_TRY(dalloc(b, &b->st.gt, F * b->P));
	RPWA_TRY(dalloc(b, &b->st.ft, F));
	RPWA_HIP_CHECK(hipMemcpy(b->d_binAmps, ampPtr.data(), nBins * sizeof(float*), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemcpy(b->d_chunkBin, chunkBin.data(), b->nChunks * sizeof(uint32_t), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemcpy(b->d_chunkTile0, chunkTile0.data(), b->nChunks * sizeof(uint32_t),
	                         hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemcpy(b->d_binChunk0, binChunk0.data(), (nBins + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemcpy(b->d_nEvents, nEv.data(), nBins * sizeof(uint32_t), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemcpy(b->d_acc, acc.data(), acc.size() * sizeof(double), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemcpy(b->d_totAcc, totAcc.data(), nBins * sizeof(double), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemset(b->d_noStatus, 0, F * sizeof(int32_t)));
	// T = 0 (padding waves stay zero), flat^2 = 1 for every start
	RPWA_HIP_CHECK(hipMemset(b->d_T32, 0, F * b->rowLen * sizeof(float)));
	std::vector<double> ones(F, 1.);
	RPWA_HIP_CHECK(hipMemcpy(b->d_flat2, ones.data(), F * sizeof(double), hipMemcpyHostToDevice));
	b->finished = true;

	// autotuner: event-kernel time of an evaluation of every start at the
	// drivers' start values, for each number of workgroups per CU the kernel's
	// occupancy allows
	std::vector<double> start, padded;
	host_start_values(b, 1u, start);
	pad_params(b, start.data(), padded);
	RPWA_TRY(prepare_eval(b, padded));
	uint32_t best = 1;
	if (b->wgPerCuFixed) {
		best = std::min(b->wgPerCuFixed, b->maxWgPerCu);
	} else if (b->maxWgPerCu > 1) {
		double bestMs = 1e300;
		for (uint32_t k = 1; k <= b->maxWgPerCu; ++k) {
			b->nWg = b->nCu * k;
			for (int w = 0; w < 2; ++w)  // warm-up
				RPWA_TRY(launch_event(b));
			double ms = 0.;
			RPWA_TRY(time_event(b, 3, &ms));
			if (std::getenv("RPWA_VERBOSE"))
				fprintf(stderr, "[rpwa batch autotune] %u WG/CU: %.3f ms\n", k, ms);
			if (ms < bestMs) {
				bestMs = ms;
				best = k;
			}
		}
	}
	b->wgPerCu = best;
	b->nWg = b->nCu * best;
	RPWA_HIP_CHECK(hipStreamSynchronize(b->stream));
	return RPWA_OK;
}

extern "C" int rpwa_batch_nmb_pars(const rpwa_batch* b, uint32_t* nmbPars)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_nmb_pars", false));
	if (!nmbPars) {
		set_error("rpwa_batch_nmb_pars: null pointer");
		return RPWA_ERR_INVALID;
	}
	*nmbPars = b->P;
	return RPWA_OK;
}

extern "C" int rpwa_batch_nmb_bins(const rpwa_batch* b, uint32_t* nmbBins)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_nmb_bins", false));
	if (!nmbBins) {
		set_error("rpwa_batch_nmb_bins: null pointer");
		return RPWA_ERR_INVALID;
	}
	*nmbBins = (uint32_t)b->bins.size();
	return RPWA_OK;
}

extern "C" int rpwa_batch_eval(rpwa_batch* b, const double* par, double* value, double* grad)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_eval", true));
	if (!par || !value) {
		set_error("rpwa_batch_eval: null pointer");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(b->device));
	const size_t F = nFits(b), P = b->P, nBins = b->bins.size();
	std::vector<double> padded;
	pad_params(b, par, padded);
	RPWA_TRY(prepare_eval(b, padded));
	RPWA_TRY(launch_event(b));
	RPWA_TRY(launch_reduce(b));
	std::vector<double> ft(F), gt;
	RPWA_HIP_CHECK(hipMemcpyAsync(ft.data(), b->st.ft, F * sizeof(double), hipMemcpyDeviceToHost, b->stream));
	if (grad) {
		gt.resize(F * P);
		RPWA_HIP_CHECK(hipMemcpyAsync(gt.data(), b->st.gt, F * P * sizeof(double), hipMemcpyDeviceToHost, b->stream));
	}
	RPWA_HIP_CHECK(hipStreamSynchronize(b->stream));
	for (size_t bi = 0; bi < nBins; ++bi)
		for (uint32_t s = 0; s < b->nStarts; ++s) {
			value[bi * b->nStarts + s] = ft[bi * kCols + s];
			if (grad)
				std::memcpy(grad + (bi * b->nStarts + s) * P, &gt[(bi * kCols + s) * P], P * sizeof(double));
		}
	return RPWA_OK;
}

extern "C" int rpwa_batch_start_values(const rpwa_batch* b, uint32_t seed, double* out)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_start_values", false));
	if (!out) {
		set_error("rpwa_batch_start_values: null pointer");
		return RPWA_ERR_INVALID;
	}
	std::vector<double> v;
	host_start_values(b, seed, v);
	std::memcpy(out, v.data(), v.size() * sizeof(double));
	return RPWA_OK;
}

extern "C" int rpwa_batch_fit(rpwa_batch* b, const double* startPar, uint32_t seed, const rpwa_fit_options* opts,
                              double* outPar, double* outValue, int32_t* outStatus, uint32_t* outNmbEval)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_fit", true));
	uint32_t memory = 10, maxEval = 50000;
	double xtolRel = 1e-4, ftolAbs = 1e-6;
	if (opts) {
		memory = opts->memory;
		maxEval = opts->maxEvaluations;
		xtolRel = opts->xtolRel;
		ftolAbs = opts->ftolAbs;
	}
	if (memory == 0 || memory > (uint32_t)kMaxMem || maxEval == 0) {
		set_error("rpwa_batch_fit: need 1 <= memory <= 16 and maxEvaluations > 0");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(b->device));
	RPWA_TRY(alloc_state(b, memory));
	const size_t F = nFits(b), P = b->P, nBins = b->bins.size();
	std::vector<double> start;
	if (!startPar) {
		host_start_values(b, seed, start);
		startPar = start.data();
	}
	std::vector<double> padded;
	pad_params(b, startPar, padded);
	RPWA_HIP_CHECK(hipMemcpyAsync(b->st.xt, padded.data(), padded.size() * sizeof(double), hipMemcpyHostToDevice,
	                              b->stream));
	BatchInitArgs ia{b->st, b->d_T32, b->d_flat2, b->W, b->P, b->rowLen, b->nStarts};
	hipLaunchKernelGGL(batch_init_kernel, dim3((uint32_t)F), dim3(64), 0, b->stream, ia);
	RPWA_HIP_CHECK(hipGetLastError());
	RPWA_TRY(launch_plan(b, false));

	// one iteration = evaluate every running fit, reduce, minimiser step, next plan
	if (!b->graph || b->graphMem != memory || b->graphMaxEval != maxEval || b->graphXtol != xtolRel ||
	    b->graphFtol != ftolAbs) {
		if (b->graph) {
			(void)hipGraphExecDestroy(b->graph);
			b->graph = nullptr;
		}
		hipGraph_t g = nullptr;
		RPWA_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
		int rc = launch_event(b);
		if (rc == RPWA_OK)
			rc = launch_reduce(b);
		if (rc == RPWA_OK)
			rc = launch_lbfgs(b, maxEval, xtolRel, ftolAbs);
		if (rc == RPWA_OK)
			rc = launch_plan(b, false);
		const hipError_t ec = hipStreamEndCapture(b->stream, &g);
		if (rc != RPWA_OK)
			return rc;
		RPWA_HIP_CHECK(ec);
		const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
		(void)hipGraphDestroy(g);
		RPWA_HIP_CHECK(ei);
		b->graphMem = memory;
		b->graphMaxEval = maxEval;
		b->graphXtol = xtolRel;
		b->graphFtol = ftolAbs;
	}
	int32_t nRunning = 1;
	uint64_t iterations = 0;
	const uint64_t maxIterations = (uint64_t)maxEval + 1;
	while (nRunning > 0 && iterations < maxIterations) {
		const uint64_t chunk = std::min<uint64_t>(8, maxIterations - iterations);
		for (uint64_t k = 0; k < chunk; ++k)
			RPWA_HIP_CHECK(hipGraphLaunch(b->graph, b->stream));
		iterations += chunk;
		RPWA_HIP_CHECK(hipMemcpyAsync(&nRunning, b->d_nRunning, sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
		RPWA_HIP_CHECK(hipStreamSynchronize(b->stream));
		if (std::getenv("RPWA_VERBOSE"))
			fprintf(stderr, "[rpwa batch] iteration %llu: %d fits running\n", (unsigned long long)iterations, nRunning);
	}
	b->lastEvaluations = iterations;

	std::vector<double> x(F * P), f(F);
	std::vector<int32_t> status(F), nEval(F);
	RPWA_HIP_CHECK(hipMemcpy(x.data(), b->st.x, F * P * sizeof(double), hipMemcpyDeviceToHost));
	RPWA_HIP_CHECK(hipMemcpy(f.data(), b->st.f, F * sizeof(double), hipMemcpyDeviceToHost));
	RPWA_HIP_CHECK(hipMemcpy(status.data(), b->st.status, F * sizeof(int32_t), hipMemcpyDeviceToHost));
	RPWA_HIP_CHECK(hipMemcpy(nEval.data(), b->st.nEval, F * sizeof(int32_t), hipMemcpyDeviceToHost));
	for (size_t bi = 0; bi < nBins; ++bi)
		for (uint32_t s = 0; s < b->nStarts; ++s) {
			const size_t fi = bi * kCols + s, o = bi * b->nStarts + s;
			double* xp = &x[fi * P];
			if (status[fi] == ST_BADSTART)  // the start point itself
				std::memcpy(xp, &padded[fi * P], P * sizeof(double));
			// CorrectParamSigns (pwaLikelihood.cc:986-1010) for rank 1, one sector:
			// a negative anchor flips the sector; the flat amplitude is |flat|
			if (xp[0] < 0.)
				for (size_t k = 0; k + 1 < P; ++k)
					xp[k] = -xp[k];
			xp[P - 1] = std::fabs(xp[P - 1]);
			if (outPar)
				std::memcpy(outPar + o * P, xp, P * sizeof(double));
			if (outValue)
				outValue[o] = status[fi] == ST_BADSTART ? NAN : f[fi];
			if (outStatus)
				outStatus[o] = status[fi] == ST_SEARCH || status[fi] == ST_INIT ? ST_MAXEVAL : status[fi];
			if (outNmbEval)
				outNmbEval[o] = (uint32_t)nEval[fi];
		}
	return RPWA_OK;
}

extern "C" int rpwa_batch_set_tile(rpwa_batch* b, uint32_t wgPerCu)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_set_tile", false));
	b->wgPerCuFixed = wgPerCu;
	if (b->finished) {
		const uint32_t k = wgPerCu ? std::min(wgPerCu, b->maxWgPerCu) : b->wgPerCu;
		b->wgPerCu = k;
		b->nWg = b->nCu * k;
		if (b->graph) {
			(void)hipGraphExecDestroy(b->graph);
			b->graph = nullptr;
		}
	}
	return RPWA_OK;
}

extern "C" int rpwa_batch_info(const rpwa_batch* b, uint64_t info[8])
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_info", false));
	if (!info) {
		set_error("rpwa_batch_info: null pointer");
		return RPWA_ERR_INVALID;
	}
	info[0] = b->bins.size();
	info[1] = b->nTiles;
	info[2] = b->nChunks;
	info[3] = b->nWg;
	info[4] = b->wgPerCu;
	info[5] = b->Wp;
	info[6] = b->totalEvents;
	info[7] = b->lastEvaluations;
	return RPWA_OK;
}

extern "C" int rpwa_batch_time_event_kernel(rpwa_batch* b, uint32_t reps, double* ms)
{
	clear_error();
	RPWA_TRY(check_batch(b, "rpwa_batch_time_event_kernel", true));
	if (!ms || reps == 0) {
		set_error("rpwa_batch_time_event_kernel: null pointer or zero repetitions");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(b->device));
	RPWA_TRY(launch_plan(b, true));  // every start at the current trial points
	for (int w = 0; w < 5; ++w)      // untimed: clocks and caches reach their steady state
		RPWA_TRY(launch_event(b));
	return time_event(b, reps, ms);
}
