// stage_kernel.hip -- the LDS-staged event kernel: one lane per event.
//
// Computes exactly what the register-tile kernels of lik_kernels.hip compute
// (sum_e -ln sigma_e, sum_e -2/sigma_e and the event part of the
// gradient, pwaLikelihood.cc:213-244), with a different work decomposition
// that keeps neither the production amplitudes nor a slice of the amplitude
// row in the registers of every lane:
//
//   * a tile of E consecutive events is one contiguous block of the amplitude
//     array; the 256 threads of a workgroup fetch it with coalesced 16-byte
//     loads one tile ahead (the loads of tile t+1 are in flight while tile t is
//     computed) and store it into LDS with an odd row stride in 16-byte units,
//     so that lanes reading the same column of consecutive rows hit distinct
//     banks;
//   * S phase: wave w contracts column slice w of every row, lane = event.
//     The production amplitudes of a column are the same for every lane and
//     are read through the scalar unit, from the kernel arguments when they
//     fit (3 KB), else from memory; waves 1..3 leave their partial S in LDS
//     and wave 0 adds them in the order 0,1,2,3;
//   * sigma phase (wave 0, lane = event): sigma, the per-lane running product
//     of sigma as m 2^e (its log is taken once at the end: the phase is on the
//     critical path of every tile, a log per event is not), w = -2/sigma, and
//     u = w S stored in LDS;
//   * G phase: thread = (group of 2-4 consecutive 16-byte units of one sector,
//     event subset); it reads u of an event once for its group, adds
//     u conj(A) over its events of the tile in FP64 and keeps the sums across
//     tiles;
//   * end: the event subsets are added in fixed order and written as the
//     workgroup's partial row, in the format of the register-tile kernels, so
//     event_reduce and everything after it is shared.
// Every sum has a fixed order for a fixed grid: results are bitwise
// reproducible and invariant under T -> -T of a (rank, sector) block.
//
// Gradient on the matrix cores (mixed precision, rows of at most 64 units):
// the G phase can instead run on v_mfma_f32_4x4x1f32 (16 blocks of 4x4x1).
// Block b is one 16-byte unit of the row (its four floats, the real and
// imaginary parts of two waves, are the rows i), the columns j are the real
// and imaginary parts of u for two ranks, K is one event: D[b][i][j] =
// sum_e x_i(e) u_j(e).  Layout on gfx950: A[b][i] in lane 4b+i, B[b][j] in
// lane 4b+j, D[b][i][j] in register i of lane 4b+j.  Wave w takes a quarter of
// the tile's events for all units; u is rounded to FP32 (the amplitudes are
// FP32 already), the products of a wave's events of one tile are summed in
// FP32 by the matrix core and added into FP64 sums per tile; at the end the
// waves' sums are added in order and combined into u conj(x).  S, sigma, the
// log sum and the weights stay FP64 as in the VALU variant.
//
// Ranks above the instantiation's R (at most 4 in registers) run in blocks of
// R ranks over several launches (STAGE_SIGMA, then STAGE_GRAD per block):
// sigma_e = sum over the blocks of sum_{r in block} |S_r(e)|^2 is accumulated
// in a per-event device array, block by block in order, then every block's
// launch reads the complete sigma_e, forms u = -2 S/sigma for its own ranks
// and adds its gradient entries to the workgroup rows; the first block also
// takes the log sum and the weight sum.  Each amplitude tile is read 2 x blocks
// times; the sums keep a fixed order.
#include "event_kernel.h"

#include <cstddef>

namespace rpwa_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxLoads = 12;  // 16-byte loads per thread and tile (E * rowUnits <= 3072)
constexpr int kInlineT = 192;  // complex values of T that travel in the kernel arguments (3 KB)

typedef float f32x4 __attribute__((ext_vector_type(4)));  // one 16-byte unit

// G phase: consecutive units of one sector handled by one thread (the
// weights u of an event are read once for all of them)
template <bool MIXED, int R>
constexpr int group_units() { return (MIXED ? 2 : 1) * R <= 2 ? 4 : 2; }
// LDS exchange area: complex values per event (2 sectors x R ranks, plus one
// so that lanes reading consecutive events hit distinct banks)
template <int R>
constexpr int part_stride() { return 2 * R + 1; }
typedef const __attribute__((address_space(4))) double cdouble;

// MFMA gradient: column sets of four (u_re, u_im per rank) and unit groups of 16
template <int R>
constexpr int gm_col_sets() { return (2 * R + 3) / 4; }
constexpr int kGmUnits = 16;     // units per MFMA (one per block)
constexpr int kGmMaxGroups = 4;  // rows of at most 64 units
typedef float gm_f4 __attribute__((ext_vector_type(4)));

// T of one row column through the scalar unit (the index is the same in every lane)
__device__ __forceinline__ double2 scalar_t(cdouble* tp, uint32_t i)
{
	return make_double2(tp[2 * i], tp[2 * i + 1]);
}

struct StageArgs {
	const f32x4* amps;      // nEvents rows of rowUnits 16-byte units
	int mode;               // STAGE_FUSED, STAGE_SIGMA or STAGE_GRAD
	uint32_t rank0;         // first rank of this launch's block (STAGE_GRAD)
	uint32_t accumulate;    // STAGE_SIGMA: add to sigma (blocks after the first)
	double* sigma;          // STAGE_SIGMA / STAGE_GRAD: per-event sum over the ranks [nEvents]
	uint64_t nEvents;
	uint64_t lastUnit;      // nEvents * rowUnits - 1
	uint32_t nTiles;
	uint32_t rowUnits;      // 16-byte units per row
	uint32_t rowUnitsP;     // LDS row stride in units (odd)
	uint32_t tileUnits;     // E * rowUnits
	uint32_t sliceUnits;    // units per wave in the S phase
	uint32_t secUnit;       // first unit of the refl +1 sector
	uint32_t gGroups0;      // G-phase unit groups of the refl -1 sector
	uint32_t gGroups;       // G-phase unit groups of both sectors (<= 256)
	uint32_t gSubs;         // event subsets per unit group in the G phase
	uint32_t gmGroups;      // MFMA gradient: unit groups of 16 (ceil(rowUnits / 16))
	uint32_t stageBytes;    // (E * rowUnitsP + 1) * 16: the rows and the dump unit
	uint32_t secOff1, nW0, nW1, wmax;
	const double2* tpad;    // [rowUnits * P][R]: T of each row column, zero on padding
	double flat2;
	double* partials;       // [nWg][nElem]
	uint32_t nElem;
	uint32_t tInline;       // 1: T is tin below (read from the kernel arguments), not tpad
	double2 tin[kInlineT];
};

static_assert(sizeof(StageArgs) <= 4096, "the staged kernel's arguments exceed the 4 KiB kernel-argument segment");

// all LDS writes of this wave done, then the workgroup barrier (no wait for
// the global loads in flight)
__device__ __forceinline__ void lds_barrier()
{
	asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <bool MIXED>
__device__ __forceinline__ double2 unit_cplx(const f32x4 v, int k)
{
	if (MIXED)
		return k == 0 ? make_double2((double)v.x, (double)v.y) : make_double2((double)v.z, (double)v.w);
	return make_double2(__hiloint2double(__float_as_int(v.y), __float_as_int(v.x)),
	                    __hiloint2double(__float_as_int(v.w), __float_as_int(v.z)));
}

// The three phases of one staged tile: S (all waves), sigma (wave 0), G (all
// threads).  The stage holds the tile; part holds the S partials of waves
// 1..NW-1 ([NW-1][E][PS]); u is the exchange area of u = -2 S / sigma ([E][PS]).

// ---- S phase: wave wid, slice of units [u0, u1), lane = event ----
template <bool MIXED, int R, int E>
__device__ __forceinline__ void s_phase(const StageArgs& a, const f32x4* stage, double2* part, int lane, int wid,
                                        double2 (&S)[2][R])
{
	constexpr int P = MIXED ? 2 : 1;
	constexpr int PS = part_stride<R>();
#pragma unroll
	for (int s = 0; s < 2; ++s)
#pragma unroll
		for (int r = 0; r < R; ++r)
			S[s][r] = make_double2(0., 0.);
	if (E == 64 || lane < E) {
		const f32x4* row = stage + (uint32_t)lane * a.rowUnitsP;
		const uint32_t u0 = (uint32_t)wid * a.sliceUnits;
		const uint32_t u1 = min(u0 + a.sliceUnits, a.rowUnits);
		const uint32_t um = min(max(a.secUnit, u0), u1);
		// T from the kernel arguments (the argument block starts the kernarg segment) or from memory
		typedef const __attribute__((address_space(4))) char cchar;
		cdouble* tp = a.tInline ? (cdouble*)((cchar*)__builtin_amdgcn_kernarg_segment_ptr() + offsetof(StageArgs, tin))
		                        : (cdouble*)(const double*)a.tpad;
#pragma unroll 4
		for (uint32_t u = u0; u < um; ++u) {
			const f32x4 v = row[u];
#pragma unroll
			for (int k = 0; k < P; ++k) {
				const double2 x = unit_cplx<MIXED>(v, k);
#pragma unroll
				for (int r = 0; r < R; ++r)
					ek::cmac(S[0][r], scalar_t(tp, (u * P + k) * R + r), x);
			}
		}
#pragma unroll 4
		for (uint32_t u = um; u < u1; ++u) {
			const f32x4 v = row[u];
#pragma unroll
			for (int k = 0; k < P; ++k) {
				const double2 x = unit_cplx<MIXED>(v, k);
#pragma unroll
				for (int r = 0; r < R; ++r)
					ek::cmac(S[1][r], scalar_t(tp, (u * P + k) * R + r), x);
			}
		}
		if (wid > 0) {
			double2* p = part + ((size_t)(wid - 1) * E + lane) * PS;
#pragma unroll
			for (int s = 0; s < 2; ++s)
#pragma unroll
				for (int r = 0; r < R; ++r)
					p[s * R + r] = S[s][r];
		}
	}
}

// ---- sigma phase: wave 0, lane = event; S holds wave 0's own partial ----
template <bool MIXED, int R, int E, int NW, bool GM>
__device__ __forceinline__ void sigma_phase(const StageArgs& a, const double2* part, double2* uOut, int lane,
                                            uint32_t tile, double2 (&S)[2][R], double& mant, int& expo,
                                            double& wsum)
{
	constexpr int PS = part_stride<R>();
	if (!(E == 64 || lane < E))
		return;
#pragma unroll
	for (int w = 1; w < NW; ++w) {
		const double2* p = part + ((size_t)(w - 1) * E + lane) * PS;
#pragma unroll
		for (int s = 0; s < 2; ++s)
#pragma unroll
			for (int r = 0; r < R; ++r) {
				const double2 q = p[s * R + r];
				S[s][r].x += q.x;
				S[s][r].y += q.y;
			}
	}
	double sigma = 0.;
#pragma unroll
	for (int r = 0; r < R; ++r) {
		sigma += S[0][r].x * S[0][r].x + S[0][r].y * S[0][r].y;
		sigma += S[1][r].x * S[1][r].x + S[1][r].y * S[1][r].y;
	}
	const uint64_t ev = (uint64_t)tile * E + lane;
	const bool valid = ev < a.nEvents;
	if (a.mode == STAGE_SIGMA) {
		// this block's share of sigma_e, added in block order
		if (valid)
			a.sigma[ev] = a.accumulate ? a.sigma[ev] + sigma : sigma;
	}
	if (a.mode == STAGE_GRAD)
		sigma = valid ? a.sigma[ev] : 0.;
	sigma += a.flat2;
	double w = 0.;
	if (valid && a.mode != STAGE_SIGMA) {
		// sum of ln sigma as the log of a running product m 2^e, m in [1, 2)
		// (one log per lane at the end instead of one per event); the rank
		// blocks after the first leave it to the first
		mant *= a.rank0 == 0 ? sigma : 1.;
		const uint64_t ub = (uint64_t)__double_as_longlong(mant);
		const int ex = (int)((ub >> 52) & 0x7FF);
		if (ex != 0 && ex != 0x7FF) {
			expo += ex - 1023;
			mant = __longlong_as_double((long long)((ub & 0x800FFFFFFFFFFFFFull) | (1023ull << 52)));
		}
		// -2/sigma from the hardware reciprocal and one Newton step
		double r = __builtin_amdgcn_rcp(sigma);
		r = fma(fma(-sigma, r, 1.), r, r);
		w = -2. * r;
		wsum += w;
	}
	if (GM) {
		// FP32 u by [sector][column][event], columns (re, im) per rank, zero past 2R
		constexpr int NCOL = 4 * gm_col_sets<R>();
		float* uf = reinterpret_cast<float*>(uOut);
#pragma unroll
		for (int s = 0; s < 2; ++s) {
#pragma unroll
			for (int r = 0; r < R; ++r) {
				uf[(s * NCOL + 2 * r) * E + lane] = (float)(w * S[s][r].x);
				uf[(s * NCOL + 2 * r + 1) * E + lane] = (float)(w * S[s][r].y);
			}
#pragma unroll
			for (int c = 2 * R; c < NCOL; ++c)
				uf[(s * NCOL + c) * E + lane] = 0.f;
		}
		return;
	}
	double2* p = uOut + (size_t)lane * PS;
#pragma unroll
	for (int s = 0; s < 2; ++s)
#pragma unroll
		for (int r = 0; r < R; ++r)
			p[s * R + r] = make_double2(w * S[s][r].x, w * S[s][r].y);
}

// ---- G phase on the matrix cores (mixed): wave wid, events [wid E/4, (wid+1) E/4),
// every unit group; aOff / bOff are the lane's byte offset in a staged row and
// float offset in the u area for each group ----
template <int R, int E, int NG>
__device__ __forceinline__ void gm_phase(const StageArgs& a, const f32x4* stage, const float* uf, int wid,
                                         const uint32_t (&aOff)[NG], const uint32_t (&bOff)[NG],
                                         double (&GD)[NG][gm_col_sets<R>()][4])
{
	constexpr int NCS = gm_col_sets<R>();
	constexpr int EW = E / 4;
	if (a.mode == STAGE_SIGMA)
		return;
	gm_f4 acc[NG][NCS];
#pragma unroll
	for (int g = 0; g < NG; ++g)
#pragma unroll
		for (int c = 0; c < NCS; ++c)
			acc[g][c] = gm_f4{0.f, 0.f, 0.f, 0.f};
	const char* sb = reinterpret_cast<const char*>(stage);
	const uint32_t rowBytes = a.rowUnitsP * 16u;
#pragma unroll
	for (int q = 0; q < EW; q += 4) {
		const int e0 = wid * EW + q;
		gm_f4 bv[NG][NCS];
#pragma unroll
		for (int g = 0; g < NG; ++g)
#pragma unroll
			for (int c = 0; c < NCS; ++c)
				bv[g][c] = *reinterpret_cast<const gm_f4*>(uf + bOff[g] + c * 4 * E + e0);
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			const char* row = sb + (uint32_t)(e0 + k) * rowBytes;
#pragma unroll
			for (int g = 0; g < NG; ++g) {
				const float av = *reinterpret_cast<const float*>(row + aOff[g]);
#pragma unroll
				for (int c = 0; c < NCS; ++c)
					acc[g][c] = __builtin_amdgcn_mfma_f32_4x4x1f32(av, bv[g][c][k], acc[g][c], 0, 0, 0);
			}
		}
	}
#pragma unroll
	for (int g = 0; g < NG; ++g)
#pragma unroll
		for (int c = 0; c < NCS; ++c)
#pragma unroll
			for (int i = 0; i < 4; ++i)
				GD[g][c][i] += (double)acc[g][c][i];
}

// ---- G phase: thread = (group of UG units of one sector, event subset) ----
template <bool MIXED, int R, int E>
__device__ __forceinline__ void g_phase(const StageArgs& a, const f32x4* stage, const double2* uIn, bool gOn,
                                        uint32_t gU0, uint32_t gSub, int gSec,
                                        double2 (&G)[group_units<MIXED, R>()][MIXED ? 2 : 1][R])
{
	constexpr int P = MIXED ? 2 : 1;
	constexpr int UG = group_units<MIXED, R>();
	constexpr int PS = part_stride<R>();
	if (!gOn || a.mode == STAGE_SIGMA)
		return;
	// The slots of a group are consecutive units, read at fixed offsets from one
	// pointer per event.  Slots past the sector (or the row) read the next units,
	// the next row or the LDS areas after the stage, all inside the allocation;
	// their sums are discarded.
	const f32x4* ap = stage + gSub * a.rowUnitsP + gU0;
	const double2* up = uIn + gSub * PS + gSec * R;
	const uint32_t aStep = a.gSubs * a.rowUnitsP, uStep = a.gSubs * PS;
	for (uint32_t e = gSub; e < (uint32_t)E; e += a.gSubs, ap += aStep, up += uStep) {
		double2 u[R];
#pragma unroll
		for (int r = 0; r < R; ++r)
			u[r] = up[r];
#pragma unroll
		for (int j = 0; j < UG; ++j) {
			const f32x4 v = ap[j];
#pragma unroll
			for (int k = 0; k < P; ++k) {
				const double2 x = unit_cplx<MIXED>(v, k);
#pragma unroll
				for (int r = 0; r < R; ++r)
					ek::cmac_conj(G[j][k][r], u[r], x);
			}
		}
	}
}

// ---- wave 0 at the end: the lanes' -ln sigma sums and weight sums, combined
// by a fixed butterfly: at each step the lower lane's value is the left operand
// in both partners, and the rounding error of each addition (TwoSum) goes to
// the compensation ----
template <int E>
__device__ __forceinline__ void log_sums_butterfly(int lane, double mant, int expo, double wsum, double& s, double& c,
                                                   double& ws)
{
	const bool ev = E == 64 || lane < E;
	s = ev ? -(log(mant) + (double)expo * 0.69314718055994530942) : 0.;
	c = 0.;
	ws = ev ? wsum : 0.;
#pragma unroll
	for (int m = 32; m >= 1; m >>= 1) {
		const double so = __shfl_xor(s, m), co = __shfl_xor(c, m), wo = __shfl_xor(ws, m);
		const bool upper = (lane & m) != 0;
		const double x = upper ? so : s, y = upper ? s : so;
		const double t = x + y;
		const double yp = t - x;
		const double err = (x - (t - yp)) + (y - yp);
		c = (upper ? co + c : c + co) + err;
		ws = upper ? wo + ws : ws + wo;
		s = t;
	}
}

template <int E>
__device__ __forceinline__ void write_log_sums(const StageArgs& a, double* out, int lane, double mant, int expo,
                                               double wsum)
{
	double s, c, ws;
	log_sums_butterfly<E>(lane, mant, expo, wsum, s, c, ws);
	if (lane == 0 && a.rank0 == 0) {
		out[0] = s;
		out[1] = c;
		out[2] = ws;
	}
}

// every wave holds log products of its own events: the waves' butterfly sums
// are added in wave order (TwoSum into the compensation) by thread 0; lsum is
// a [waves][3] LDS area; ends with a workgroup barrier
__device__ __forceinline__ void write_log_sums_all_waves(const StageArgs& a, double* out, double (*lsum)[3], int tid,
                                                         int lane, int wid, double mant, int expo, double wsum)
{
	double s, c, ws;
	log_sums_butterfly<64>(lane, mant, expo, wsum, s, c, ws);
	if (lane == 0) {
		lsum[wid][0] = s;
		lsum[wid][1] = c;
		lsum[wid][2] = ws;
	}
	lds_barrier();
	if (tid == 0 && a.rank0 == 0) {
		double S = lsum[0][0], C = lsum[0][1], W = lsum[0][2];
#pragma unroll
		for (int w = 1; w < kWaves; ++w) {
			const double x = S, y = lsum[w][0];
			const double t = x + y;
			const double yp = t - x;
			C = (C + lsum[w][1]) + ((x - (t - yp)) + (y - yp));
			W += lsum[w][2];
			S = t;
		}
		out[0] = S;
		out[1] = C;
		out[2] = W;
	}
}

// ---- end of the MFMA variant: every wave's FP64 sums D[unit][column][row] go
// to LDS, then each gradient entry is formed from them, waves in order:
// u conj(x) = (u_re x_re + u_im x_im) + i (u_im x_re - u_re x_im) ----
template <int R, int E, int NG, int NT, bool ALLW = false>
__device__ __forceinline__ void write_partial_row_gm(const StageArgs& a, f32x4* smem, int tid, int lane, int wid,
                                                     const double (&GD)[NG][gm_col_sets<R>()][4], double mant,
                                                     int expo, double wsum, double (*lsum)[3] = nullptr)
{
	constexpr int NCS = gm_col_sets<R>();
	constexpr int NCOL = 4 * NCS;
	constexpr int UN = kGmUnits * NG;  // units per wave in the buffer
	lds_barrier();
	double* gbuf = reinterpret_cast<double*>(smem);  // [wave][UN][NCOL][4]
#pragma unroll
	for (int g = 0; g < NG; ++g)
#pragma unroll
		for (int c = 0; c < NCS; ++c) {
			double* p = gbuf + (((size_t)wid * UN + g * kGmUnits + lane / 4) * NCOL + c * 4 + lane % 4) * 4;
#pragma unroll
			for (int i = 0; i < 4; ++i)
				p[i] = GD[g][c][i];
		}
	double* out = a.partials + (size_t)blockIdx.x * a.nElem;
	if constexpr (ALLW) {
		write_log_sums_all_waves(a, out, lsum, tid, lane, wid, mant, expo, wsum);
	} else {
		if (wid == 0)
			write_log_sums<E>(a, out, lane, mant, expo, wsum);
		lds_barrier();
	}
	const uint32_t nCplx = (a.nElem - 3) / 2;
	for (uint32_t idx = tid; idx < nCplx; idx += NT) {
		const uint32_t j = idx % a.wmax;
		const uint32_t rs = idx / a.wmax;
		const int s = (int)(rs & 1);
		const int rg = (int)(rs >> 1);
		if (rg < (int)a.rank0 || rg >= (int)a.rank0 + R)
			continue;
		const int r = rg - (int)a.rank0;
		double2 v = make_double2(0., 0.);
		if (j < (s ? a.nW1 : a.nW0)) {
			const uint32_t c = (s ? a.secOff1 : 0u) + j;
			const uint32_t unit = c / 2;
			const uint32_t k = c % 2;
#pragma unroll
			for (int w = 0; w < kWaves; ++w) {
				const double* q = gbuf + ((size_t)w * UN + unit) * NCOL * 4;
				const double* ur = q + (2 * r) * 4;      // column u_re
				const double* ui = q + (2 * r + 1) * 4;  // column u_im
				v.x += ur[2 * k] + ui[2 * k + 1];
				v.y += ui[2 * k] - ur[2 * k + 1];
			}
		}
		out[3 + 2 * idx] = v.x;
		out[3 + 2 * idx + 1] = v.y;
	}
}

// ---- end: the workgroup's partial row (lnL sum and its compensation, weight
// sum, gradient entries), in the format of the register-tile kernels ----
template <bool MIXED, int R, int E, int NT>
__device__ __forceinline__ void write_partial_row(const StageArgs& a, f32x4* smem, int tid, int lane, int wid, bool gOn,
                                                  uint32_t gU0, uint32_t gEnd, uint32_t gSub,
                                                  double2 (&G)[group_units<MIXED, R>()][MIXED ? 2 : 1][R],
                                                  double mant, int expo, double wsum)
{
	constexpr int P = MIXED ? 2 : 1;
	constexpr int UG = group_units<MIXED, R>();
	lds_barrier();
	double2* gbuf = reinterpret_cast<double2*>(smem);  // [gSubs][rowUnits][P][R]
	if (gOn) {
#pragma unroll
		for (int j = 0; j < UG; ++j) {
			const uint32_t unit = gU0 + (uint32_t)j;
			if (unit < gEnd) {
				double2* p = gbuf + ((size_t)gSub * a.rowUnits + unit) * P * R;
#pragma unroll
				for (int k = 0; k < P; ++k)
#pragma unroll
					for (int r = 0; r < R; ++r)
						p[k * R + r] = G[j][k][r];
			}
		}
	}
	double* out = a.partials + (size_t)blockIdx.x * a.nElem;
	if (wid == 0)
		write_log_sums<E>(a, out, lane, mant, expo, wsum);
	lds_barrier();
	// element 3 + 2*((r*2 + s)*wmax + j) (+1 for the imaginary part)
	const uint32_t nCplx = (a.nElem - 3) / 2;
	for (uint32_t idx = tid; idx < nCplx; idx += NT) {
		const uint32_t j = idx % a.wmax;
		const uint32_t rs = idx / a.wmax;
		const int s = (int)(rs & 1);
		const int rg = (int)(rs >> 1);  // rank of the element
		if (rg < (int)a.rank0 || rg >= (int)a.rank0 + R)
			continue;  // another rank block's element
		const int r = rg - (int)a.rank0;
		double2 v = make_double2(0., 0.);
		if (j < (s ? a.nW1 : a.nW0)) {
			const uint32_t c = (s ? a.secOff1 : 0u) + j;
			const uint32_t unit = c / P;
			const uint32_t k = c % P;
			for (uint32_t sub = 0; sub < a.gSubs; ++sub) {
				const double2 q = gbuf[((size_t)sub * a.rowUnits + unit) * P * R + k * R + r];
				v.x += q.x;
				v.y += q.y;
			}
		}
		out[3 + 2 * idx] = v.x;
		out[3 + 2 * idx + 1] = v.y;
	}
}

// Loads of one tile into the staging registers buf[KL]: tiles before the
// array's last are complete, so their loads are the tile base plus each
// thread's fixed byte offsets (a load past the tile reads the tile's first
// unit); in the last tile indices past the array are clamped to its last unit
// (rows past the last event hold copies of real amplitudes and get weight zero).
// The loads are non-temporal: the array is streamed once per evaluation and is
// far larger than the caches (streaming reads reach ~7 TB/s this way on an
// MI355X, ~6.2 TB/s through the caches).
template <int KL>
__device__ __forceinline__ void fetch_tile(const StageArgs& a, uint32_t t, uint32_t lastTile, const uint32_t (&ldOff)[KL],
                                           f32x4 (&buf)[KL])
{
	if (t < lastTile) {
		const char* tb = reinterpret_cast<const char*>(a.amps + (uint64_t)t * a.tileUnits);
#pragma unroll
		for (int k = 0; k < KL; ++k)
			buf[k] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tb + ldOff[k]));
	} else {
		const uint64_t base = (uint64_t)t * a.tileUnits;
#pragma unroll
		for (int k = 0; k < KL; ++k) {
			const uint64_t g = base + ldOff[k] / 16u;
			buf[k] = __builtin_nontemporal_load(a.amps + (g < a.lastUnit ? g : a.lastUnit));
		}
	}
}

// LDS positions and tile byte offsets of a thread's KL loads (the same for
// every tile); a load past the tile goes to the dump unit after the stage and
// reads the tile's first unit
template <int NT, int KL, int E>
__device__ __forceinline__ void load_plan(const StageArgs& a, int tid, uint32_t (&ldsPos)[KL], uint32_t (&ldOff)[KL])
{
	const uint32_t dumpUnit = (uint32_t)E * a.rowUnitsP;
#pragma unroll
	for (int k = 0; k < KL; ++k) {
		const uint32_t g = (uint32_t)tid + (uint32_t)NT * k;
		const uint32_t row = g / a.rowUnits;
		const bool in = g < a.tileUnits;
		ldsPos[k] = in ? row * a.rowUnitsP + (g - row * a.rowUnits) : dumpUnit;
		ldOff[k] = in ? g * 16u : 0u;
	}
}

// Register-lean forms of the above for a kernel that holds two staging sets:
// the tile offsets are recomputed from the thread index (load k of thread tid
// is unit tid + NT k of the tile, or the tile's first unit past the tile) and
// the LDS positions are kept as 16-bit pairs.
template <int NT, int KL>
__device__ __forceinline__ void fetch_tile_lean(const StageArgs& a, uint32_t t, uint32_t lastTile, int tid,
                                                f32x4 (&buf)[KL])
{
	const uint64_t base = (uint64_t)t * a.tileUnits;
	if (t < lastTile) {
		const f32x4* tb = a.amps + base;
#pragma unroll
		for (int k = 0; k < KL; ++k) {
			const uint32_t g = (uint32_t)tid + (uint32_t)NT * k;
			buf[k] = __builtin_nontemporal_load(tb + (g < a.tileUnits ? g : 0u));
		}
	} else {
#pragma unroll
		for (int k = 0; k < KL; ++k) {
			const uint32_t g = (uint32_t)tid + (uint32_t)NT * k;
			const uint64_t gg = base + (g < a.tileUnits ? g : 0u);
			buf[k] = __builtin_nontemporal_load(a.amps + (gg < a.lastUnit ? gg : a.lastUnit));
		}
	}
}

template <int KL>
__device__ __forceinline__ void pack_positions(const uint32_t (&ldsPos)[KL], uint32_t (&packed)[(KL + 1) / 2])
{
#pragma unroll
	for (int k = 0; k < KL; k += 2)
		packed[k / 2] = ldsPos[k] | ((k + 1 < KL ? ldsPos[k + 1] : 0u) << 16);
}

template <int KL>
__device__ __forceinline__ void store_tile_packed(f32x4* stage, const uint32_t (&packed)[(KL + 1) / 2],
                                                  const f32x4 (&buf)[KL])
{
#pragma unroll
	for (int k = 0; k < KL; ++k)
		stage[(packed[k / 2] >> (16 * (k & 1))) & 0xFFFFu] = buf[k];
}

template <int KL>
__device__ __forceinline__ void store_tile(f32x4* stage, const uint32_t (&ldsPos)[KL], const f32x4 (&buf)[KL])
{
#pragma unroll
	for (int k = 0; k < KL; ++k)
		stage[ldsPos[k]] = buf[k];
}

// The staged kernel: 4 waves, one tile stage; per tile: stage the prefetched
// tile, S, sigma, G, each phase behind a workgroup barrier.  NG > 0 (mixed
// precision only): the G phase on the matrix cores with NG unit groups of 16.
template <bool MIXED, int R, int E, int NG>
__global__ __launch_bounds__(kThreads, 2) void stage_kernel(StageArgs a)
{
	constexpr int P = MIXED ? 2 : 1;  // complex values per 16-byte unit
	constexpr int UG = group_units<MIXED, R>();
	constexpr bool GM = NG > 0;
	constexpr int NGX = GM ? NG : 1;
	constexpr int NCS = gm_col_sets<R>();
	static_assert(!GM || MIXED, "the MFMA gradient needs complex64 amplitudes");
	extern __shared__ f32x4 smem[];
	f32x4* stage = smem;  // [E][rowUnitsP] units, then the dump unit
	double2* part = reinterpret_cast<double2*>(reinterpret_cast<char*>(smem) + a.stageBytes);  // [3][E][PS]

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
	uint32_t ldsPos[kMaxLoads], ldOff[kMaxLoads];
	load_plan<kThreads, kMaxLoads, E>(a, tid, ldsPos, ldOff);

	// G-phase role: unit group gi of sector gSec, units [gU0, min(gU0 + UG, gEnd))
	const uint32_t gi = (uint32_t)tid % a.gGroups;
	const uint32_t gSub = (uint32_t)tid / a.gGroups;
	const bool gOn = gSub < a.gSubs;
	const int gSec = gi >= a.gGroups0 ? 1 : 0;
	const uint32_t gU0 = gSec ? a.secUnit + (gi - a.gGroups0) * UG : gi * UG;
	const uint32_t gEnd = gSec ? a.rowUnits : a.secUnit;

	double2 G[UG][P][R];
#pragma unroll
	for (int j = 0; j < UG; ++j)
#pragma unroll
		for (int k = 0; k < P; ++k)
#pragma unroll
			for (int r = 0; r < R; ++r)
				G[j][k][r] = make_double2(0., 0.);
	// MFMA G phase: per unit group the lane's byte offset in a staged row (units
	// past the row read its last unit; their sums are never used) and its float
	// offset in the u area (sector of the unit, column lane % 4)
	double GD[NGX][NCS][4];
	uint32_t aOff[NGX], bOff[NGX];
	if constexpr (GM) {
#pragma unroll
		for (int g = 0; g < NG; ++g) {
			const uint32_t unit = (uint32_t)(g * kGmUnits + lane / 4);
			aOff[g] = min(unit, a.rowUnits - 1) * 16u + (uint32_t)(lane % 4) * 4u;
			const uint32_t sec = unit >= a.secUnit ? 1u : 0u;
			bOff[g] = (sec * 4u * NCS + (uint32_t)(lane % 4)) * (uint32_t)E;
#pragma unroll
			for (int c = 0; c < NCS; ++c)
#pragma unroll
				for (int i = 0; i < 4; ++i)
					GD[g][c][i] = 0.;
		}
	}
	double mant = 1., wsum = 0.;
	int expo = 0;

	// Every thread issues kMaxLoads loads per tile, so that the staging registers
	// are a fixed set.  The loads of the next tile are in flight while a tile is
	// computed (the fetch after the block's last tile re-reads the array's last
	// tile, so that every stage write waits for the same loads).
	f32x4 buf[kMaxLoads];
	const uint32_t step = gridDim.x;
	const uint32_t lastTile = a.nTiles ? a.nTiles - 1 : 0;
	uint32_t tile = blockIdx.x;
	if (tile < a.nTiles) {
		fetch_tile(a, tile, lastTile, ldOff, buf);
		const uint32_t nMine = (a.nTiles - 1 - tile) / step + 1;  // tiles of this workgroup
		for (uint32_t i = 0; i < nMine; ++i, tile += step) {
			lds_barrier();  // the previous tile's G phase is done with the stage
			store_tile(stage, ldsPos, buf);
			fetch_tile(a, min(tile + step, lastTile), lastTile, ldOff, buf);
			lds_barrier();
			double2 S[2][R];
			s_phase<MIXED, R, E>(a, stage, part, lane, wid, S);
			lds_barrier();
			if (wid == 0)
				sigma_phase<MIXED, R, E, kWaves, GM>(a, part, part, lane, tile, S, mant, expo, wsum);  // u replaces wave 1's partial
			lds_barrier();
			if constexpr (GM)
				gm_phase<R, E, NGX>(a, stage, reinterpret_cast<const float*>(part), wid, aOff, bOff, GD);
			else
				g_phase<MIXED, R, E>(a, stage, part, gOn, gU0, gSub, gSec, G);
		}
	}
	if (a.mode == STAGE_SIGMA)
		return;  // uniform: the whole workgroup leaves
	if constexpr (GM)
		write_partial_row_gm<R, E, NGX, kThreads>(a, smem, tid, lane, wid, GD, mant, expo, wsum);
	else
		write_partial_row<MIXED, R, E, kThreads>(a, smem, tid, lane, wid, gOn, gU0, gEnd, gSub, G, mant, expo, wsum);
}

// ---- The wave-tile variant (complex64, rank <= 2, rows of at most 16 NG units,
// tiles of 64 events): each wave stages its own 16 events of the tile and
// computes their complete S on v_mfma_f64_4x4x4f64, sigma and u inside the wave,
// and the gradient of those events on v_mfma_f32_4x4x1f32 -- no workgroup
// barrier in the tile loop (with the tile staged by the whole workgroup, the two
// barriers per tile cost 3.5% of the kernel time) and no single-wave step.
//
// S: 4 blocks of 4x4x4 per MFMA, K = the four floats of one 16-byte unit (Re, Im
// of its two waves), block b and row i = event 4b + i of the wave's 16, column
// j = (Re S, Im S) of rank 0, then rank 1.  Layout on gfx950: A[b][i][k] in lane
// 16k + 4b + i, B[b][k][j] in lane 16k + 4b + j, D[b][i][j] in lane 16i + 4b + j.
// B is T re-laid as a real 4x4 matrix per unit, the same for every block, kept
// in LDS after the stage for the whole kernel (16 doubles per unit):
//   Re S += Re x Re T - Im x Im T,  Im S += Re x Im T + Im x Re T.
// The two sectors have their own accumulators; sigma of an event is the sum of
// its four lanes' squares (a quad, combined by symmetric exchanges so that all
// four hold the same bits).
// S of 16 events over units [u0, u1) on v_mfma_f64_4x4x4f64 in four independent
// accumulator chains (units u0 + 4m + c go to chain c), added in a fixed order
// at the end: the chains keep the matrix pipe busy where one chain would wait
// for each result
__device__ __forceinline__ double s_mfma(const float* xa, const double* tLane, uint32_t u0, uint32_t u1)
{
	double c0 = 0., c1 = 0., c2 = 0., c3 = 0.;
	uint32_t u = u0;
	for (; u + 4 <= u1; u += 4) {
		const double x0 = (double)xa[u * 4], x1 = (double)xa[u * 4 + 4];
		const double x2 = (double)xa[u * 4 + 8], x3 = (double)xa[u * 4 + 12];
		const double t0 = tLane[u * 16], t1 = tLane[u * 16 + 16], t2 = tLane[u * 16 + 32], t3 = tLane[u * 16 + 48];
		c0 = __builtin_amdgcn_mfma_f64_4x4x4f64(x0, t0, c0, 0, 0, 0);
		c1 = __builtin_amdgcn_mfma_f64_4x4x4f64(x1, t1, c1, 0, 0, 0);
		c2 = __builtin_amdgcn_mfma_f64_4x4x4f64(x2, t2, c2, 0, 0, 0);
		c3 = __builtin_amdgcn_mfma_f64_4x4x4f64(x3, t3, c3, 0, 0, 0);
	}
	if (u < u1)
		c0 = __builtin_amdgcn_mfma_f64_4x4x4f64((double)xa[u * 4], tLane[u * 16], c0, 0, 0, 0);
	if (u + 1 < u1)
		c1 = __builtin_amdgcn_mfma_f64_4x4x4f64((double)xa[u * 4 + 4], tLane[u * 16 + 16], c1, 0, 0, 0);
	if (u + 2 < u1)
		c2 = __builtin_amdgcn_mfma_f64_4x4x4f64((double)xa[u * 4 + 8], tLane[u * 16 + 32], c2, 0, 0, 0);
	return (c0 + c1) + (c2 + c3);
}

// G += u conj(x) of 16 staged events (rows at rows + e * rowBytes) on
// v_mfma_f32_4x4x1f32, one accumulator chain per unit group, added into the
// FP64 sums GD at the end of the block
template <int NG>
__device__ __forceinline__ void g_mfma_block(const char* rows, uint32_t rowBytes, const float* uw,
                                             const uint32_t (&aOff)[NG], const uint32_t (&bOff)[NG],
                                             double (&GD)[NG][1][4])
{
	constexpr int EW = 16;
	gm_f4 acc[NG];
#pragma unroll
	for (int g = 0; g < NG; ++g)
		acc[g] = gm_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
	for (int q4 = 0; q4 < EW; q4 += 4) {
		gm_f4 bv[NG];
#pragma unroll
		for (int g = 0; g < NG; ++g)
			bv[g] = *reinterpret_cast<const gm_f4*>(uw + bOff[g] + q4);
#pragma unroll
		for (int kk = 0; kk < 4; ++kk) {
			const char* row = rows + (uint32_t)(q4 + kk) * rowBytes;
#pragma unroll
			for (int g = 0; g < NG; ++g) {
				const float av = *reinterpret_cast<const float*>(row + aOff[g]);
				acc[g] = __builtin_amdgcn_mfma_f32_4x4x1f32(av, bv[g][kk], acc[g], 0, 0, 0);
			}
		}
	}
#pragma unroll
	for (int g = 0; g < NG; ++g)
#pragma unroll
		for (int i = 0; i < 4; ++i)
			GD[g][0][i] += (double)acc[g][i];
}

// Per-wave staging of the wave-tile kernel: each wave stages only its own 16
// rows of a tile (16 rowUnits units, KL loads of 16 bytes per lane) into its
// rows of the stage and contracts them; no other wave reads them, so the tile
// loop needs no workgroup barrier (a wave's LDS operations complete in order).
template <int KL>
__device__ __forceinline__ void wave_load_plan(const StageArgs& a, int lane, int wid, uint32_t (&ldsPos)[KL])
{
	const uint32_t waveUnits = 16u * a.rowUnits;
	const uint32_t dumpUnit = 64u * a.rowUnitsP;
#pragma unroll
	for (int k = 0; k < KL; ++k) {
		const uint32_t g = (uint32_t)lane + 64u * k;
		const uint32_t row = g / a.rowUnits;
		ldsPos[k] = g < waveUnits ? (16u * (uint32_t)wid + row) * a.rowUnitsP + (g - row * a.rowUnits) : dumpUnit;
	}
}

template <int KL>
__device__ __forceinline__ void fetch_wave_lean(const StageArgs& a, uint32_t t, uint32_t lastTile, int lane, int wid,
                                                f32x4 (&buf)[KL])
{
	const uint32_t waveUnits = 16u * a.rowUnits;
	const uint64_t base = (uint64_t)t * a.tileUnits + (uint64_t)wid * waveUnits;
	if (t < lastTile) {
		const f32x4* tb = a.amps + base;
#pragma unroll
		for (int k = 0; k < KL; ++k) {
			const uint32_t g = (uint32_t)lane + 64u * k;
			buf[k] = __builtin_nontemporal_load(tb + (g < waveUnits ? g : 0u));
		}
	} else {
#pragma unroll
		for (int k = 0; k < KL; ++k) {
			const uint32_t g = (uint32_t)lane + 64u * k;
			const uint64_t gg = base + (g < waveUnits ? g : 0u);
			buf[k] = __builtin_nontemporal_load(a.amps + (gg < a.lastUnit ? gg : a.lastUnit));
		}
	}
}

template <int R, int NG, int KL>
__global__ __launch_bounds__(kThreads, 2) void wave_tile_kernel(StageArgs a)
{
	constexpr int E = 64;
	constexpr int NU = kGmUnits * NG;   // units of T' in LDS (rows of at most NU units)
	constexpr int EW = 16;              // events per wave
	extern __shared__ f32x4 smem[];
	__shared__ double lsum[kWaves][3];
	f32x4* stage = smem;  // [E][rowUnitsP] units, then the dump unit

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
	// u of every wave's events: [wave][sector][column][16] floats after the stage,
	// then T' [unit][k][j] doubles (B of the S MFMA)
	float* uw = reinterpret_cast<float*>(reinterpret_cast<char*>(smem) + a.stageBytes) + wid * (2 * 4 * EW);
	double* tL = reinterpret_cast<double*>(reinterpret_cast<char*>(smem) + a.stageBytes + kWaves * 2 * 4 * EW * 4);
	// KL loads of 16 bytes per thread and tile (64 rows of at most 4 KL units)
	uint32_t ldsPacked[(KL + 1) / 2];
	{
		uint32_t ldsPos[KL];
		wave_load_plan<KL>(a, lane, wid, ldsPos);
		pack_positions<KL>(ldsPos, ldsPacked);
	}
	{
		typedef const __attribute__((address_space(4))) char cchar;
		cdouble* tp = a.tInline ? (cdouble*)((cchar*)__builtin_amdgcn_kernarg_segment_ptr() + offsetof(StageArgs, tin))
		                        : (cdouble*)(const double*)a.tpad;
		for (uint32_t idx = (uint32_t)tid; idx < (uint32_t)NU * 16u; idx += kThreads) {
			const uint32_t u = idx / 16u, k = (idx / 4u) % 4u, j = idx % 4u;
			const uint32_t r = j >> 1;
			double v = 0.;
			if (u < a.rowUnits && r < (uint32_t)R) {
				const uint32_t col = 2u * u + (k >> 1);
				const double tre = tp[2 * (col * R + r)], tim = tp[2 * (col * R + r) + 1];
				if ((j & 1) == 0)
					v = (k & 1) == 0 ? tre : -tim;  // Re S
				else
					v = (k & 1) == 0 ? tim : tre;   // Im S
			}
			tL[idx] = v;
		}
	}
	// this lane's entry of T' in every unit: T'[k = lane / 16][j = lane % 4]
	const double* tLane = tL + (lane >> 4) * 4 + (lane & 3);
	// G MFMA: per unit group the lane's byte offset in a staged row and the
	// float offset of its B operand in uw (sector of the unit, column lane % 4)
	uint32_t aOff[NG], bOff[NG];
	double GD[NG][1][4];
#pragma unroll
	for (int g = 0; g < NG; ++g) {
		const uint32_t unit = (uint32_t)(g * kGmUnits + lane / 4);
		aOff[g] = min(unit, a.rowUnits - 1) * 16u + (uint32_t)(lane % 4) * 4u;
		const uint32_t sec = unit >= a.secUnit ? 1u : 0u;
		bOff[g] = (sec * 4u + (uint32_t)(lane % 4)) * (uint32_t)EW;
#pragma unroll
		for (int i = 0; i < 4; ++i)
			GD[g][0][i] = 0.;
	}
	double mant = 1., wsum = 0.;
	int expo = 0;
	const uint32_t rowBytes = a.rowUnitsP * 16u;
	// S operand A: event (lane % 16) of the wave's 16, float (lane / 16) of each unit
	const uint32_t aRow = ((uint32_t)(wid * EW + (lane & 15))) * rowBytes + (uint32_t)(lane >> 4) * 4u;
	// D layout: this lane's event of the wave's 16 and its column
	const int dEv = 4 * ((lane >> 2) & 3) + (lane >> 4);
	const int dCol = lane & 3;

	// two staging register sets: the loads of the next two tiles are in flight
	// while a tile is contracted (no fetch past the workgroup's last tile)
	f32x4 bufA[KL], bufB[KL];
	const uint32_t step = gridDim.x;
	const uint32_t lastTile = a.nTiles ? a.nTiles - 1 : 0;
	const uint32_t tile0 = blockIdx.x;
	const uint32_t nMine = tile0 < a.nTiles ? (a.nTiles - 1 - tile0) / step + 1 : 0;
	auto contract = [&](uint32_t tile) {
			const char* sb = reinterpret_cast<const char*>(stage);
			// ---- S of the wave's 16 events ----
			const float* xa = reinterpret_cast<const float*>(sb + aRow);
			const double s0 = s_mfma(xa, tLane, 0, a.secUnit);
			const double s1 = s_mfma(xa, tLane, a.secUnit, a.rowUnits);
			// ---- sigma, the log product, the weights and u, inside the wave ----
			double q = s0 * s0 + s1 * s1;
			q = q + __shfl_xor(q, 1);
			q = q + __shfl_xor(q, 2);
			const double sigma = q + a.flat2;
			const uint64_t ev = (uint64_t)tile * E + (uint64_t)(wid * EW + dEv);
			double w = 0.;
			if (ev < a.nEvents) {
				double rc = __builtin_amdgcn_rcp(sigma);
				rc = fma(fma(-sigma, rc, 1.), rc, rc);
				w = -2. * rc;
				if (dCol == 0) {
					mant *= sigma;
					const uint64_t ub = (uint64_t)__double_as_longlong(mant);
					const int ex = (int)((ub >> 52) & 0x7FF);
					if (ex != 0 && ex != 0x7FF) {
						expo += ex - 1023;
						mant = __longlong_as_double((long long)((ub & 0x800FFFFFFFFFFFFFull) | (1023ull << 52)));
					}
					wsum += w;
				}
			}
			uw[(0 * 4 + dCol) * EW + dEv] = (float)(w * s0);
			uw[(1 * 4 + dCol) * EW + dEv] = (float)(w * s1);
			// the wave's own u is read back below: LDS operations of a wave complete in order
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			// ---- G of the wave's 16 events on the matrix cores ----
			g_mfma_block<NG>(sb + (uint32_t)(wid * EW) * rowBytes, rowBytes, uw, aOff, bOff, GD);
	};
	if (nMine > 0)
		fetch_wave_lean<KL>(a, tile0, lastTile, lane, wid, bufA);
	if (nMine > 1)
		fetch_wave_lean<KL>(a, tile0 + step, lastTile, lane, wid, bufB);
	lds_barrier();  // T' of every unit is in LDS
	for (uint32_t it = 0; it < nMine; it += 2) {
		store_tile_packed<KL>(stage, ldsPacked, bufA);
		if (it + 2 < nMine)
			fetch_wave_lean<KL>(a, tile0 + (it + 2) * step, lastTile, lane, wid, bufA);
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		contract(tile0 + it * step);
		if (it + 1 < nMine) {
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			store_tile_packed<KL>(stage, ldsPacked, bufB);
			if (it + 3 < nMine)
				fetch_wave_lean<KL>(a, tile0 + (it + 3) * step, lastTile, lane, wid, bufB);
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			contract(tile0 + (it + 1) * step);
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	}
	write_partial_row_gm<R, E, NG, kThreads, true>(a, smem, tid, lane, wid, GD, mant, expo, wsum, lsum);
}

using StageKernel = void (*)(StageArgs);

template <bool MIXED, int E, int NG>
StageKernel pick_r(int rank)
{
	if (rank <= 1)
		return stage_kernel<MIXED, 1, E, NG>;
	if (rank <= 2)
		return stage_kernel<MIXED, 2, E, NG>;
	return stage_kernel<MIXED, 4, E, NG>;
}

template <bool MIXED>
StageKernel pick_e(int events, int rank)
{
	switch (events) {
	case 64: return pick_r<MIXED, 64, 0>(rank);
	case 32: return pick_r<MIXED, 32, 0>(rank);
	default: return pick_r<MIXED, 16, 0>(rank);
	}
}

template <int E>
StageKernel pick_gm(int groups, int rank)
{
	switch (groups) {
	case 1: return pick_r<true, E, 1>(rank);
	case 2: return pick_r<true, E, 2>(rank);
	case 3: return pick_r<true, E, 3>(rank);
	default: return pick_r<true, E, 4>(rank);
	}
}

template <int R, int NG>
StageKernel pick_wt_kl(int kl)
{
	// rows of 16 (NG - 1) + 1 .. 16 NG units: KL = ceil(rowUnits / 4) in 4 NG - 3 .. 4 NG
	switch (kl - 4 * NG) {
	case -3: return wave_tile_kernel<R, NG, 4 * NG - 3>;
	case -2: return wave_tile_kernel<R, NG, 4 * NG - 2>;
	case -1: return wave_tile_kernel<R, NG, 4 * NG - 1>;
	default: return wave_tile_kernel<R, NG, 4 * NG>;
	}
}

template <int R>
StageKernel pick_wt(int rowUnits)
{
	const int groups = (rowUnits + kGmUnits - 1) / kGmUnits;
	const int kl = (rowUnits + 3) / 4;
	switch (groups) {
	case 1: return pick_wt_kl<R, 1>(kl);
	case 2: return pick_wt_kl<R, 2>(kl);
	default: return pick_wt_kl<R, 3>(kl);
	}
}

StageKernel pick_stage(const AmpLayout& lay, int events, int gm, int rank)
{
	if (gm == 2) {
		const int rowUnits = (int)(lay.rowLen * 8u / 16u);
		return rank <= 1 ? pick_wt<1>(rowUnits) : pick_wt<2>(rowUnits);
	}
	if (gm) {
		const int groups = (int)((lay.rowLen * 8u / 16u + kGmUnits - 1) / kGmUnits);
		return events == 64 ? pick_gm<64>(groups, rank) : pick_gm<32>(groups, rank);
	}
	return lay.precision == RPWA_PREC_MIXED ? pick_e<true>(events, rank) : pick_e<false>(events, rank);
}

}  // namespace

int stage_rank_cap(int rank) { return rank <= 1 ? 1 : (rank <= 2 ? 2 : 4); }

bool stage_t_inline(const AmpLayout& lay, int rank)
{
	return (size_t)lay.rowLen * stage_rank_cap(rank) <= (size_t)kInlineT;
}

bool stage_plan(const AmpLayout& lay, int events, int rank, StagePlan& p, int gm)
{
	if (events != 16 && events != 32 && events != 64)
		return false;
	// the MFMA gradient: complex64 rows of at most 64 units, tiles of 64 or 32 events;
	// the wave-tile variant (S on MFMA too): rows of at most 48 units, rank <= 2, 64 events
	if (gm && (lay.precision != RPWA_PREC_MIXED || events == 16 || lay.rowLen * 8u > 16u * kGmUnits * kGmMaxGroups))
		return false;
	if (gm == 2 && (events != 64 || rank > 2 || lay.rowLen * 8u > 16u * kGmUnits * 3u))
		return false;
	if (gm < 0 || gm > 2)
		return false;
	p.gm = gm;
	const int threads = kThreads;
	const int waves = kWaves;
	p.threads = threads;
	const uint32_t bpc = lay.precision == RPWA_PREC_MIXED ? 8u : 16u;
	const uint32_t P = 16u / bpc;
	p.events = events;
	p.rowUnits = lay.rowLen * bpc / 16u;
	if (p.rowUnits == 0 || p.rowUnits > (uint32_t)kThreads)
		return false;
	p.sliceUnits = (p.rowUnits + waves - 1) / waves;
	const uint32_t tileUnits = (uint32_t)events * p.rowUnits;
	if (tileUnits > (uint32_t)(kThreads * kMaxLoads))
		return false;
	p.rowUnitsP = p.rowUnits | 1u;
	p.ranks = stage_rank_cap(rank);
	const bool mixed = lay.precision == RPWA_PREC_MIXED;
	const uint32_t ug = (uint32_t)(mixed ? (p.ranks == 1 ? group_units<true, 1>() : p.ranks == 2 ? group_units<true, 2>()
	                                                                                     : group_units<true, 4>())
	                                     : (p.ranks == 1 ? group_units<false, 1>() : p.ranks == 2 ? group_units<false, 2>()
	                                                                                     : group_units<false, 4>()));
	const uint32_t secUnit = lay.secOff[1] * bpc / 16u;
	p.gGroups0 = (secUnit + ug - 1) / ug;
	p.gGroups = p.gGroups0 + (p.rowUnits - secUnit + ug - 1) / ug;
	if (p.gGroups > (uint32_t)threads)
		return false;
	p.gSubs = (uint32_t)threads / p.gGroups;
	p.stageBytes = ((uint32_t)events * p.rowUnitsP + 1u) * 16u;
	const int ps = p.ranks == 1 ? part_stride<1>() : p.ranks == 2 ? part_stride<2>() : part_stride<4>();
	// S partials of waves 1..3 (u replaces wave 1's)
	const size_t partBytes = (size_t)(waves - 1) * events * ps * sizeof(double2);
	const size_t endBytes = (size_t)p.gSubs * p.rowUnits * P * p.ranks * sizeof(double2) +
	                        (size_t)events * 3 * sizeof(double);
	p.ldsBytes = std::max((size_t)p.stageBytes + partBytes, endBytes);
	if (gm) {
		// the waves' FP64 sums [wave][units][columns][4] at the end
		const uint32_t groups = (p.rowUnits + kGmUnits - 1) / kGmUnits;
		const int ncs = p.ranks == 1 ? gm_col_sets<1>() : p.ranks == 2 ? gm_col_sets<2>() : gm_col_sets<4>();
		const size_t gmBytes = (size_t)waves * groups * kGmUnits * 4 * ncs * 4 * sizeof(double);
		p.ldsBytes = std::max(p.ldsBytes, gmBytes);
	}
	if (gm == 2) {  // the stage, every wave's u [2][4][16] floats, T' [units][4][4] doubles
		const uint32_t units = (p.rowUnits + kGmUnits - 1) / kGmUnits * kGmUnits;
		p.ldsBytes = std::max(p.ldsBytes, (size_t)p.stageBytes + (size_t)waves * 2 * 4 * 16 * sizeof(float) +
		                                      (size_t)units * 16 * sizeof(double));
	}

	return p.ldsBytes <= 160u * 1024u;
}

hipError_t launch_stage_kernel(const AmpLayout& lay, int events, int gm, const EventArgs& ea, int nWg, hipStream_t s,
                               hipEvent_t start, hipEvent_t stop)
{
	StagePlan p;
	if (!stage_plan(lay, events, ea.rank, p, gm))
		return hipErrorInvalidValue;
	StageArgs a;
	a.amps = reinterpret_cast<const f32x4*>(ea.amps);
	a.mode = ea.stageMode;
	a.rank0 = ea.rank0;
	a.accumulate = ea.accumulate;
	a.sigma = ea.sigma;
	a.nEvents = lay.nEvents;
	a.lastUnit = lay.nEvents ? lay.nEvents * p.rowUnits - 1 : 0;  // no tile, no load when there are no events
	a.nTiles = (uint32_t)((lay.nEvents + events - 1) / events);
	a.rowUnits = p.rowUnits;
	a.rowUnitsP = p.rowUnitsP;
	a.tileUnits = (uint32_t)events * p.rowUnits;
	a.sliceUnits = p.sliceUnits;
	const uint32_t bpc = lay.precision == RPWA_PREC_MIXED ? 8u : 16u;
	a.secUnit = lay.secOff[1] * bpc / 16u;
	a.gGroups0 = p.gGroups0;
	a.gGroups = p.gGroups;
	a.gSubs = p.gSubs;
	a.gmGroups = (p.rowUnits + kGmUnits - 1) / kGmUnits;
	a.stageBytes = p.stageBytes;
	a.secOff1 = lay.secOff[1];
	a.nW0 = lay.nW[0];
	a.nW1 = lay.nW[1];
	a.wmax = lay.wmax;
	a.tpad = ea.tpad;
	a.flat2 = ea.flat2;
	a.partials = ea.partials;
	a.nElem = ea.nElem;
	a.tInline = 0;
	if (ea.tpadHost && stage_t_inline(lay, ea.rank)) {
		a.tInline = 1;
		const size_t n = (size_t)lay.rowLen * stage_rank_cap(ea.rank);
		for (size_t i = 0; i < n; ++i)
			a.tin[i] = ea.tpadHost[i];
	}
	StageKernel k = pick_stage(lay, events, gm, ea.rank);
	if (p.ldsBytes > 64u * 1024u) {
		hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
		                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.ldsBytes);
		if (e != hipSuccess)
			return e;
	}
	if (start || stop)
		hipExtLaunchKernelGGL(k, dim3(nWg), dim3(p.threads), (uint32_t)p.ldsBytes, s, start, stop, 0u, a);
	else
		hipLaunchKernelGGL(k, dim3(nWg), dim3(p.threads), p.ldsBytes, s, a);
	return hipGetLastError();
}

int stage_kernel_occupancy(const AmpLayout& lay, int events, int gm, int rank, int* wgPerCu)
{
	StagePlan p;
	if (!stage_plan(lay, events, rank, p, gm))
		return RPWA_ERR_UNSUPPORTED;
	StageKernel k = pick_stage(lay, events, gm, rank);
	if (p.ldsBytes > 64u * 1024u)
		if (hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
		                        (int)p.ldsBytes) != hipSuccess)
			return RPWA_ERR_HIP;
	int n = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, p.threads, p.ldsBytes) != hipSuccess)
		return RPWA_ERR_HIP;
	*wgPerCu = n > 0 ? n : 1;
	return RPWA_OK;
}

}  // namespace rpwa_amd
