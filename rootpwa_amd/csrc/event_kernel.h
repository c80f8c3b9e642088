// event_kernel.h -- the fused event kernel template (see lik_kernels.hip for
// the algorithm); instantiated per (precision, lanes) in event_kernels_*.hip so
// the variants compile in parallel.
#pragma once

#include "internal.h"

namespace rpwa_amd {
namespace ek {

constexpr int kWgThreads = 256;
constexpr int kWavesPerWg = kWgThreads / 64;

__device__ __forceinline__ double shfl_xor_d(double v, int mask)
{
	return __shfl_xor(v, mask, 64);
}

// Neumaier step: (s, c) += x with the rounding error of the addition kept in c
__device__ __forceinline__ void neumaier_add(double& s, double& c, double x)
{
	const double t = s + x;
	if (fabs(s) >= fabs(x))
		c += (s - t) + x;
	else
		c += (x - t) + s;
	s = t;
}

// acc += t * x
__device__ __forceinline__ void cmac(double2& acc, const double2 t, const double2 x)
{
	acc.x = fma(t.x, x.x, acc.x);
	acc.x = fma(-t.y, x.y, acc.x);
	acc.y = fma(t.x, x.y, acc.y);
	acc.y = fma(t.y, x.x, acc.y);
}

// acc += u * conj(x)
__device__ __forceinline__ void cmac_conj(double2& acc, const double2 u, const double2 x)
{
	acc.x = fma(u.x, x.x, acc.x);
	acc.x = fma(u.y, x.y, acc.x);
	acc.y = fma(u.y, x.x, acc.y);
	acc.y = fma(-u.x, x.y, acc.y);
}

template <bool MIXED>
struct SlotData;

// Amplitude loads are non-temporal: the array is streamed once per evaluation
// and is far larger than the caches (streaming reads reach ~7 TB/s this way on
// an MI355X, ~6.2 TB/s through the caches).
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));

template <>
struct SlotData<false> {
	double2 v;
	__device__ __forceinline__ void zero() { v = make_double2(0., 0.); }
	__device__ __forceinline__ void load(const void* p)
	{
		const nt_d2 x = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
		v = make_double2(x.x, x.y);
	}
	__device__ __forceinline__ double2 get(int) const { return v; }
};

template <>
struct SlotData<true> {
	float4 v;
	__device__ __forceinline__ void zero() { v = make_float4(0.f, 0.f, 0.f, 0.f); }
	__device__ __forceinline__ void load(const void* p)
	{
		const nt_f4 x = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
		v = make_float4(x.x, x.y, x.z, x.w);
	}
	__device__ __forceinline__ double2 get(int p) const
	{
		return p == 0 ? make_double2((double)v.x, (double)v.y) : make_double2((double)v.z, (double)v.w);
	}
};

// a double moved across lanes with a DPP row permutation (both 32-bit halves)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
	const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
	const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
	return __hiloint2double(hi, lo);
}

// v + (v of the partner lane) for the butterfly step that merges blocks of
// M lanes inside a group: quad_perm swaps, then the half-row and row mirrors
// (DPP, no LDS traffic), then cross-row exchanges.  Each step pairs lanes
// symmetrically, so both partners compute the same sum bit for bit.
template <int M>
__device__ __forceinline__ double bfly_add(double v)
{
	if (M == 1)
		return v + dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
	if (M == 2)
		return v + dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
	if (M == 4)
		return v + dpp_d<0x141>(v);  // row_half_mirror
	if (M == 8)
		return v + dpp_d<0x140>(v);  // row_mirror
	return v + __shfl_xor(v, M, 64);
}

template <int L>
__device__ __forceinline__ double group_sum(double v)
{
	if (L >= 2) v = bfly_add<1>(v);
	if (L >= 4) v = bfly_add<2>(v);
	if (L >= 8) v = bfly_add<4>(v);
	if (L >= 16) v = bfly_add<8>(v);
	if (L >= 32) v = bfly_add<16>(v);
	if (L >= 64) v = bfly_add<32>(v);
	return v;
}

// Tile: a group of L lanes owns one event.  The first lanes0 lanes hold the
// refl -1 sector, the next lanes1 lanes the refl +1 sector (a lane never mixes
// sectors); lane ls of a sector holds K slots of P consecutive waves,
// wave(k, p) = (ls + lanesOfSector * k) * P + p, so each slot load of a sector
// is one contiguous row segment.  R ranks are held in registers (T = 0 for
// ranks beyond the requested one).
template <bool MIXED, int L, int K, int R>
__global__ __launch_bounds__(kWgThreads) void event_kernel(EventArgs a)
{
	constexpr int P = MIXED ? 2 : 1;
	constexpr int GPW = 64 / L;             // events (groups) per wavefront
	constexpr size_t BPC = MIXED ? 8 : 16;  // bytes per complex
	extern __shared__ double2 lds[];        // [L][R][K][P]

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wid = tid >> 6;
	const int lig = lane & (L - 1);
	const int grp = lane / L;

	const int sec = lig >= a.lanes0 ? 1 : 0;
	const int ls = sec ? lig - a.lanes0 : lig;
	const int laneStride = sec ? a.lanes1 : a.lanes0;
	const int nw = (sec ? (int)a.nW1 : (int)a.nW0) * ((sec ? ls < a.lanes1 : true) ? 1 : 0);
	const uint32_t secOff = sec ? a.secOff1 : a.secOff0;

	int nValid[K];
	uint32_t slotOff[K];
#pragma unroll
	for (int k = 0; k < K; ++k) {
		const int j = (ls + laneStride * k) * P;
		int nv = nw - j;
		nv = nv < 0 ? 0 : (nv > P ? P : nv);
		nValid[k] = nv;
		slotOff[k] = secOff + (uint32_t)(nv > 0 ? j : 0);
	}

	double2 T[R][K][P];
	double2 G[R][K][P];
#pragma unroll
	for (int r = 0; r < R; ++r)
#pragma unroll
		for (int k = 0; k < K; ++k)
#pragma unroll
			for (int p = 0; p < P; ++p) {
				double2 t = make_double2(0., 0.);
				if (r < a.rank && p < nValid[k])
					t = a.prod[((size_t)r * 2 + sec) * a.wmax + (ls + laneStride * k) * P + p];
				T[r][k][p] = t;
				G[r][k][p] = make_double2(0., 0.);
			}

	double mant = 1.;  // running product of sigma_e as mant 2^expo, mant in [1, 2) (group leaders);
	int expo = 0;      // its log is taken once at the end instead of one log per event
	double wsum = 0.;           // sum of -2/sigma_e (group leaders)
	const uint64_t nWavesGrid = (uint64_t)gridDim.x * kWavesPerWg;
	const uint64_t rowBytes = (uint64_t)a.rowLen * BPC;
	const char* base = reinterpret_cast<const char*>(a.amps);
	const uint64_t stride = nWavesGrid * GPW;

	auto load_slots = [&](SlotData<MIXED>* dst, uint64_t e) {
		const bool ok = e < a.nEvents;
		const char* row = base + e * rowBytes;
#pragma unroll
		for (int k = 0; k < K; ++k) {
			if (ok && nValid[k] > 0)
				dst[k].load(row + (size_t)slotOff[k] * BPC);
			else
				dst[k].zero();
		}
	};
	const uint64_t eFirst = ((uint64_t)blockIdx.x * kWavesPerWg + wid) * GPW;
	SlotData<MIXED> An[K];  // slots of the next event group, loaded one iteration ahead
	load_slots(An, eFirst + grp);
	for (uint64_t e0 = eFirst; e0 < a.nEvents; e0 += stride) {
		const uint64_t e = e0 + grp;
		const bool ev = e < a.nEvents;
		SlotData<MIXED> A[K];
#pragma unroll
		for (int k = 0; k < K; ++k)
			A[k] = An[k];
		if (e0 + stride < a.nEvents)
			load_slots(An, e + stride);

		// this lane's part of S_r(e) for its sector
		double2 Sa[R];
#pragma unroll
		for (int r = 0; r < R; ++r)
			Sa[r] = make_double2(0., 0.);
#pragma unroll
		for (int k = 0; k < K; ++k)
#pragma unroll
			for (int p = 0; p < P; ++p) {
				const double2 x = A[k].get(p);
#pragma unroll
				for (int r = 0; r < R; ++r)
					cmac(Sa[r], T[r][k][p], x);
			}
		// sector sums over the group: every lane ends with identical bits
		double2 S0[R], S1[R];
#pragma unroll
		for (int r = 0; r < R; ++r) {
			S0[r].x = group_sum<L>(sec ? 0. : Sa[r].x);
			S0[r].y = group_sum<L>(sec ? 0. : Sa[r].y);
			S1[r].x = group_sum<L>(sec ? Sa[r].x : 0.);
			S1[r].y = group_sum<L>(sec ? Sa[r].y : 0.);
		}
		double sigma = 0.;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			sigma += S0[r].x * S0[r].x + S0[r].y * S0[r].y;
			sigma += S1[r].x * S1[r].x + S1[r].y * S1[r].y;
		}
		sigma += a.flat2;
		double w = 0.;
		if (ev) {
			if (lig == 0) {
				mant *= sigma;
				const uint64_t ub = (uint64_t)__double_as_longlong(mant);
				const int ex = (int)((ub >> 52) & 0x7FF);
				if (ex != 0 && ex != 0x7FF) {
					expo += ex - 1023;
					mant = __longlong_as_double((long long)((ub & 0x800FFFFFFFFFFFFFull) | (1023ull << 52)));
				}
			}
			// -2/sigma from the hardware reciprocal and one Newton step
			double r = __builtin_amdgcn_rcp(sigma);
			r = fma(fma(-sigma, r, 1.), r, r);
			w = -2. * r;
			if (lig == 0)
				wsum += w;
		}
		// gradient: G_i += (w S_sector) conj(A_i)
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const double2 Ssec = sec ? S1[r] : S0[r];
			const double2 u = make_double2(w * Ssec.x, w * Ssec.y);
#pragma unroll
			for (int k = 0; k < K; ++k)
#pragma unroll
				for (int p = 0; p < P; ++p)
					cmac_conj(G[r][k][p], u, A[k].get(p));
		}
	}

	// ---- epilogue: wavefront butterfly over the groups, then the workgroup ----
#pragma unroll
	for (int m = L; m < 64; m <<= 1)
#pragma unroll
		for (int r = 0; r < R; ++r)
#pragma unroll
			for (int k = 0; k < K; ++k)
#pragma unroll
				for (int p = 0; p < P; ++p) {
					G[r][k][p].x += shfl_xor_d(G[r][k][p].x, m);
					G[r][k][p].y += shfl_xor_d(G[r][k][p].y, m);
				}
	__shared__ double scal[kWavesPerWg][GPW][3];
	if (lig == 0) {
		scal[wid][grp][0] = -(log(mant) + (double)expo * 0.69314718055994530942);
		scal[wid][grp][1] = 0.;
		scal[wid][grp][2] = wsum;
	}
	// waves add their gradient sums into LDS in fixed order 0,1,2,3
	for (int wv = 0; wv < kWavesPerWg; ++wv) {
		if (wid == wv && grp == 0) {
#pragma unroll
			for (int r = 0; r < R; ++r)
#pragma unroll
				for (int k = 0; k < K; ++k)
#pragma unroll
					for (int p = 0; p < P; ++p) {
						double2* d = &lds[((lig * R + r) * K + k) * P + p];
						if (wv == 0) {
							*d = G[r][k][p];
						} else {
							double2 v = *d;
							v.x += G[r][k][p].x;
							v.y += G[r][k][p].y;
							*d = v;
						}
					}
		}
		__syncthreads();
	}
	double* out = a.partials + (size_t)blockIdx.x * a.nElem;
	if (tid == 0) {
		double s = 0., c = 0., ws = 0.;
		for (int wv = 0; wv < kWavesPerWg; ++wv)
			for (int g = 0; g < GPW; ++g) {
				neumaier_add(s, c, scal[wv][g][0]);
				neumaier_add(s, c, scal[wv][g][1]);
				ws += scal[wv][g][2];
			}
		out[0] = s;
		out[1] = c;
		out[2] = ws;
	}
	// element 3 + 2*((r*2 + s)*wmax + j) (+1 for the imaginary part)
	const uint32_t nCplx = (a.nElem - 3) / 2;
	for (uint32_t idx = tid; idx < nCplx; idx += kWgThreads) {
		const uint32_t j = idx % a.wmax;
		const uint32_t rs = idx / a.wmax;
		const int s = (int)(rs & 1);
		const int r = (int)(rs >> 1);
		double2 v = make_double2(0., 0.);
		const uint32_t nws = s ? a.nW1 : a.nW0;
		if (r < a.rank && r < R && j < nws) {
			const int lsn = s ? a.lanes1 : a.lanes0;
			const int q = (int)(j / P);
			const int p = (int)(j % P);
			const int k = q / lsn;
			const int l = (s ? a.lanes0 : 0) + q % lsn;
			v = lds[((l * R + r) * K + k) * P + p];
		}
		out[3 + 2 * idx] = v.x;
		out[3 + 2 * idx + 1] = v.y;
	}
}

}  // namespace ek

using EventKernel = void (*)(EventArgs);

// kernel for (slots, ranks) of one (precision, lanes) family
#define RPWA_DECLARE_EVENT_FAMILY(NAME) EventKernel NAME(int slots, int ranks)
RPWA_DECLARE_EVENT_FAMILY(event_family_f64_l8);
RPWA_DECLARE_EVENT_FAMILY(event_family_f64_l16);
RPWA_DECLARE_EVENT_FAMILY(event_family_f64_l32);
RPWA_DECLARE_EVENT_FAMILY(event_family_f64_l64);
RPWA_DECLARE_EVENT_FAMILY(event_family_c64_l8);
RPWA_DECLARE_EVENT_FAMILY(event_family_c64_l16);
RPWA_DECLARE_EVENT_FAMILY(event_family_c64_l32);
RPWA_DECLARE_EVENT_FAMILY(event_family_c64_l64);

}  // namespace rpwa_amd
