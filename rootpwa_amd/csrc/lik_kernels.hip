// lik_kernels.hip -- gfx950 event kernels of the unbinned extended likelihood.
//
// One pass over the decay amplitudes computes, for every event e,
//   S_r,s(e) = sum_i T_r,s,i A_s,i(e)            (sector s = reflectivity)
//   sigma_e  = sum_r sum_s |S_r,s(e)|^2 + flat^2
// and accumulates  sum_e -ln sigma_e  (compensated) and the event part of the
// gradient  G_r,s,i = sum_e (-2/sigma_e) S_r,s(e) conj(A_s,i(e)),  i.e. the
// quantities of pwaLikelihood::FdF's event loop (partialWaveFit/pwaLikelihood.cc:213-244)
// and of cuda/likelihoodKernel.cu:84-259, fused into one kernel.
//
// Work decomposition (single fit start: the problem is HBM-bound, 2-4 flop/B):
//   * a group of L lanes of a 64-wide wavefront owns one event; lane l of the
//     group holds `slots` slots of the event's amplitude row in registers, each
//     slot P complex values of one sector (P = 1 for complex128, 2 for complex64,
//     so every lane load is 16 bytes and a group reads a contiguous row segment);
//   * the production amplitudes T of the lane's slots live in registers;
//   * S is reduced across the L lanes with a butterfly (all lanes end with the
//     same bits), sigma and the log are computed once per event group;
//   * the amplitudes are still in registers for the gradient term, so every
//     amplitude byte is read from HBM exactly once per evaluation;
//   * each lane keeps its gradient accumulators across the grid-stride loop; the
//     epilogue reduces them across the wavefront (butterfly) and the workgroup
//     (LDS, fixed wave order) into one partial row per workgroup, and
//     event_reduce sums the rows in fixed order.
// The grid is fixed per context, so every reduction has a fixed order: the
// value is bitwise reproducible, identical between value-only and value+gradient
// calls (same kernel), and invariant under T -> -T of any (rank, sector) block.
// No floating-point atomics are used.
#include "internal.h"

namespace rpwa_amd {
namespace {

constexpr int kWgThreads = 256;
constexpr int kWavesPerWg = kWgThreads / 64;

__device__ __forceinline__ double shfl_xor_d(double v, int mask)
{
	return __shfl_xor(v, mask, 64);
}

// compensated accumulation of one term (utilities/sumAccumulators.hpp:166-177)
__device__ __forceinline__ void kahan_add(double& s, double& c, double x)
{
	const double y = x - c;
	const double t = s + y;
	c = (t - s) - y;
	s = t;
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

template <>
struct SlotData<false> {
	double2 v;
	__device__ __forceinline__ void zero() { v = make_double2(0., 0.); }
	__device__ __forceinline__ void load(const void* p) { v = *reinterpret_cast<const double2*>(p); }
	__device__ __forceinline__ double2 get(int) const { return v; }
};

template <>
struct SlotData<true> {
	float4 v;
	__device__ __forceinline__ void zero() { v = make_float4(0.f, 0.f, 0.f, 0.f); }
	__device__ __forceinline__ void load(const void* p) { v = *reinterpret_cast<const float4*>(p); }
	__device__ __forceinline__ double2 get(int p) const
	{
		return p == 0 ? make_double2((double)v.x, (double)v.y) : make_double2((double)v.z, (double)v.w);
	}
};

template <bool MIXED, int L, int KMAX, int RMAX>
__global__ __launch_bounds__(kWgThreads) void event_kernel(EventArgs a)
{
	constexpr int P = MIXED ? 2 : 1;
	constexpr int GPW = 64 / L;             // events (groups) per wavefront
	constexpr size_t BPC = MIXED ? 8 : 16;  // bytes per complex
	extern __shared__ double2 lds[];        // [L][RMAX][KMAX][P]

	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wid = tid >> 6;
	const int lig = lane & (L - 1);
	const int grp = lane / L;

	// slot geometry: slot k belongs to sector 1 iff k >= slot1
	int slotWave[KMAX];   // first wave index (within its sector) of the lane's slot
	int slotValid[KMAX];  // number of valid complex values in the lane's slot (0..P)
	uint32_t slotOff[KMAX];
#pragma unroll
	for (int k = 0; k < KMAX; ++k) {
		const int s = k >= a.slot1 ? 1 : 0;
		const int ks = s ? k - a.slot1 : k;
		const int j = (ks * L + lig) * P;
		const int nw = s ? (int)a.nW1 : (int)a.nW0;
		int nv = nw - j;
		nv = nv < 0 ? 0 : (nv > P ? P : nv);
		if (k >= a.slots)
			nv = 0;
		slotWave[k] = j;
		slotValid[k] = nv;
		slotOff[k] = (s ? a.secOff1 : a.secOff0) + (uint32_t)j;
	}

	// production amplitudes of the lane's slots
	double2 T[RMAX][KMAX][P];
	double2 G[RMAX][KMAX][P];
#pragma unroll
	for (int r = 0; r < RMAX; ++r)
#pragma unroll
		for (int k = 0; k < KMAX; ++k)
#pragma unroll
			for (int p = 0; p < P; ++p) {
				const int s = k >= a.slot1 ? 1 : 0;
				double2 t = make_double2(0., 0.);
				if (r < a.rank && p < slotValid[k])
					t = a.prod[((size_t)r * 2 + s) * a.wmax + slotWave[k] + p];
				T[r][k][p] = t;
				G[r][k][p] = make_double2(0., 0.);
			}

	double lnS = 0., lnC = 0.;  // compensated sum of -ln sigma_e (group leaders)
	double wsum = 0.;           // sum of -2/sigma_e (group leaders)
	const uint64_t nWavesGrid = (uint64_t)gridDim.x * kWavesPerWg;
	const uint64_t rowBytes = (uint64_t)a.rowLen * BPC;
	const char* base = reinterpret_cast<const char*>(a.amps);

	for (uint64_t e0 = ((uint64_t)blockIdx.x * kWavesPerWg + wid) * GPW; e0 < a.nEvents;
	     e0 += nWavesGrid * GPW) {
		const uint64_t e = e0 + grp;
		const bool ev = e < a.nEvents;
		SlotData<MIXED> A[KMAX];
		const char* row = base + e * rowBytes;
#pragma unroll
		for (int k = 0; k < KMAX; ++k) {
			if (ev && slotValid[k] > 0)
				A[k].load(row + (size_t)slotOff[k] * BPC);
			else
				A[k].zero();
		}
		// S = sum_i T_i A_i for the lane's slots
		double2 S[RMAX][2];
#pragma unroll
		for (int r = 0; r < RMAX; ++r) {
			S[r][0] = make_double2(0., 0.);
			S[r][1] = make_double2(0., 0.);
		}
#pragma unroll
		for (int k = 0; k < KMAX; ++k) {
			if (k >= a.slots)
				break;
			// the sector of slot k is uniform across the wavefront: a scalar branch
			// keeps every register index a compile-time constant
			if (k < a.slot1) {
#pragma unroll
				for (int p = 0; p < P; ++p)
#pragma unroll
					for (int r = 0; r < RMAX; ++r)
						if (r < a.rank)
							cmac(S[r][0], T[r][k][p], A[k].get(p));
			} else {
#pragma unroll
				for (int p = 0; p < P; ++p)
#pragma unroll
					for (int r = 0; r < RMAX; ++r)
						if (r < a.rank)
							cmac(S[r][1], T[r][k][p], A[k].get(p));
			}
		}
		// butterfly over the L lanes of the group: every lane gets identical sums
#pragma unroll
		for (int m = 1; m < L; m <<= 1)
#pragma unroll
			for (int r = 0; r < RMAX; ++r) {
				if (r >= a.rank)
					break;
#pragma unroll
				for (int s = 0; s < 2; ++s) {
					S[r][s].x += shfl_xor_d(S[r][s].x, m);
					S[r][s].y += shfl_xor_d(S[r][s].y, m);
				}
			}
		double sigma = 0.;
#pragma unroll
		for (int r = 0; r < RMAX; ++r) {
			if (r >= a.rank)
				break;
#pragma unroll
			for (int s = 0; s < 2; ++s)
				sigma += S[r][s].x * S[r][s].x + S[r][s].y * S[r][s].y;
		}
		sigma += a.flat2;
		double w = 0.;
		if (ev) {
			if (lig == 0)
				kahan_add(lnS, lnC, -log(sigma));
			w = -2. / sigma;
			if (lig == 0)
				wsum += w;
		}
		// gradient: G_i += (w S) conj(A_i)
		double2 U[RMAX][2];
#pragma unroll
		for (int r = 0; r < RMAX; ++r)
#pragma unroll
			for (int s = 0; s < 2; ++s)
				U[r][s] = make_double2(w * S[r][s].x, w * S[r][s].y);
#pragma unroll
		for (int k = 0; k < KMAX; ++k) {
			if (k >= a.slots)
				break;
			if (k < a.slot1) {
#pragma unroll
				for (int p = 0; p < P; ++p)
#pragma unroll
					for (int r = 0; r < RMAX; ++r)
						if (r < a.rank)
							cmac_conj(G[r][k][p], U[r][0], A[k].get(p));
			} else {
#pragma unroll
				for (int p = 0; p < P; ++p)
#pragma unroll
					for (int r = 0; r < RMAX; ++r)
						if (r < a.rank)
							cmac_conj(G[r][k][p], U[r][1], A[k].get(p));
			}
		}
	}

	// ---- epilogue: wavefront butterfly over the groups, then the workgroup ----
#pragma unroll
	for (int m = L; m < 64; m <<= 1)
#pragma unroll
		for (int r = 0; r < RMAX; ++r) {
			if (r >= a.rank)
				break;
#pragma unroll
			for (int k = 0; k < KMAX; ++k) {
				if (k >= a.slots)
					break;
#pragma unroll
				for (int p = 0; p < P; ++p) {
					G[r][k][p].x += shfl_xor_d(G[r][k][p].x, m);
					G[r][k][p].y += shfl_xor_d(G[r][k][p].y, m);
				}
			}
		}
	__shared__ double scal[kWavesPerWg][GPW][3];
	if (lig == 0) {
		scal[wid][grp][0] = lnS;
		scal[wid][grp][1] = -lnC;  // Kahan keeps c = (rounded - exact); value = s - c
		scal[wid][grp][2] = wsum;
	}
	// waves add their gradient sums into LDS in fixed order 0,1,2,3
	for (int wv = 0; wv < kWavesPerWg; ++wv) {
		if (wid == wv && grp == 0) {
#pragma unroll
			for (int r = 0; r < RMAX; ++r)
#pragma unroll
				for (int k = 0; k < KMAX; ++k)
#pragma unroll
					for (int p = 0; p < P; ++p) {
						double2* d = &lds[((lig * RMAX + r) * KMAX + k) * P + p];
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
		const uint32_t nw = s ? a.nW1 : a.nW0;
		if (r < a.rank && j < nw) {
			const int ks = (int)(j / (uint32_t)(L * P));
			const int rem = (int)(j % (uint32_t)(L * P));
			const int l = rem / P;
			const int p = rem % P;
			const int k = ks + (s ? a.slot1 : 0);
			v = lds[((l * RMAX + r) * KMAX + k) * P + p];
		}
		out[3 + 2 * idx] = v.x;
		out[3 + 2 * idx + 1] = v.y;
	}
}

// Second stage: fixed-order sum of the workgroup partial rows.
// Block = 64 elements x 16 chunks of workgroups; chunks are combined in order.
constexpr int kRedElems = 64;
constexpr int kRedChunks = 16;

__global__ __launch_bounds__(kRedElems* kRedChunks) void event_reduce_kernel(
    const double* __restrict__ partials, int nWg, uint32_t nElem, double flat,
    double* __restrict__ out)
{
	__shared__ double sh[kRedChunks][kRedElems];
	__shared__ double shc[kRedChunks][kRedElems];
	const int el = threadIdx.x % kRedElems;
	const int ch = threadIdx.x / kRedElems;
	const uint32_t elem = blockIdx.x * kRedElems + el;
	const int w0 = (int)((int64_t)nWg * ch / kRedChunks);
	const int w1 = (int)((int64_t)nWg * (ch + 1) / kRedChunks);
	double s = 0., c = 0.;
	if (elem < nElem) {
		if (elem < 2) {
			// lnL: compensated over both the sums and their corrections
			for (int w = w0; w < w1; ++w) {
				neumaier_add(s, c, partials[(size_t)w * nElem + 0]);
				neumaier_add(s, c, partials[(size_t)w * nElem + 1]);
			}
		} else {
			int w = w0;
			for (; w + 4 <= w1; w += 4) {
				const double v0 = partials[(size_t)(w + 0) * nElem + elem];
				const double v1 = partials[(size_t)(w + 1) * nElem + elem];
				const double v2 = partials[(size_t)(w + 2) * nElem + elem];
				const double v3 = partials[(size_t)(w + 3) * nElem + elem];
				s += v0;
				s += v1;
				s += v2;
				s += v3;
			}
			for (; w < w1; ++w)
				s += partials[(size_t)w * nElem + elem];
		}
	}
	sh[ch][el] = s;
	shc[ch][el] = c;
	__syncthreads();
	if (ch == 0 && elem < nElem) {
		if (elem < 2) {
			double ts = 0., tc = 0.;
			for (int k = 0; k < kRedChunks; ++k) {
				neumaier_add(ts, tc, sh[k][el]);
				neumaier_add(ts, tc, shc[k][el]);
			}
			if (elem == 0)
				out[0] = ts + tc;
		} else {
			double t = 0.;
			for (int k = 0; k < kRedChunks; ++k)
				t += sh[k][el];
			// out[1] = flat * sum_e(-2/sigma_e); out[2..] = derivatives
			if (elem == 2)
				out[1] = flat * t;
			else
				out[elem - 1] = t;
		}
	}
}

// Amplitude layout transform: source FP64 (either host layout) -> device rows.
template <bool MIXED>
__global__ void transform_kernel(const double2* __restrict__ src, int srcLayout, uint64_t nEvents,
                                 uint32_t nW0, uint32_t nW1, uint32_t wmax, uint32_t secOff1,
                                 uint32_t rowLen, const double* __restrict__ scale, void* dst)
{
	const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= nEvents * rowLen)
		return;
	const uint64_t e = idx / rowLen;
	const uint32_t pos = (uint32_t)(idx % rowLen);
	const int s = pos >= secOff1 ? 1 : 0;
	const uint32_t j = s ? pos - secOff1 : pos;
	const uint32_t nw = s ? nW1 : nW0;
	double2 v = make_double2(0., 0.);
	if (j < nw) {
		const uint32_t gw = s ? nW0 + j : j;
		if (srcLayout == RPWA_LAYOUT_EVT_WAVE)
			v = src[e * (nW0 + nW1) + gw];
		else
			v = src[((uint64_t)s * wmax + j) * nEvents + e];
		if (scale) {
			const double f = scale[gw];
			v.x *= f;
			v.y *= f;
		}
	}
	if (MIXED)
		reinterpret_cast<float2*>(dst)[idx] = make_float2((float)v.x, (float)v.y);
	else
		reinterpret_cast<double2*>(dst)[idx] = v;
}

using EventKernel = void (*)(EventArgs);

template <bool MIXED, int L, int KMAX>
EventKernel pick_rank(int rank)
{
	if (rank <= 1)
		return event_kernel<MIXED, L, KMAX, 1>;
	if (rank <= 2)
		return event_kernel<MIXED, L, KMAX, 2>;
	return event_kernel<MIXED, L, KMAX, 4>;
}

template <bool MIXED, int L>
EventKernel pick_k(int kmax, int rank)
{
	if (kmax <= 2)
		return pick_rank<MIXED, L, 2>(rank);
	if (kmax <= 4)
		return pick_rank<MIXED, L, 4>(rank);
	return pick_rank<MIXED, L, 8>(rank);
}

template <bool MIXED>
EventKernel pick_l(int lanes, int kmax, int rank)
{
	switch (lanes) {
	case 8: return pick_k<MIXED, 8>(kmax, rank);
	case 16: return pick_k<MIXED, 16>(kmax, rank);
	case 32: return pick_k<MIXED, 32>(kmax, rank);
	default: return pick_k<MIXED, 64>(kmax, rank);
	}
}

EventKernel pick_kernel(const AmpLayout& lay, const TileShape& ts, int rank)
{
	if (lay.precision == RPWA_PREC_MIXED)
		return pick_l<true>(ts.lanes, ts.kmax, rank);
	return pick_l<false>(ts.lanes, ts.kmax, rank);
}

int rank_cap(int rank) { return rank <= 1 ? 1 : (rank <= 2 ? 2 : 4); }

size_t lds_bytes(const AmpLayout& lay, const TileShape& ts, int rank)
{
	return (size_t)ts.lanes * rank_cap(rank) * ts.kmax * lay.perSlot * sizeof(double2);
}

}  // namespace

hipError_t launch_event_kernel(const AmpLayout& lay, const TileShape& ts, const EventArgs& a,
                               bool /*grad*/, int nWg, hipStream_t s)
{
	EventKernel k = pick_kernel(lay, ts, a.rank);
	hipLaunchKernelGGL(k, dim3(nWg), dim3(kWgThreads), lds_bytes(lay, ts, a.rank), s, a);
	return hipGetLastError();
}

int event_kernel_occupancy(const AmpLayout& lay, const TileShape& ts, bool /*grad*/, int* wgPerCu)
{
	// occupancy of the rank-2 variant (the common case) decides the grid
	EventKernel k = pick_kernel(lay, ts, 2);
	int n = 0;
	hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kWgThreads, lds_bytes(lay, ts, 2));
	if (e != hipSuccess)
		return RPWA_ERR_HIP;
	*wgPerCu = n > 0 ? n : 1;
	return RPWA_OK;
}

hipError_t launch_event_reduce(const double* partials, int nWg, uint32_t nElem, double flat,
                               double* out, hipStream_t s)
{
	const int nb = (int)((nElem + kRedElems - 1) / kRedElems);
	hipLaunchKernelGGL(event_reduce_kernel, dim3(nb), dim3(kRedElems * kRedChunks), 0, s, partials,
	                   nWg, nElem, flat, out);
	return hipGetLastError();
}

hipError_t launch_transform_amps(const AmpLayout& lay, const double* d_src, int srcLayout,
                                 const double* d_scale, void* d_dst, hipStream_t s)
{
	const uint64_t n = lay.nEvents * lay.rowLen;
	const int bs = 256;
	const uint64_t nb = (n + bs - 1) / bs;
	if (lay.precision == RPWA_PREC_MIXED)
		hipLaunchKernelGGL(transform_kernel<true>, dim3((uint32_t)nb), dim3(bs), 0, s,
		                   reinterpret_cast<const double2*>(d_src), srcLayout, lay.nEvents, lay.nW[0],
		                   lay.nW[1], lay.wmax, lay.secOff[1], lay.rowLen, d_scale, d_dst);
	else
		hipLaunchKernelGGL(transform_kernel<false>, dim3((uint32_t)nb), dim3(bs), 0, s,
		                   reinterpret_cast<const double2*>(d_src), srcLayout, lay.nEvents, lay.nW[0],
		                   lay.nW[1], lay.wmax, lay.secOff[1], lay.rowLen, d_scale, d_dst);
	return hipGetLastError();
}

}  // namespace rpwa_amd
