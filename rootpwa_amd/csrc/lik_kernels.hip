// lik_kernels.hip -- gfx950 event kernels of the unbinned extended likelihood.
//
// One pass over the decay amplitudes computes, for every event e,
//   S_r,s(e) = sum_i T_r,s,i A_s,i(e)            (sector s = reflectivity)
//   sigma_e  = sum_r sum_s |S_r,s(e)|^2 + flat^2
// and accumulates  sum_e -ln sigma_e  and the event part of the
// gradient  G_r,s,i = sum_e (-2/sigma_e) S_r,s(e) conj(A_s,i(e)),  i.e. the
// quantities of pwaLikelihood::FdF's event loop (partialWaveFit/pwaLikelihood.cc:213-244)
// and of cuda/likelihoodKernel.cu:84-259, fused into one kernel.
//
// Work decomposition (single fit start: the problem is HBM-bound, 2-4 flop/B):
//   * a group of L lanes of a 64-wide wavefront owns one event; each lane holds
//     K slots of the event's amplitude row of one sector in registers, each slot
//     P complex values (P = 1 for complex128, 2 for complex64, so every lane load
//     is 16 bytes and a sector's lanes read a contiguous row segment per slot);
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
#include "event_kernel.h"

#include <algorithm>

namespace rpwa_amd {
namespace {

constexpr int kWgThreads = 256;

__device__ __forceinline__ void neumaier_add(double& s, double& c, double x)
{
	const double t = s + x;
	if (fabs(s) >= fabs(x))
		c += (s - t) + x;
	else
		c += (x - t) + s;
	s = t;
}

// Second stage: fixed-order sum of the workgroup partial rows.
// Block = 16 elements x 64 chunks of workgroups: each thread adds the rows of
// its chunk (loads issued kRedBatch at a time), then the 64 chunk sums of an
// element are combined by a fixed pairwise tree in LDS.  Element 0 (the lnL
// sum, with element 1 its compensation) is summed with Neumaier steps in the
// chunks and combined with error-free additions (TwoSum) in the tree.
constexpr int kRedElems = 16;
constexpr int kRedChunks = 64;
constexpr int kRedBatch = 8;

// Completion: after its results (each written with a system-scope fence) every
// block b stores seq into doneFlag[b] (host-visible memory), so a host can wait
// for the results by watching the flags of all blocks instead of a blocking
// stream synchronisation; no block waits for another.
__global__ __launch_bounds__(kRedElems* kRedChunks) void event_reduce_kernel(
    const double* __restrict__ partials, int nWg, uint32_t nElem, double flat,
    double* __restrict__ out, unsigned long long* doneFlag, unsigned long long seq)
{
	__shared__ double sh[kRedChunks][kRedElems];
	__shared__ double shc[kRedChunks][kRedElems];
	const int el = threadIdx.x % kRedElems;
	const int ch = threadIdx.x / kRedElems;
	const uint32_t elem = blockIdx.x * kRedElems + el;
	const int w0 = (int)((int64_t)nWg * ch / kRedChunks);
	const int w1 = (int)((int64_t)nWg * (ch + 1) / kRedChunks);
	double s = 0., c = 0.;
	if (elem == 0) {
		// lnL: compensated over both the sums and their corrections (the loads of a
		// batch are issued before the dependent additions)
		const double* q = partials + (size_t)w0 * nElem;
		int w = w0;
		for (; w + kRedBatch <= w1; w += kRedBatch, q += (size_t)kRedBatch * nElem) {
			double2 v[kRedBatch];
#pragma unroll
			for (int k = 0; k < kRedBatch; ++k)
				v[k] = make_double2(q[(size_t)k * nElem], q[(size_t)k * nElem + 1]);
#pragma unroll
			for (int k = 0; k < kRedBatch; ++k) {
				neumaier_add(s, c, v[k].x);
				neumaier_add(s, c, v[k].y);
			}
		}
		for (; w < w1; ++w, q += nElem) {
			neumaier_add(s, c, q[0]);
			neumaier_add(s, c, q[1]);
		}
	} else if (elem > 1 && elem < nElem) {
		const double* q = partials + (size_t)w0 * nElem + elem;
		int w = w0;
		for (; w + kRedBatch <= w1; w += kRedBatch, q += (size_t)kRedBatch * nElem) {
			double v[kRedBatch];
#pragma unroll
			for (int k = 0; k < kRedBatch; ++k)
				v[k] = q[(size_t)k * nElem];
#pragma unroll
			for (int k = 0; k < kRedBatch; ++k)
				s += v[k];
		}
		for (; w < w1; ++w, q += nElem)
			s += q[0];
	}
	sh[ch][el] = s;
	shc[ch][el] = c;
	__syncthreads();
	// chunk k absorbs chunk k + step: ((0+32)+(16+48))+... a fixed tree
	for (int step = kRedChunks / 2; step > 0; step >>= 1) {
		if (ch < step) {
			const double a = sh[ch][el], b = sh[ch + step][el];
			const double t = a + b;
			if (elem == 0) {
				const double bp = t - a;
				const double err = (a - (t - bp)) + (b - bp);  // TwoSum: a + b = t + err exactly
				shc[ch][el] = (shc[ch][el] + shc[ch + step][el]) + err;
			}
			sh[ch][el] = t;
		}
		__syncthreads();
	}
	if (ch == 0 && elem < nElem && elem != 1) {
		const double t = sh[0][el];
		// out[0] = lnL, out[1] = flat * sum_e(-2/sigma_e), out[2..] = derivatives
		if (elem == 0)
			out[0] = t + shc[0][el];
		else if (elem == 2)
			out[1] = flat * t;
		else
			out[elem - 1] = t;
		__threadfence_system();  // this thread's results are visible to the host
	}
	if (doneFlag) {
		__syncthreads();
		if (threadIdx.x == 0)
			__hip_atomic_store(doneFlag + blockIdx.x, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
	}
}

// Amplitude layout transform: source FP64 (either host layout) -> device rows.
template <bool MIXED>
__global__ void transform_kernel(const double2* __restrict__ src, int srcLayout, uint64_t nEvents,
                                 uint32_t nW0, uint32_t nW1, uint32_t wmax, uint32_t secOff1,
                                 uint32_t rowLen, const double* __restrict__ norm, void* dst)
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
		if (norm) {
			// complex / real component-wise, correctly rounded: the host's a / sqrt(I_ii)
			const double f = norm[gw];
			v.x = v.x / f;
			v.y = v.y / f;
		}
	}
	if (MIXED)
		reinterpret_cast<float2*>(dst)[idx] = make_float2((float)v.x, (float)v.y);
	else
		reinterpret_cast<double2*>(dst)[idx] = v;
}

// One wave's column of [nEvents][W] complex128 amplitudes: is any element
// non-zero (NaN counts as non-zero, as `a != 0` does)?  One integer OR per
// workgroup.
__global__ __launch_bounds__(256) void wave_nonzero_kernel(const double2* __restrict__ amps, uint64_t nEvents,
                                                           uint32_t W, uint32_t wave, unsigned int* flag)
{
	int nz = 0;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nEvents; e += stride) {
		const double2 v = amps[e * W + wave];
		nz |= (v.x != 0. || v.y != 0.) ? 1 : 0;
	}
	if (__syncthreads_or(nz) && threadIdx.x == 0)
		atomicOr(flag, 1u);
}

EventKernel pick_kernel(const AmpLayout& lay, const TileShape& ts, int rank)
{
	const bool mixed = lay.precision == RPWA_PREC_MIXED;
	switch (ts.lanes) {
	case 8: return mixed ? event_family_c64_l8(ts.slots, rank) : event_family_f64_l8(ts.slots, rank);
	case 16: return mixed ? event_family_c64_l16(ts.slots, rank) : event_family_f64_l16(ts.slots, rank);
	case 32: return mixed ? event_family_c64_l32(ts.slots, rank) : event_family_f64_l32(ts.slots, rank);
	default: return mixed ? event_family_c64_l64(ts.slots, rank) : event_family_f64_l64(ts.slots, rank);
	}
}

int rank_cap(int rank) { return rank <= 1 ? 1 : (rank <= 2 ? 2 : 4); }

size_t lds_bytes(const AmpLayout& lay, const TileShape& ts, int rank)
{
	return (size_t)ts.lanes * rank_cap(rank) * ts.slots * lay.perSlot * sizeof(double2);
}

}  // namespace

hipError_t launch_event_kernel(const AmpLayout& lay, const TileShape& ts, const EventArgs& a,
                               bool /*grad*/, int nWg, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
	if (ts.events > 0)
		return launch_stage_kernel(lay, ts.events, ts.gm, a, nWg, s, start, stop);
	EventKernel k = pick_kernel(lay, ts, a.rank);
	if (start || stop)
		hipExtLaunchKernelGGL(k, dim3(nWg), dim3(kWgThreads), (uint32_t)lds_bytes(lay, ts, a.rank), s, start, stop, 0u, a);
	else
		hipLaunchKernelGGL(k, dim3(nWg), dim3(kWgThreads), lds_bytes(lay, ts, a.rank), s, a);
	return hipGetLastError();
}

int event_kernel_occupancy(const AmpLayout& lay, const TileShape& ts, int rank, int* wgPerCu)
{
	if (ts.events > 0)
		return stage_kernel_occupancy(lay, ts.events, ts.gm, rank, wgPerCu);
	EventKernel k = pick_kernel(lay, ts, rank);
	int n = 0;
	hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kWgThreads, lds_bytes(lay, ts, rank));
	if (e != hipSuccess)
		return RPWA_ERR_HIP;
	*wgPerCu = n > 0 ? n : 1;
	return RPWA_OK;
}

hipError_t launch_event_reduce(const double* partials, int nWg, uint32_t nElem, double flat,
                               double* out, hipStream_t s, unsigned long long* doneFlag,
                               unsigned long long seq, hipEvent_t start, hipEvent_t stop)
{
	const int nb = (int)reduce_blocks(nElem);
	if (start || stop)
		hipExtLaunchKernelGGL(event_reduce_kernel, dim3(nb), dim3(kRedElems * kRedChunks), 0u, s, start, stop, 0u,
		                      partials, nWg, nElem, flat, out, doneFlag, seq);
	else
		hipLaunchKernelGGL(event_reduce_kernel, dim3(nb), dim3(kRedElems * kRedChunks), 0, s, partials,
		                   nWg, nElem, flat, out, doneFlag, seq);
	return hipGetLastError();
}

uint32_t reduce_blocks(uint32_t nElem) { return (nElem + kRedElems - 1) / kRedElems; }

hipError_t launch_wave_nonzero(const double* d_amps, uint64_t nEvents, uint32_t W, uint32_t wave,
                               unsigned int* flag, int nCu, hipStream_t s)
{
	const uint64_t need = (nEvents + 255) / 256;
	const uint32_t nb = (uint32_t)std::min<uint64_t>(need, (uint64_t)std::max(nCu, 1) * 8);
	hipLaunchKernelGGL(wave_nonzero_kernel, dim3(std::max(nb, 1u)), dim3(256), 0, s,
	                   reinterpret_cast<const double2*>(d_amps), nEvents, W, wave, flag);
	return hipGetLastError();
}

hipError_t launch_transform_amps(const AmpLayout& lay, const double* d_src, int srcLayout,
                                 const double* d_norm, void* d_dst, hipStream_t s)
{
	const uint64_t n = lay.nEvents * lay.rowLen;
	const int bs = 256;
	const uint64_t nb = (n + bs - 1) / bs;
	if (lay.precision == RPWA_PREC_MIXED)
		hipLaunchKernelGGL(transform_kernel<true>, dim3((uint32_t)nb), dim3(bs), 0, s,
		                   reinterpret_cast<const double2*>(d_src), srcLayout, lay.nEvents, lay.nW[0],
		                   lay.nW[1], lay.wmax, lay.secOff[1], lay.rowLen, d_norm, d_dst);
	else
		hipLaunchKernelGGL(transform_kernel<false>, dim3((uint32_t)nb), dim3(bs), 0, s,
		                   reinterpret_cast<const double2*>(d_src), srcLayout, lay.nEvents, lay.nW[0],
		                   lay.nW[1], lay.wmax, lay.secOff[1], lay.rowLen, d_norm, d_dst);
	return hipGetLastError();
}

}  // namespace rpwa_amd
