// intensity.hip -- per-event model intensity for pseudo-data generation.
//
// modelIntensity::getIntensity (generators/modelIntensity.cc:170-187) weighs a
// generated event by
//   I(e) = sum_refl | sum_{i in refl} T_i a_i(e) |^2,   a_i = A_i / sqrt(Re I_ii)
// with the decay amplitudes of getDecayAmplitudes (:190-211: each divided by
// its phase-space integral) and the production amplitudes of a fit result; the
// flat wave is left out.  Here the sum also runs over the ranks of T
// ([rank][2][Wmax], the likelihood's production-amplitude layout), so that
//   I(e) = sum_r sum_refl | sum_i T_r,refl,i a_refl,i(e) |^2
// is the likelihood's sigma_e without the flat term; rank 1 is the reference's
// quantity.  The decay amplitudes arrive as [N][W0 + W1] complex128 rows (refl
// -1 waves first), each optionally divided by its wave's norm sqrt(Re I_ii).
//
// Work decomposition (one pass over the amplitudes, HBM-bound): a group of L
// lanes owns one event; lanes0 lanes hold the refl -1 waves, lanes1 the refl
// +1 waves, each lane K slots of one complex128 (a sector's lanes read one
// contiguous row segment per slot).  T sits in LDS; for every rank the lanes'
// partial sums of each sector are combined over the group by a butterfly and
// the group's first lane adds |S|^2 to the intensity.  Amplitudes of the next
// event group are loaded one iteration ahead.
#include "event_kernel.h"

#include <algorithm>
#include <string>
#include <vector>

namespace rpwa_amd {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

struct IntensityArgs {
	const double2* amps;   // [nEvents][W] complex128
	uint64_t nEvents;
	uint32_t W, nW0, nW1, wmax;
	uint32_t rank;
	int lanes0, lanes1;
	const double2* prod;   // [rank][2][wmax] complex
	const double* norm;    // [W] divisors or null
	double* out;           // [nEvents]
};

template <int L, int K>
__global__ __launch_bounds__(kThreads) void intensity_kernel(IntensityArgs a)
{
	constexpr int GPW = 64 / L;
	extern __shared__ double2 tsh[];  // [rank][2][wmax]
	const int tid = threadIdx.x;
	const uint32_t nT = a.rank * 2u * a.wmax;
	for (uint32_t i = tid; i < nT; i += kThreads)
		tsh[i] = a.prod[i];
	__syncthreads();

	const int lane = tid & 63;
	const int wid = tid >> 6;
	const int lig = lane & (L - 1);
	const int grp = lane / L;
	const int sec = lig >= a.lanes0 ? 1 : 0;
	const int ls = sec ? lig - a.lanes0 : lig;
	const int laneStride = sec ? a.lanes1 : a.lanes0;
	const int nw = (ls < laneStride) ? (int)(sec ? a.nW1 : a.nW0) : 0;

	bool valid[K];
	uint32_t col[K], tIdx[K];
	double nrm[K];
#pragma unroll
	for (int k = 0; k < K; ++k) {
		const int j = ls + laneStride * k;
		valid[k] = j < nw;
		col[k] = valid[k] ? (sec ? a.nW0 : 0u) + (uint32_t)j : 0u;
		tIdx[k] = valid[k] ? (uint32_t)sec * a.wmax + (uint32_t)j : 0u;
		nrm[k] = (valid[k] && a.norm) ? a.norm[col[k]] : 1.;
	}

	const uint64_t stride = (uint64_t)gridDim.x * kWaves * GPW;
	const uint64_t eFirst = ((uint64_t)blockIdx.x * kWaves + wid) * GPW + grp;
	double2 An[K];
	auto load = [&](uint64_t e) {
#pragma unroll
		for (int k = 0; k < K; ++k)
			An[k] = (e < a.nEvents && valid[k]) ? a.amps[e * a.W + col[k]] : make_double2(0., 0.);
	};
	load(eFirst);
	for (uint64_t e = eFirst; e - grp < a.nEvents; e += stride) {
		double2 A[K];
#pragma unroll
		for (int k = 0; k < K; ++k)
			A[k] = An[k];
		if (e - grp + stride < a.nEvents)
			load(e + stride);
		if (a.norm) {
			// the reference's a_i = A_i / sqrt(I_ii), component-wise
#pragma unroll
			for (int k = 0; k < K; ++k) {
				A[k].x = A[k].x / nrm[k];
				A[k].y = A[k].y / nrm[k];
			}
		}
		double inten = 0.;
		for (uint32_t r = 0; r < a.rank; ++r) {
			double2 Sa = make_double2(0., 0.);
			const double2* tr = tsh + (size_t)r * 2 * a.wmax;
#pragma unroll
			for (int k = 0; k < K; ++k) {
				const double2 t = valid[k] ? tr[tIdx[k]] : make_double2(0., 0.);
				ek::cmac(Sa, t, A[k]);
			}
			const double s0x = ek::group_sum<L>(sec ? 0. : Sa.x);
			const double s0y = ek::group_sum<L>(sec ? 0. : Sa.y);
			const double s1x = ek::group_sum<L>(sec ? Sa.x : 0.);
			const double s1y = ek::group_sum<L>(sec ? Sa.y : 0.);
			inten += s0x * s0x + s0y * s0y;
			inten += s1x * s1x + s1y * s1y;
		}
		if (lig == 0 && e < a.nEvents)
			a.out[e] = inten;
	}
}

using IntensityKernel = void (*)(IntensityArgs);

template <int L>
IntensityKernel pick_k(int k)
{
	switch (k) {
	case 1: return intensity_kernel<L, 1>;
	case 2: return intensity_kernel<L, 2>;
	case 3: return intensity_kernel<L, 3>;
	case 4: return intensity_kernel<L, 4>;
	case 5: return intensity_kernel<L, 5>;
	case 6: return intensity_kernel<L, 6>;
	case 7: return intensity_kernel<L, 7>;
	default: return intensity_kernel<L, 8>;
	}
}

IntensityKernel pick(int lanes, int k)
{
	switch (lanes) {
	case 8: return pick_k<8>(k);
	case 16: return pick_k<16>(k);
	case 32: return pick_k<32>(k);
	default: return pick_k<64>(k);
	}
}

// the fewest lanes whose group holds the row with at most 4 slots per lane
// (8 at 64 lanes), each sector on its own lanes
bool intensity_tile(uint32_t nW0, uint32_t nW1, int& lanes, int& k, int& l0, int& l1)
{
	for (int L : {8, 16, 32, 64})
		for (int kk = 1; kk <= (L == 64 ? 8 : 4); ++kk) {
			const int a0 = (int)((nW0 + kk - 1) / kk), a1 = (int)((nW1 + kk - 1) / kk);
			if (a0 + a1 <= L) {
				lanes = L;
				k = kk;
				l0 = a0;
				l1 = a1;
				return true;
			}
		}
	return false;
}

int intensity_run(int device, const double* d_amps, uint64_t N, const uint32_t nW[2], const double* prodAmps,
                  uint32_t rank, const double* waveNorm, double* d_out)
{
	const uint32_t W = nW[0] + nW[1];
	const uint32_t wmax = std::max(nW[0], nW[1]);
	int lanes = 0, k = 0, l0 = 0, l1 = 0;
	if (!intensity_tile(nW[0], nW[1], lanes, k, l0, l1)) {
		set_error("rpwa_intensity: " + std::to_string(nW[0]) + "+" + std::to_string(nW[1]) +
		          " waves exceed 64 lanes x 8 slots");
		return RPWA_ERR_UNSUPPORTED;
	}
	const size_t nProd = (size_t)rank * 2 * wmax;
	const size_t lds = nProd * sizeof(double2);
	if (lds > 64u * 1024u) {
		set_error("rpwa_intensity: rank x 2 x max(nmbWavesRefl) = " + std::to_string(nProd) +
		          " production amplitudes exceed 4096");
		return RPWA_ERR_UNSUPPORTED;
	}
	RPWA_HIP_CHECK(hipSetDevice(device));
	int nCu = 0;
	RPWA_HIP_CHECK(hipDeviceGetAttribute(&nCu, hipDeviceAttributeMultiprocessorCount, device));
	double2* d_prod = nullptr;
	double* d_norm = nullptr;
	auto cleanup = [&]() {
		(void)hipFree(d_prod);
		(void)hipFree(d_norm);
	};
	hipError_t e = hipMalloc(&d_prod, nProd * sizeof(double2));
	if (e == hipSuccess)
		e = hipMemcpy(d_prod, prodAmps, nProd * sizeof(double2), hipMemcpyHostToDevice);
	if (e == hipSuccess && waveNorm) {
		e = hipMalloc(&d_norm, W * sizeof(double));
		if (e == hipSuccess)
			e = hipMemcpy(d_norm, waveNorm, W * sizeof(double), hipMemcpyHostToDevice);
	}
	if (e != hipSuccess) {
		cleanup();
		RPWA_HIP_CHECK(e);
	}
	IntensityArgs a;
	a.amps = reinterpret_cast<const double2*>(d_amps);
	a.nEvents = N;
	a.W = W;
	a.nW0 = nW[0];
	a.nW1 = nW[1];
	a.wmax = wmax;
	a.rank = rank;
	a.lanes0 = l0;
	a.lanes1 = l1;
	a.prod = d_prod;
	a.norm = d_norm;
	a.out = d_out;
	const uint64_t perWg = (uint64_t)kWaves * (64 / lanes);
	const uint64_t need = (N + perWg - 1) / perWg;
	const uint32_t nWg = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)nCu * 8));
	hipLaunchKernelGGL(pick(lanes, k), dim3(nWg), dim3(kThreads), lds, 0, a);
	e = hipGetLastError();
	if (e == hipSuccess)
		e = hipDeviceSynchronize();
	cleanup();
	RPWA_HIP_CHECK(e);
	return RPWA_OK;
}

int intensity_check(const char* fn, const void* amps, uint64_t N, const uint32_t* nW, const double* prodAmps,
                    uint32_t nmbProdAmps, uint32_t rank, const double* out)
{
	if (!amps || !nW || !prodAmps || !out || N == 0 || nW[0] + nW[1] == 0 || rank == 0) {
		set_error(std::string(fn) + ": null pointer, no events, no waves or rank 0");
		return RPWA_ERR_INVALID;
	}
	if ((uint64_t)nmbProdAmps != (uint64_t)rank * 2 * std::max(nW[0], nW[1])) {
		set_error(std::string(fn) + ": nmbProdAmps must be rank*2*nmbWavesMax");
		return RPWA_ERR_INVALID;
	}
	return RPWA_OK;
}

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

extern "C" int rpwa_intensity_device(int device, const double* d_amps, uint64_t nmbEvents,
                                     const uint32_t nmbWavesRefl[2], const double* prodAmps, uint32_t nmbProdAmps,
                                     uint32_t rank, const double* waveNorm, double* d_intensity)
{
	clear_error();
	if (int st = intensity_check("rpwa_intensity_device", d_amps, nmbEvents, nmbWavesRefl, prodAmps, nmbProdAmps,
	                             rank, d_intensity))
		return st;
	try {
		return intensity_run(device, d_amps, nmbEvents, nmbWavesRefl, prodAmps, rank, waveNorm, d_intensity);
	} catch (const std::exception& ex) {
		set_error(ex.what());
		return RPWA_ERR_NOMEM;
	}
}

extern "C" int rpwa_intensity(int device, const double* amps, uint64_t nmbEvents, const uint32_t nmbWavesRefl[2],
                              const double* prodAmps, uint32_t nmbProdAmps, uint32_t rank, const double* waveNorm,
                              double* intensity)
{
	clear_error();
	if (int st = intensity_check("rpwa_intensity", amps, nmbEvents, nmbWavesRefl, prodAmps, nmbProdAmps, rank,
	                             intensity))
		return st;
	try {
		RPWA_HIP_CHECK(hipSetDevice(device));
		const size_t W = (size_t)nmbWavesRefl[0] + nmbWavesRefl[1];
		double* d_amps = nullptr;
		double* d_out = nullptr;
		hipError_t e = hipMalloc(&d_amps, nmbEvents * W * 16);
		if (e == hipSuccess)
			e = hipMalloc(&d_out, nmbEvents * sizeof(double));
		if (e == hipSuccess)
			e = hipMemcpy(d_amps, amps, nmbEvents * W * 16, hipMemcpyHostToDevice);
		int st = RPWA_OK;
		if (e == hipSuccess) {
			st = intensity_run(device, d_amps, nmbEvents, nmbWavesRefl, prodAmps, rank, waveNorm, d_out);
			if (st == RPWA_OK)
				e = hipMemcpy(intensity, d_out, nmbEvents * sizeof(double), hipMemcpyDeviceToHost);
		}
		(void)hipFree(d_amps);
		(void)hipFree(d_out);
		if (st != RPWA_OK)
			return st;
		RPWA_HIP_CHECK(e);
		return RPWA_OK;
	} catch (const std::exception& ex) {
		set_error(ex.what());
		return RPWA_ERR_NOMEM;
	}
}
