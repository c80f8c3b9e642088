// gen.hip -- synthetic decay amplitudes and accept-reject pseudo data.
//
// Amplitude recipe (the benchmark workloads):
//   A_i(e) = (N(0,1) + i N(0,1)) * s_i * exp(0.5 * N(0,1)_e),  s_i ~ U[0.1, 1].
// Every random number is a function of (seed, stream, counter) only
// (splitmix64 finaliser of a keyed counter), so any event of any pool can be
// generated independently and in any order.  Counters of event e of a pool
// with W waves: slot i < W -> complex normal of wave i, slot W -> the per-event
// lognormal factor, slot W+1 -> the accept-reject uniform; slot q uses the
// counters 2*(e*(W+2)+q) and 2*(e*(W+2)+q)+1.
//
// Accept-reject (pseudo data, cf. generators/modelIntensity.cc:170-187 for the
// intensity): event e of the pool is accepted iff u_e < sigma_e / max sigma,
// where sigma_e = sum_r sum_refl |sum_i T_r,refl,i A_refl,i(e)|^2 and the
// maximum is taken over the whole pool; accepted events keep pool order.
#include "internal.h"
#include "rng.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rpwa_amd {
namespace {

__global__ void gen_amps_kernel(uint64_t seed, uint32_t stream, uint64_t e0, uint64_t n, uint32_t W,
                                const double* __restrict__ scales, double2* __restrict__ out)
{
	const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n * W)
		return;
	const uint64_t e = e0 + idx / W;
	const uint32_t i = (uint32_t)(idx % W);
	const double lam = rng::event_factor(seed, stream, e, W);
	const double2 z = rng::complex_normal(seed, stream, e, W, i);
	const double f = scales[i] * lam;
	out[idx] = make_double2(z.x * f, z.y * f);
}

__global__ void gen_amps_indexed_kernel(uint64_t seed, uint32_t stream, const uint64_t* __restrict__ ev,
                                        uint64_t n, uint32_t W, const double* __restrict__ scales,
                                        double2* __restrict__ out)
{
	const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n * W)
		return;
	const uint64_t e = ev[idx / W];
	const uint32_t i = (uint32_t)(idx % W);
	const double lam = rng::event_factor(seed, stream, e, W);
	const double2 z = rng::complex_normal(seed, stream, e, W, i);
	const double f = scales[i] * lam;
	out[idx] = make_double2(z.x * f, z.y * f);
}

// intensity of every pool event; one thread per event, waves in order
__global__ void intensity_kernel(uint64_t seed, uint32_t stream, uint64_t e0, uint64_t n, uint32_t W0,
                                 uint32_t W1,
                                 uint32_t wmax, int rank, const double2* __restrict__ prod,
                                 const double* __restrict__ scales, double* __restrict__ sig,
                                 double* __restrict__ blockMax)
{
	__shared__ double sh[256];
	const uint64_t le = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t e = e0 + le;
	const uint32_t W = W0 + W1;
	double sigma = 0.;
	if (le < n) {
		const double lam = rng::event_factor(seed, stream, e, W);
		for (int r = 0; r < rank; ++r)
			for (int s = 0; s < 2; ++s) {
				const uint32_t nw = s ? W1 : W0;
				double2 S = make_double2(0., 0.);
				for (uint32_t j = 0; j < nw; ++j) {
					const uint32_t gw = s ? W0 + j : j;
					const double2 z = rng::complex_normal(seed, stream, e, W, gw);
					const double f = scales[gw] * lam;
					const double2 a = make_double2(z.x * f, z.y * f);
					const double2 t = prod[((size_t)r * 2 + s) * wmax + j];
					S.x += t.x * a.x - t.y * a.y;
					S.y += t.x * a.y + t.y * a.x;
				}
				sigma += S.x * S.x + S.y * S.y;
			}
		sig[le] = sigma;
	}
	sh[threadIdx.x] = sigma;
	__syncthreads();
	for (int m = 128; m > 0; m >>= 1) {
		if ((int)threadIdx.x < m)
			sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + m]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		blockMax[blockIdx.x] = sh[0];
}

__device__ __forceinline__ bool accepted(uint64_t seed, uint32_t stream, uint64_t e, uint32_t W,
                                         double sigma, double smax)
{
	return rng::accept_uniform(seed, stream, e, W) < sigma / smax;
}

__global__ void accept_count_kernel(uint64_t seed, uint32_t stream, uint64_t e0, uint64_t n, uint32_t W,
                                    const double* __restrict__ sig, double smax, uint32_t* __restrict__ cnt)
{
	__shared__ uint32_t c;
	if (threadIdx.x == 0)
		c = 0;
	__syncthreads();
	const uint64_t le = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (le < n && accepted(seed, stream, e0 + le, W, sig[le], smax))
		atomicAdd(&c, 1u);
	__syncthreads();
	if (threadIdx.x == 0)
		cnt[blockIdx.x] = c;
}

// ordered compaction: block b writes its accepted indices from offset[b] on
__global__ void accept_scatter_kernel(uint64_t seed, uint32_t stream, uint64_t e0, uint64_t n, uint32_t W,
                                      const double* __restrict__ sig, double smax,
                                      const uint64_t* __restrict__ offset, uint64_t nWanted,
                                      uint64_t* __restrict__ idxOut)
{
	__shared__ uint32_t flags[256];
	const uint64_t le = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const uint64_t e = e0 + le;
	const uint32_t f = (le < n && accepted(seed, stream, e, W, sig[le], smax)) ? 1u : 0u;
	flags[threadIdx.x] = f;
	__syncthreads();
	if (f) {
		uint32_t pos = 0;
		for (uint32_t t = 0; t < threadIdx.x; ++t)
			pos += flags[t];
		const uint64_t o = offset[blockIdx.x] + pos;
		if (o < nWanted)
			idxOut[o] = e;
	}
}

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

extern "C" int rpwa_gen_wave_scales(uint64_t seed, uint32_t nmbWaves, double* scales)
{
	clear_error();
	if (!scales) {
		set_error("rpwa_gen_wave_scales: null output");
		return RPWA_ERR_INVALID;
	}
	for (uint32_t i = 0; i < nmbWaves; ++i)
		scales[i] = rng::wave_scale(seed, i);
	return RPWA_OK;
}

extern "C" int rpwa_gen_prod_amps(uint64_t seed, uint32_t stream, uint32_t rank,
                                  const uint32_t nmbWavesRefl[2], double* prodAmps)
{
	clear_error();
	if (!prodAmps || !nmbWavesRefl || rank == 0) {
		set_error("rpwa_gen_prod_amps: invalid argument");
		return RPWA_ERR_INVALID;
	}
	const uint32_t wmax = std::max(nmbWavesRefl[0], nmbWavesRefl[1]);
	for (uint32_t r = 0; r < rank; ++r)
		for (uint32_t s = 0; s < 2; ++s)
			for (uint32_t j = 0; j < wmax; ++j) {
				const size_t k = ((size_t)r * 2 + s) * wmax + j;
				double re = 0., im = 0.;
				if (j < nmbWavesRefl[s] && j >= r) {
					rng::host_complex_normal(seed, stream, k, re, im);
					if (j == r)
						im = 0.;
				}
				prodAmps[2 * k] = re;
				prodAmps[2 * k + 1] = im;
			}
	return RPWA_OK;
}

extern "C" int rpwa_gen_amps_device(int device, uint64_t seed, uint32_t stream, uint64_t eventOffset,
                                    uint64_t nmbEvents, uint32_t nmbWaves, const double* scales,
                                    double* d_out)
{
	clear_error();
	if (!scales || !d_out || nmbWaves == 0) {
		set_error("rpwa_gen_amps_device: invalid argument");
		return RPWA_ERR_INVALID;
	}
	if (nmbEvents == 0)
		return RPWA_OK;
	RPWA_HIP_CHECK(hipSetDevice(device));
	double* d_scales = nullptr;
	RPWA_HIP_CHECK(hipMalloc(&d_scales, nmbWaves * sizeof(double)));
	hipError_t e = hipMemcpy(d_scales, scales, nmbWaves * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess) {
		const uint64_t n = nmbEvents * nmbWaves;
		const uint64_t nb = (n + 255) / 256;
		hipLaunchKernelGGL(gen_amps_kernel, dim3((uint32_t)nb), dim3(256), 0, 0, seed, stream,
		                   eventOffset, nmbEvents, nmbWaves, d_scales, reinterpret_cast<double2*>(d_out));
		e = hipGetLastError();
		if (e == hipSuccess)
			e = hipDeviceSynchronize();
	}
	hipFree(d_scales);
	RPWA_HIP_CHECK(e);
	return RPWA_OK;
}

extern "C" int rpwa_gen_accept_reject_device(int device, uint64_t seed, uint32_t stream,
                                             uint64_t nmbWanted, const uint32_t nmbWavesRefl[2],
                                             uint32_t rank, const double* prodAmps,
                                             const double* scales, double* d_out, uint64_t* poolSize,
                                             double* maxIntensity)
{
	clear_error();
	if (!nmbWavesRefl || !prodAmps || !scales || !d_out || rank == 0 || rank > 64 || nmbWanted == 0) {
		set_error("rpwa_gen_accept_reject_device: invalid argument");
		return RPWA_ERR_INVALID;
	}
	const uint32_t W0 = nmbWavesRefl[0], W1 = nmbWavesRefl[1], W = W0 + W1;
	const uint32_t wmax = std::max(W0, W1);
	if (W == 0) {
		set_error("rpwa_gen_accept_reject_device: no waves");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(device));
	// the pool is a list of segments; doubling the pool appends segments, so
	// intensities are computed once per event
	struct Segment {
		uint64_t e0 = 0, n = 0, nb = 0;
		double* sig = nullptr;
		uint32_t* cnt = nullptr;
		uint64_t* off = nullptr;
	};
	std::vector<Segment> segs;
	double *d_scales = nullptr, *d_prod = nullptr, *d_bmax = nullptr;
	uint64_t* d_idx = nullptr;
	auto cleanup = [&]() {
		hipFree(d_scales); hipFree(d_prod); hipFree(d_bmax); hipFree(d_idx);
		for (Segment& g : segs) {
			hipFree(g.sig); hipFree(g.cnt); hipFree(g.off);
		}
	};
#define AR_CHECK(expr)                                                                 \
	do {                                                                               \
		hipError_t e_ = (expr);                                                        \
		if (e_ != hipSuccess) {                                                        \
			set_error(std::string(#expr) + ": " + hipGetErrorString(e_));             \
			cleanup();                                                                 \
			return RPWA_ERR_HIP;                                                       \
		}                                                                              \
	} while (0)
	AR_CHECK(hipMalloc(&d_scales, W * sizeof(double)));
	AR_CHECK(hipMemcpy(d_scales, scales, W * sizeof(double), hipMemcpyHostToDevice));
	const size_t nProd = (size_t)rank * 2 * wmax;
	AR_CHECK(hipMalloc(&d_prod, nProd * 2 * sizeof(double)));
	AR_CHECK(hipMemcpy(d_prod, prodAmps, nProd * 2 * sizeof(double), hipMemcpyHostToDevice));
	AR_CHECK(hipMalloc(&d_idx, nmbWanted * sizeof(uint64_t)));
	const uint64_t kSeg = 1ull << 26;  // events per segment
	AR_CHECK(hipMalloc(&d_bmax, (kSeg / 256) * sizeof(double)));
	uint64_t pool = 0;
	uint64_t target = std::max<uint64_t>(nmbWanted * 64, 4096);
	double smax = 0.;
	for (int attempt = 0; attempt < 24; ++attempt) {
		while (pool < target) {
			Segment g;
			g.e0 = pool;
			g.n = std::min(kSeg, target - pool);
			g.nb = (g.n + 255) / 256;
			segs.push_back(g);
			Segment& sg = segs.back();
			AR_CHECK(hipMalloc(&sg.sig, sg.n * sizeof(double)));
			AR_CHECK(hipMalloc(&sg.cnt, sg.nb * sizeof(uint32_t)));
			AR_CHECK(hipMalloc(&sg.off, sg.nb * sizeof(uint64_t)));
			hipLaunchKernelGGL(intensity_kernel, dim3((uint32_t)sg.nb), dim3(256), 0, 0, seed, stream, sg.e0,
			                   sg.n, W0, W1, wmax, (int)rank, reinterpret_cast<const double2*>(d_prod), d_scales,
			                   sg.sig, d_bmax);
			AR_CHECK(hipGetLastError());
			std::vector<double> bmax(sg.nb);
			AR_CHECK(hipMemcpy(bmax.data(), d_bmax, sg.nb * sizeof(double), hipMemcpyDeviceToHost));
			for (double v : bmax)
				smax = std::max(smax, v);
			pool += sg.n;
		}
		if (!(smax > 0.)) {
			set_error("rpwa_gen_accept_reject_device: intensity is zero everywhere");
			cleanup();
			return RPWA_ERR_INVALID;
		}
		uint64_t tot = 0;
		std::vector<std::vector<uint64_t>> offs(segs.size());
		for (size_t k = 0; k < segs.size(); ++k) {
			Segment& sg = segs[k];
			hipLaunchKernelGGL(accept_count_kernel, dim3((uint32_t)sg.nb), dim3(256), 0, 0, seed, stream, sg.e0,
			                   sg.n, W, sg.sig, smax, sg.cnt);
			AR_CHECK(hipGetLastError());
			std::vector<uint32_t> cnt(sg.nb);
			AR_CHECK(hipMemcpy(cnt.data(), sg.cnt, sg.nb * sizeof(uint32_t), hipMemcpyDeviceToHost));
			offs[k].resize(sg.nb);
			for (uint64_t b = 0; b < sg.nb; ++b) {
				offs[k][b] = tot;
				tot += cnt[b];
			}
		}
		if (tot >= nmbWanted) {
			for (size_t k = 0; k < segs.size(); ++k) {
				Segment& sg = segs[k];
				if (offs[k][0] >= nmbWanted)
					break;
				AR_CHECK(hipMemcpy(sg.off, offs[k].data(), sg.nb * sizeof(uint64_t), hipMemcpyHostToDevice));
				hipLaunchKernelGGL(accept_scatter_kernel, dim3((uint32_t)sg.nb), dim3(256), 0, 0, seed, stream,
				                   sg.e0, sg.n, W, sg.sig, smax, sg.off, nmbWanted, d_idx);
				AR_CHECK(hipGetLastError());
			}
			const uint64_t n = nmbWanted * W;
			hipLaunchKernelGGL(gen_amps_indexed_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0,
			                   seed, stream, d_idx, nmbWanted, W, d_scales, reinterpret_cast<double2*>(d_out));
			AR_CHECK(hipGetLastError());
			AR_CHECK(hipDeviceSynchronize());
			if (poolSize)
				*poolSize = pool;
			if (maxIntensity)
				*maxIntensity = smax;
			cleanup();
			return RPWA_OK;
		}
		target = pool * 2;
	}
#undef AR_CHECK
	cleanup();
	set_error("rpwa_gen_accept_reject_device: acceptance too low");
	return RPWA_ERR_INVALID;
}
