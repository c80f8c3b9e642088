// bandwidth.hip -- HBM bandwidth probe (read-only and copy kernels).
//
// Plays the role of the reference's memory-bandwidth microbenchmarks
// (cuda/test/benchmarkMemoryBandwidth.cu:85-260): it measures what a streaming
// kernel reaches on this device, the practical ceiling of the HBM-bound event
// kernels, and gives a copy kernel of known size to calibrate the byte counters
// of a profiler against.
#include "internal.h"

#include <algorithm>
#include <vector>

namespace rpwa_amd {
namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // one 16-byte load

// Streaming read: every thread keeps U independent 16-byte loads in flight per
// step; the loads of a step cover U x blockDim consecutive units, so each load
// instruction of a wave is one contiguous 1 KiB; NT marks them non-temporal (no
// cache allocation).  The integer fold is cheap and keeps every load alive.
template <int U, bool NT>
__global__ __launch_bounds__(256) void read_kernel(const u32x4* __restrict__ src, uint64_t n, unsigned int* sink)
{
	unsigned int acc = 0;
	const uint64_t span = (uint64_t)U * blockDim.x;
	for (uint64_t base = (uint64_t)blockIdx.x * span; base < n; base += (uint64_t)gridDim.x * span) {
		u32x4 v[U];
#pragma unroll
		for (int k = 0; k < U; ++k) {
			const uint64_t i = base + (uint64_t)k * blockDim.x + threadIdx.x;
			if (i < n)
				v[k] = NT ? __builtin_nontemporal_load(src + i) : src[i];
			else
				v[k] = u32x4{0u, 0u, 0u, 0u};
		}
#pragma unroll
		for (int k = 0; k < U; ++k)
			acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
	}
	// never true for the probe's fill pattern in practice; the store keeps the loads
	if (acc == 0x9E3779B9u)
		sink[0] = acc;
}

// Copy with U loads in flight per thread, then U stores.
template <int U>
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst, uint64_t n)
{
	const uint64_t span = (uint64_t)U * blockDim.x;
	for (uint64_t base = (uint64_t)blockIdx.x * span; base < n; base += (uint64_t)gridDim.x * span) {
		u32x4 v[U];
#pragma unroll
		for (int k = 0; k < U; ++k) {
			const uint64_t i = base + (uint64_t)k * blockDim.x + threadIdx.x;
			if (i < n)
				v[k] = src[i];
		}
#pragma unroll
		for (int k = 0; k < U; ++k) {
			const uint64_t i = base + (uint64_t)k * blockDim.x + threadIdx.x;
			if (i < n)
				dst[i] = v[k];
		}
	}
}

__global__ void fill_kernel(double* p, uint64_t n)
{
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 1e-3 * (double)(i % 1000);
}

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

extern "C" int rpwa_bandwidth_probe(int device, uint64_t bytes, int repetitions, double* readGBs, double* copyGBs)
{
	clear_error();
	if (bytes < (1u << 20) || repetitions < 1) {
		set_error("rpwa_bandwidth_probe: at least 1 MiB and one repetition");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(device));
	const uint64_t n = bytes / 16;
	u32x4 *a = nullptr, *b = nullptr;
	unsigned int* sink = nullptr;
	RPWA_HIP_CHECK(hipMalloc(&a, n * 16));
	if (hipMalloc(&b, n * 16) != hipSuccess || hipMalloc(&sink, 64) != hipSuccess) {
		hipFree(a);
		hipFree(b);
		set_error("rpwa_bandwidth_probe: allocation failed");
		return RPWA_ERR_NOMEM;
	}
	int nCu = 0;
	hipDeviceGetAttribute(&nCu, hipDeviceAttributeMultiprocessorCount, device);
	nCu = std::max(1, nCu);
	hipEvent_t e0, e1;
	hipEventCreate(&e0);
	hipEventCreate(&e1);
	hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)((2 * n + 255) / 256)), dim3(256), 0, 0,
	                   reinterpret_cast<double*>(a), 2 * n);
	// median time of one launch configuration (after one untimed launch)
	auto median_ms = [&](auto launch) {
		std::vector<float> t;
		for (int r = 0; r < repetitions + 1; ++r) {
			hipEventRecord(e0, 0);
			launch();
			hipEventRecord(e1, 0);
			hipEventSynchronize(e1);
			float ms = 0.f;
			hipEventElapsedTime(&ms, e0, e1);
			if (r > 0)
				t.push_back(ms);
		}
		std::sort(t.begin(), t.end());
		return (double)t[t.size() / 2];
	};
	// the best of several read variants (loads in flight per thread, temporal or
	// not, workgroups per CU) and copy variants: the ceiling a streaming kernel reaches
	double bestRead = 1e300, bestCopy = 1e300;
	for (int wgPerCu : {4, 8, 16}) {
		const dim3 grid((uint32_t)(nCu * wgPerCu));
		bestRead = std::min(bestRead, median_ms([&] { hipLaunchKernelGGL((read_kernel<4, false>), grid, dim3(256), 0, 0, a, n, sink); }));
		bestRead = std::min(bestRead, median_ms([&] { hipLaunchKernelGGL((read_kernel<8, false>), grid, dim3(256), 0, 0, a, n, sink); }));
		bestRead = std::min(bestRead, median_ms([&] { hipLaunchKernelGGL((read_kernel<16, false>), grid, dim3(256), 0, 0, a, n, sink); }));
		bestRead = std::min(bestRead, median_ms([&] { hipLaunchKernelGGL((read_kernel<8, true>), grid, dim3(256), 0, 0, a, n, sink); }));
		bestRead = std::min(bestRead, median_ms([&] { hipLaunchKernelGGL((read_kernel<16, true>), grid, dim3(256), 0, 0, a, n, sink); }));
		bestCopy = std::min(bestCopy, median_ms([&] { hipLaunchKernelGGL((copy_kernel<4>), grid, dim3(256), 0, 0, a, b, n); }));
		bestCopy = std::min(bestCopy, median_ms([&] { hipLaunchKernelGGL((copy_kernel<8>), grid, dim3(256), 0, 0, a, b, n); }));
	}
	const hipError_t err = hipGetLastError();
	const double nbytes = (double)n * 16;
	if (readGBs)
		*readGBs = nbytes / (1e-3 * bestRead) / 1e9;
	if (copyGBs)
		*copyGBs = 2 * nbytes / (1e-3 * bestCopy) / 1e9;
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	hipFree(a);
	hipFree(b);
	hipFree(sink);
	RPWA_HIP_CHECK(err);
	return RPWA_OK;
}
