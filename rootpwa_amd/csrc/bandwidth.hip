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

__global__ __launch_bounds__(256) void read_kernel(const double4* __restrict__ src, uint64_t n,
                                                   double* __restrict__ sink)
{
	double acc = 0.;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double4 v = src[i];
		acc += v.x + v.y + v.z + v.w;
	}
	// the write keeps the loads alive; acc is never exactly this value for real data
	if (acc == 1.2345678901234567e300)
		sink[0] = acc;
}

__global__ __launch_bounds__(256) void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst,
                                                   uint64_t n)
{
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		dst[i] = src[i];
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
	const uint64_t n4 = bytes / sizeof(double4);
	double4 *a = nullptr, *b = nullptr;
	double* sink = nullptr;
	hipEvent_t e0, e1;
	RPWA_HIP_CHECK(hipMalloc(&a, n4 * sizeof(double4)));
	if (hipMalloc(&b, n4 * sizeof(double4)) != hipSuccess || hipMalloc(&sink, 64) != hipSuccess) {
		hipFree(a);
		hipFree(b);
		set_error("rpwa_bandwidth_probe: allocation failed");
		return RPWA_ERR_NOMEM;
	}
	int nCu = 0;
	hipDeviceGetAttribute(&nCu, hipDeviceAttributeMultiprocessorCount, device);
	const int grid = std::max(1, nCu) * 8;
	hipEventCreate(&e0);
	hipEventCreate(&e1);
	hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)((4 * n4 + 255) / 256)), dim3(256), 0, 0,
	                   reinterpret_cast<double*>(a), 4 * n4);
	std::vector<float> tr, tc;
	for (int r = 0; r < repetitions + 1; ++r) {
		hipEventRecord(e0, 0);
		hipLaunchKernelGGL(read_kernel, dim3(grid), dim3(256), 0, 0, a, n4, sink);
		hipEventRecord(e1, 0);
		hipEventSynchronize(e1);
		float ms = 0.f;
		hipEventElapsedTime(&ms, e0, e1);
		if (r > 0)
			tr.push_back(ms);
		hipEventRecord(e0, 0);
		hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a, b, n4);
		hipEventRecord(e1, 0);
		hipEventSynchronize(e1);
		hipEventElapsedTime(&ms, e0, e1);
		if (r > 0)
			tc.push_back(ms);
	}
	const hipError_t err = hipGetLastError();
	std::sort(tr.begin(), tr.end());
	std::sort(tc.begin(), tc.end());
	const double nbytes = (double)n4 * sizeof(double4);
	if (readGBs)
		*readGBs = nbytes / (1e-3 * tr[tr.size() / 2]) / 1e9;
	if (copyGBs)
		*copyGBs = 2 * nbytes / (1e-3 * tc[tc.size() / 2]) / 1e9;
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	hipFree(a);
	hipFree(b);
	hipFree(sink);
	RPWA_HIP_CHECK(err);
	return RPWA_OK;
}
