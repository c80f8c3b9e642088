// hessian.hip -- event part of the analytic Hessian on the GPU.
//
// pwaLikelihood::Hessian (partialWaveFit/pwaLikelihood.cc:697-753) sums per event
//   factor2 * Re/Im(d_b) * Re/Im(d_a)            with d_a = S_{r_a s_a}(e) conj(A_a(e)),
//   - factor * Re/Im(A_a conj(A_b))               for a, b of the same (rank, sector),
//   factor2 * flat * d_a  and  factor2 * flat^2 - factor,
// factor = 2 / sigma_e.  Each is a weighted rank-k update over events: a row
// kernel writes, per event, the vector (d_a..., 1) with weight factor^2 and
// the amplitude row (A..., 1) with weight -factor; the FP64-MFMA rank-k update
// of integral.hip sums x x^T for both; the constant column yields the flat-wave
// terms.  Events are processed in chunks so the row buffers stay bounded.
#include "internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace rpwa_amd {
namespace {

template <bool MIXED>
__device__ __forceinline__ double2 amp_at(const void* amps, uint64_t e, uint32_t rowLen, uint32_t pos)
{
	if (MIXED) {
		const float2 v = reinterpret_cast<const float2*>(amps)[e * rowLen + pos];
		return make_double2((double)v.x, (double)v.y);
	}
	return reinterpret_cast<const double2*>(amps)[e * rowLen + pos];
}

// One block handles nEv consecutive events: their amplitude rows are staged
// in LDS with coalesced loads (compact wave order: the refl -1 waves, then the
// refl +1 waves), then S, the weights and the rows are computed with the
// threads spread over (event, entry), so that every global store is coalesced.
template <bool MIXED>
__global__ __launch_bounds__(256) void hessian_rows_kernel(const void* amps, uint64_t e0, uint64_t n,
                                                           uint32_t nEv, uint32_t rowLen, uint32_t secOff1,
                                                           uint32_t nW0, uint32_t nW1, uint32_t wmax, int rank,
                                                           const double2* __restrict__ prod, double flat2,
                                                           uint32_t nd, double2* __restrict__ dRows,
                                                           double2* __restrict__ aRows, double* __restrict__ wD,
                                                           double* __restrict__ wA)
{
	extern __shared__ double2 lds[];
	const uint32_t W = nW0 + nW1;
	const uint32_t R2 = 2 * (uint32_t)rank;
	double2* A = lds;                           // [nEv][W]
	double2* Sx = lds + (size_t)nEv * W;        // [nEv][rank][2]
	const uint64_t k0 = (uint64_t)blockIdx.x * nEv;
	const uint32_t ne = (uint32_t)min((uint64_t)nEv, n - k0);
	for (uint32_t idx = threadIdx.x; idx < ne * W; idx += blockDim.x) {
		const uint32_t ev = idx / W, g = idx - ev * W;
		A[idx] = amp_at<MIXED>(amps, e0 + k0 + ev, rowLen, g < nW0 ? g : secOff1 + (g - nW0));
	}
	__syncthreads();
	// S_rs(e) = sum_i T_rsi A_si(e), in the order of the FdF kernels' host restatement
	for (uint32_t idx = threadIdx.x; idx < ne * R2; idx += blockDim.x) {
		const uint32_t ev = idx / R2, rs = idx - ev * R2, r = rs / 2, s = rs % 2;
		const uint32_t nw = s ? nW1 : nW0, off = s ? nW0 : 0;
		double2 acc = make_double2(0., 0.);
		for (uint32_t i = 0; i < nw; ++i) {
			const double2 t = prod[((size_t)r * 2 + s) * wmax + i];
			const double2 x = A[(size_t)ev * W + off + i];
			acc.x = fma(t.x, x.x, acc.x);
			acc.x = fma(-t.y, x.y, acc.x);
			acc.y = fma(t.x, x.y, acc.y);
			acc.y = fma(t.y, x.x, acc.y);
		}
		Sx[idx] = acc;
	}
	__syncthreads();
	if (threadIdx.x < ne) {
		const uint32_t ev = threadIdx.x;
		double sigma = 0.;
		for (uint32_t rs = 0; rs < R2; ++rs) {
			const double2 v = Sx[(size_t)ev * R2 + rs];
			sigma += v.x * v.x + v.y * v.y;
		}
		sigma += flat2;
		const double factor = 2. / sigma;
		wD[k0 + ev] = factor * factor;
		wA[k0 + ev] = -factor;
	}
	// d rows (entry q = r W + compact wave g; d = S conj(A)) and the constant column
	const uint32_t nD = nd + 1;
	for (uint32_t idx = threadIdx.x; idx < ne * nD; idx += blockDim.x) {
		const uint32_t ev = idx / nD, q = idx - ev * nD;
		double2 v = make_double2(1., 0.);
		if (q < nd) {
			const uint32_t r = q / W, g = q - r * W;
			const double2 S = Sx[(size_t)ev * R2 + 2 * r + (g >= nW0 ? 1 : 0)];
			const double2 x = A[(size_t)ev * W + g];
			v = make_double2(S.x * x.x + S.y * x.y, S.y * x.x - S.x * x.y);
		}
		dRows[(k0 + ev) * nD + q] = v;
	}
	const uint32_t nA = W + 1;
	for (uint32_t idx = threadIdx.x; idx < ne * nA; idx += blockDim.x) {
		const uint32_t ev = idx / nA, g = idx - ev * nA;
		aRows[(k0 + ev) * nA + g] = g < W ? A[(size_t)ev * W + g] : make_double2(1., 0.);
	}
}

// Per-device work area of the Hessian (the per-event rows, their weights and
// the production amplitudes), kept between calls and grown on demand: a call
// then allocates nothing on the device (allocating and freeing the 8.5 GB of
// rows at 2M events x 345 parameters took about 2 ms of a 13 ms call).  Calls
// on one device are serialised by the mutex; the area lives until the process
// ends.
struct HessWork {
	std::mutex m;
	char* buf = nullptr;
	size_t cap = 0;
};
constexpr int kHessDevices = 64;

HessWork& hess_work(int device)
{
	static HessWork work[kHessDevices];
	return work[device];
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

extern "C" int rpwa_lik_hessian(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                                double prodAmpFlat, uint32_t rank, double* hessian, double* flatTerms,
                                double* hessianFlat)
{
	clear_error();
	if (!ctx) {
		set_error("rpwa_lik_hessian: null context");
		return RPWA_ERR_INVALID;
	}
	if (!ctx->loaded) {
		set_error("rpwa_lik_hessian: decay amplitudes not loaded");
		return RPWA_ERR_STATE;
	}
	if (!prodAmps || !hessian || !flatTerms || !hessianFlat || rank == 0) {
		set_error("rpwa_lik_hessian: null output or rank 0");
		return RPWA_ERR_INVALID;
	}
	const AmpLayout& lay = ctx->lay;
	const uint32_t wmax = lay.wmax, W = lay.nW[0] + lay.nW[1];
	if ((uint64_t)nmbProdAmps != (uint64_t)rank * 2 * wmax) {
		set_error("rpwa_lik_hessian: nmbProdAmps must be rank*2*nmbWavesMax");
		return RPWA_ERR_INVALID;
	}
	try {
		RPWA_HIP_CHECK(hipSetDevice(ctx->device));
		const uint32_t nd = rank * W;  // valid (rank, sector, wave) entries
		const uint32_t dD = 2 * (nd + 1), dA = 2 * (W + 1);
		// event chunk: row buffers of at most ~8 GiB (2M events x 345 parameters in two chunks)
		const uint64_t chunk = std::max<uint64_t>(1024, (8ull << 30) / (16ull * (nd + 1)));
		const uint64_t cn = std::min<uint64_t>(chunk, lay.nEvents);
		if (ctx->device < 0 || ctx->device >= kHessDevices) {
			set_error("rpwa_lik_hessian: device index out of range");
			return RPWA_ERR_INVALID;
		}
		// one allocation: rows d [cn][nd + 1], rows a [cn][W + 1] (complex), T, weights wD, wA [cn]
		const size_t offA = align256(cn * (nd + 1) * sizeof(double2));
		const size_t offP = offA + align256(cn * (W + 1) * sizeof(double2));
		const size_t offWD = offP + align256((size_t)nmbProdAmps * sizeof(double2));
		const size_t offWA = offWD + align256(cn * sizeof(double));
		const size_t need = offWA + align256(cn * sizeof(double));
		HessWork& work = hess_work(ctx->device);
		std::lock_guard<std::mutex> lock(work.m);
		if (need > work.cap) {
			(void)hipFree(work.buf);
			work.buf = nullptr;
			work.cap = 0;
			if (hipMalloc(&work.buf, need) != hipSuccess) {
				work.buf = nullptr;
				set_error("rpwa_lik_hessian: device allocation failed");
				return RPWA_ERR_NOMEM;
			}
			work.cap = need;
		}
		double2* const dRows = reinterpret_cast<double2*>(work.buf);
		double2* const aRows = reinterpret_cast<double2*>(work.buf + offA);
		double2* const prod = reinterpret_cast<double2*>(work.buf + offP);
		double* const wD = reinterpret_cast<double*>(work.buf + offWD);
		double* const wA = reinterpret_cast<double*>(work.buf + offWA);
		hipError_t e = hipMemcpy(prod, prodAmps, (size_t)nmbProdAmps * sizeof(double2), hipMemcpyHostToDevice);
		// events per block of the row kernel: their rows and S within 48 KB of LDS
		const size_t perEvent = ((size_t)W + 2 * rank) * sizeof(double2);
		const uint32_t nEv = (uint32_t)std::max<size_t>(1, std::min<size_t>(16, (48u << 10) / perEvent));
		const size_t ldsBytes = (size_t)nEv * perEvent;
		std::vector<double> CD((size_t)dD * dD, 0.), CA((size_t)dA * dA, 0.), part;
		int st = RPWA_OK;
		for (uint64_t e0 = 0; e == hipSuccess && st == RPWA_OK && e0 < lay.nEvents; e0 += cn) {
			const uint64_t n = std::min<uint64_t>(cn, lay.nEvents - e0);
			const uint32_t nb = (uint32_t)((n + nEv - 1) / nEv);
			if (lay.precision == RPWA_PREC_MIXED)
				hipLaunchKernelGGL(hessian_rows_kernel<true>, dim3(nb), dim3(256), ldsBytes, 0, ctx->d_amps, e0, n,
				                   nEv, lay.rowLen, lay.secOff[1], lay.nW[0], lay.nW[1], wmax, (int)rank, prod,
				                   prodAmpFlat * prodAmpFlat, nd, dRows, aRows, wD, wA);
			else
				hipLaunchKernelGGL(hessian_rows_kernel<false>, dim3(nb), dim3(256), ldsBytes, 0, ctx->d_amps, e0, n,
				                   nEv, lay.rowLen, lay.secOff[1], lay.nW[0], lay.nW[1], wmax, (int)rank, prod,
				                   prodAmpFlat * prodAmpFlat, nd, dRows, aRows, wD, wA);
			e = hipGetLastError();
			if (e == hipSuccess)
				e = hipDeviceSynchronize();
			if (e != hipSuccess)
				break;
			st = syrk_run(ctx->device, reinterpret_cast<const double*>(dRows), n, dD, wD, true, nullptr, nullptr,
			              part, nullptr);
			if (st == RPWA_OK) {
				for (size_t i = 0; i < CD.size(); ++i)
					CD[i] += part[i];
				st = syrk_run(ctx->device, reinterpret_cast<const double*>(aRows), n, dA, wA, true, nullptr,
				              nullptr, part, nullptr);
				if (st == RPWA_OK)
					for (size_t i = 0; i < CA.size(); ++i)
						CA[i] += part[i];
			}
		}
		RPWA_HIP_CHECK(e);
		if (st != RPWA_OK)
			return st;
		// assemble [rank][2][wmax] x [rank][2][wmax] x 3 (pwaLikelihood.cc:697-753)
		const size_t ndFull = (size_t)rank * 2 * wmax;
		std::vector<int> valid(ndFull, -1), gwave(ndFull, 0);
		{
			uint32_t q = 0;
			for (uint32_t r = 0; r < rank; ++r)
				for (uint32_t s = 0; s < 2; ++s)
					for (uint32_t i = 0; i < lay.nW[s]; ++i) {
						const size_t a = ((size_t)r * 2 + s) * wmax + i;
						valid[a] = (int)q++;
						gwave[a] = (int)(s ? lay.nW[0] + i : i);
					}
		}
		auto cd = [&](size_t r, size_t c) { return CD[r * dD + c]; };
		auto ca = [&](size_t r, size_t c) { return CA[r * dA + c]; };
		std::memset(hessian, 0, ndFull * ndFull * 3 * sizeof(double));
		for (size_t a = 0; a < ndFull; ++a) {
			const int va = valid[a];
			if (va < 0) {
				flatTerms[2 * a] = flatTerms[2 * a + 1] = 0.;
				continue;
			}
			const size_t ra = a / (2 * wmax), sa = (a / wmax) % 2;
			for (size_t b = 0; b < ndFull; ++b) {
				const int vb = valid[b];
				if (vb < 0)
					continue;
				const size_t rb = b / (2 * wmax), sb = (b / wmax) % 2;
				double* h = hessian + (a * ndFull + b) * 3;
				h[0] = cd(2 * va, 2 * vb);
				h[1] = cd(2 * vb + 1, 2 * va);
				h[2] = cd(2 * va + 1, 2 * vb + 1);
				if (ra == rb && sa == sb) {
					const size_t ga = gwave[a], gb = gwave[b];
					const double re = ca(2 * ga, 2 * gb) + ca(2 * ga + 1, 2 * gb + 1);
					const double im = ca(2 * ga + 1, 2 * gb) - ca(2 * ga, 2 * gb + 1);
					h[0] += re;
					h[1] += im;
					h[2] += re;
				}
			}
			flatTerms[2 * a] = prodAmpFlat * cd(2 * va, 2 * nd);
			flatTerms[2 * a + 1] = prodAmpFlat * cd(2 * va + 1, 2 * nd);
		}
		*hessianFlat = prodAmpFlat * prodAmpFlat * cd(2 * nd, 2 * nd) + ca(2 * W, 2 * W);
		return RPWA_OK;
	} catch (const std::exception& ex) {
		set_error(std::string("rpwa_lik_hessian: ") + ex.what());
		return RPWA_ERR_NOMEM;
	}
}
