// mlp.hip -- the learned initializer's network (include/rootpwa_amd_mlp.h).
//
// Products run on FP32 MFMA (v_mfma_f32_32x32x2f32, one 32x32 output tile per
// wave, operands straight from global memory: the matrices are small and stay
// in L2); bias, GELU, loss and Adam are elementwise kernels.  Every reduction
// (the K loop of a product, the bias gradients, the loss) runs in a fixed
// order, so training is deterministic for a given seed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"
#include "rootpwa_amd_mlp.h"

namespace rpwa_amd {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr float kInvSqrt2 = 0.70710678118654752440f;
constexpr float kInvSqrt2Pi = 0.39894228040143267794f;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * kInvSqrt2)); }

__device__ __forceinline__ float gelu_grad(float z)
{
	return 0.5f * (1.f + erff(z * kInvSqrt2)) + z * kInvSqrt2Pi * expf(-0.5f * z * z);
}

enum Epilogue { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2 };

// C[M][N] = op(A) op(B) (+ epilogue); op(A)(m, k) = TA ? A[k lda + m] : A[m lda + k],
// op(B)(k, n) = TB ? B[n ldb + k] : B[k ldb + n].  Block: 4 waves, 64 x 64 tile.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(const float* __restrict__ A, int lda, const float* __restrict__ B,
                                                   int ldb, float* __restrict__ C, int ldc, int M, int N, int K,
                                                   const float* __restrict__ bias, float* __restrict__ H)
{
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int m0 = blockIdx.y * 64 + 32 * (wave >> 1), n0 = blockIdx.x * 64 + 32 * (wave & 1);
	const int r = lane & 31, kk = lane >> 5;
	const int am = m0 + r, bn = n0 + r;
	f32x16 acc = f32x16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
	for (int k = 0; k < K; k += 2) {
		const int ka = k + kk;
		float a = 0.f, b = 0.f;
		if (am < M && ka < K)
			a = TA ? A[(size_t)ka * lda + am] : A[(size_t)am * lda + ka];
		if (bn < N && ka < K)
			b = TB ? B[(size_t)bn * ldb + ka] : B[(size_t)ka * ldb + bn];
		acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
	}
	// D[i][j]: column j = lane % 32, row i = 8 (v / 4) + 4 (lane / 32) + v % 4
#pragma unroll
	for (int v = 0; v < 16; ++v) {
		const int i = m0 + 8 * (v >> 2) + 4 * kk + (v & 3), jn = n0 + r;
		if (i < M && jn < N) {
			float z = acc[v];
			if (EPI != EPI_STORE)
				z += bias[jn];
			C[(size_t)i * ldc + jn] = z;
			if (EPI == EPI_BIAS_GELU)
				H[(size_t)i * ldc + jn] = gelu(z);
		}
	}
}

// dZ = dH * GELU'(Z) in place
__global__ void gelu_backward_kernel(float* __restrict__ dH, const float* __restrict__ Z, size_t n)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dH[i] *= gelu_grad(Z[i]);
}

// db[n] = sum over rows of dZ[row][n], rows in order
__global__ void column_sum_kernel(const float* __restrict__ dZ, int rows, int cols, float* __restrict__ db)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= cols)
		return;
	float s = 0.f;
	for (int r = 0; r < rows; ++r)
		s += dZ[(size_t)r * cols + c];
	db[c] = s;
}

// dY = 2 (Y - T) / n and per-row squared errors (summed in order afterwards)
__global__ void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, int rows, int cols,
                           float* __restrict__ dY, float* __restrict__ rowLoss)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= rows)
		return;
	const float scale = 2.f / ((float)rows * (float)cols);
	float s = 0.f;
	for (int c = 0; c < cols; ++c) {
		const float d = Y[(size_t)r * cols + c] - T[(size_t)r * cols + c];
		dY[(size_t)r * cols + c] = scale * d;
		s += d * d;
	}
	rowLoss[r] = s;
}

__global__ void loss_total_kernel(const float* __restrict__ rowLoss, int rows, int cols, float* __restrict__ loss,
                                  double* __restrict__ epochSum)
{
	if (threadIdx.x != 0 || blockIdx.x != 0)
		return;
	double s = 0.;
	for (int r = 0; r < rows; ++r)
		s += rowLoss[r];
	const float l = (float)(s / ((double)rows * (double)cols));
	*loss = l;
	if (epochSum)
		*epochSum += l;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, size_t n, float lr, float c1, float c2)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float gi = g[i];
	const float a = 0.9f * m1[i] + 0.1f * gi;
	const float b = 0.999f * m2[i] + 0.001f * gi * gi;
	m1[i] = a;
	m2[i] = b;
	p[i] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
}

// rows idx[r] of X [n][cols] -> out [r][cols]
__global__ void gather_kernel(const float* __restrict__ X, const uint32_t* __restrict__ idx, int rows, int cols,
                              float* __restrict__ out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (size_t)rows * cols)
		return;
	const size_t r = i / cols, c = i % cols;
	out[i] = X[(size_t)idx[r] * cols + c];
}

uint64_t mix64(uint64_t x)
{
	x += 0x9E3779B97F4A7C15ull;
	x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
	x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
	return x ^ (x >> 31);
}

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

struct rpwa_mlp {
	int device = 0;
	std::vector<int> dims;         // nIn, hidden..., nOut
	std::vector<size_t> wOff, bOff;  // parameter offsets
	size_t nParams = 0;
	hipStream_t stream = nullptr;
	float* d_p = nullptr;   // parameters
	float* d_g = nullptr;   // gradient
	float* d_m1 = nullptr;  // Adam moments
	float* d_m2 = nullptr;
	uint64_t adamStep = 0;
	// activations of the current capacity (rows)
	uint32_t capRows = 0;
	std::vector<float*> d_Z, d_H;  // per layer: pre-activation [rows][out], activation (hidden)
	float* d_x = nullptr;
	float* d_t = nullptr;
	float* d_dA = nullptr;  // backward buffers [rows][max width]
	float* d_dB = nullptr;
	float* d_rowLoss = nullptr;
	float* d_loss = nullptr;
	double* d_epoch = nullptr;
	int maxWidth = 0;

	~rpwa_mlp() { release(); }
	void release_acts()
	{
		for (float* p : d_Z)
			(void)hipFree(p);
		for (float* p : d_H)
			(void)hipFree(p);
		d_Z.clear();
		d_H.clear();
		for (float** p : {&d_x, &d_t, &d_dA, &d_dB, &d_rowLoss})
			if (*p) {
				(void)hipFree(*p);
				*p = nullptr;
			}
		capRows = 0;
	}
	void release()
	{
		release_acts();
		for (float** p : {&d_p, &d_g, &d_m1, &d_m2, &d_loss})
			if (*p) {
				(void)hipFree(*p);
				*p = nullptr;
			}
		if (d_epoch) {
			(void)hipFree(d_epoch);
			d_epoch = nullptr;
		}
		if (stream) {
			(void)hipStreamDestroy(stream);
			stream = nullptr;
		}
	}
	int nLayers() const { return (int)dims.size() - 1; }
};

namespace {

int check_mlp(const rpwa_mlp* m, const char* fn)
{
	if (!m) {
		set_error(std::string(fn) + ": null network");
		return RPWA_ERR_INVALID;
	}
	return RPWA_OK;
}

int ensure_rows(rpwa_mlp* m, uint32_t rows)
{
	if (rows <= m->capRows)
		return RPWA_OK;
	m->release_acts();
	const int L = m->nLayers();
	for (int l = 0; l < L; ++l) {
		float *z = nullptr, *h = nullptr;
		RPWA_HIP_CHECK(hipMalloc(&z, (size_t)rows * m->dims[l + 1] * sizeof(float)));
		m->d_Z.push_back(z);
		if (l + 1 < L) {
			RPWA_HIP_CHECK(hipMalloc(&h, (size_t)rows * m->dims[l + 1] * sizeof(float)));
		}
		m->d_H.push_back(h);
	}
	RPWA_HIP_CHECK(hipMalloc(&m->d_x, (size_t)rows * m->dims[0] * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_t, (size_t)rows * m->dims[L] * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_dA, (size_t)rows * m->maxWidth * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_dB, (size_t)rows * m->maxWidth * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_rowLoss, (size_t)rows * sizeof(float)));
	m->capRows = rows;
	return RPWA_OK;
}

template <bool TA, bool TB, int EPI>
int gemm(rpwa_mlp* m, const float* A, int lda, const float* B, int ldb, float* C, int ldc, int M, int N, int K,
         const float* bias = nullptr, float* H = nullptr)
{
	const dim3 grid((N + 63) / 64, (M + 63) / 64);
	hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, m->stream, A, lda, B, ldb, C, ldc, M, N, K,
	                   bias, H);
	RPWA_HIP_CHECK(hipGetLastError());
	return RPWA_OK;
}

// forward pass of rows in d_x; output in d_Z[L - 1]
int forward(rpwa_mlp* m, uint32_t rows)
{
	const int L = m->nLayers();
	const float* in = m->d_x;
	for (int l = 0; l < L; ++l) {
		const int nin = m->dims[l], nout = m->dims[l + 1];
		const float* W = m->d_p + m->wOff[l];
		const float* b = m->d_p + m->bOff[l];
		int rc;
		if (l + 1 < L)
			rc = gemm<false, false, EPI_BIAS_GELU>(m, in, nin, W, nout, m->d_Z[l], nout, (int)rows, nout, nin, b,
			                                       m->d_H[l]);
		else
			rc = gemm<false, false, EPI_BIAS>(m, in, nin, W, nout, m->d_Z[l], nout, (int)rows, nout, nin, b);
		if (rc != RPWA_OK)
			return rc;
		in = m->d_H[l];
	}
	return RPWA_OK;
}

// loss (into d_loss, and added to *epochSum if given) and gradient into d_g, for rows of d_x / d_t
int backward(rpwa_mlp* m, uint32_t rows, double* epochSum)
{
	const int L = m->nLayers();
	const int nOut = m->dims[L];
	float* dZ = m->d_dA;
	float* dH = m->d_dB;
	hipLaunchKernelGGL(mse_kernel, dim3((rows + 255) / 256), dim3(256), 0, m->stream, m->d_Z[L - 1], m->d_t,
	                   (int)rows, nOut, dZ, m->d_rowLoss);
	RPWA_HIP_CHECK(hipGetLastError());
	hipLaunchKernelGGL(loss_total_kernel, dim3(1), dim3(64), 0, m->stream, m->d_rowLoss, (int)rows, nOut, m->d_loss,
	                   epochSum);
	RPWA_HIP_CHECK(hipGetLastError());
	for (int l = L - 1; l >= 0; --l) {
		const int nin = m->dims[l], nout = m->dims[l + 1];
		const float* in = l == 0 ? m->d_x : m->d_H[l - 1];
		// dW = in^T dZ  ([nin][nout]), db = column sums of dZ
		int rc = gemm<true, false, EPI_STORE>(m, in, nin, dZ, nout, m->d_g + m->wOff[l], nout, nin, nout, (int)rows);
		if (rc != RPWA_OK)
			return rc;
		hipLaunchKernelGGL(column_sum_kernel, dim3((nout + 255) / 256), dim3(256), 0, m->stream, dZ, (int)rows, nout,
		                   m->d_g + m->bOff[l]);
		RPWA_HIP_CHECK(hipGetLastError());
		if (l > 0) {
			// dH = dZ W^T ([rows][nin]), then dZ_{l-1} = dH * GELU'(Z_{l-1})
			rc = gemm<false, true, EPI_STORE>(m, dZ, nout, m->d_p + m->wOff[l], nout, dH, nin, (int)rows, nin, nout);
			if (rc != RPWA_OK)
				return rc;
			const size_t n = (size_t)rows * nin;
			hipLaunchKernelGGL(gelu_backward_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, m->stream, dH,
			                   m->d_Z[l - 1], n);
			RPWA_HIP_CHECK(hipGetLastError());
			std::swap(dZ, dH);
		}
	}
	return RPWA_OK;
}

}  // namespace

extern "C" int rpwa_mlp_create(rpwa_mlp** out, int device, uint32_t nIn, uint32_t nHidden, uint32_t nHiddenLayers,
                               uint32_t nOut, uint64_t seed)
{
	clear_error();
	if (!out || nIn == 0 || nOut == 0 || (nHiddenLayers > 0 && nHidden == 0) || nHiddenLayers > 16) {
		set_error("rpwa_mlp_create: null pointer or invalid shape");
		return RPWA_ERR_INVALID;
	}
	*out = nullptr;
	int nDev = 0;
	if (hipGetDeviceCount(&nDev) != hipSuccess || device < 0 || device >= nDev) {
		set_error("rpwa_mlp_create: no GPU with index " + std::to_string(device));
		return RPWA_ERR_INVALID;
	}
	std::unique_ptr<rpwa_mlp> m(new rpwa_mlp);
	m->device = device;
	m->dims.push_back((int)nIn);
	for (uint32_t l = 0; l < nHiddenLayers; ++l)
		m->dims.push_back((int)nHidden);
	m->dims.push_back((int)nOut);
	m->maxWidth = *std::max_element(m->dims.begin(), m->dims.end());
	std::vector<float> p;
	for (int l = 0; l < m->nLayers(); ++l) {
		const int nin = m->dims[l], nout = m->dims[l + 1];
		m->wOff.push_back(p.size());
		const double a = std::sqrt(6.0 / (nin + nout));  // Glorot uniform
		for (int k = 0; k < nin * nout; ++k) {
			const uint64_t h = mix64(seed ^ mix64(((uint64_t)l << 40) + (uint64_t)k));
			const double u = ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
			p.push_back((float)((2. * u - 1.) * a));
		}
		m->bOff.push_back(p.size());
		p.insert(p.end(), nout, 0.f);
	}
	m->nParams = p.size();
	RPWA_HIP_CHECK(hipSetDevice(device));
	RPWA_HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
	RPWA_HIP_CHECK(hipMalloc(&m->d_p, m->nParams * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_g, m->nParams * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_m1, m->nParams * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_m2, m->nParams * sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_loss, sizeof(float)));
	RPWA_HIP_CHECK(hipMalloc(&m->d_epoch, sizeof(double)));
	RPWA_HIP_CHECK(hipMemcpy(m->d_p, p.data(), m->nParams * sizeof(float), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemset(m->d_m1, 0, m->nParams * sizeof(float)));
	RPWA_HIP_CHECK(hipMemset(m->d_m2, 0, m->nParams * sizeof(float)));
	*out = m.release();
	return RPWA_OK;
}

extern "C" int rpwa_mlp_destroy(rpwa_mlp* m)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_destroy") != RPWA_OK)
		return RPWA_ERR_INVALID;
	(void)hipSetDevice(m->device);
	if (m->stream)
		(void)hipStreamSynchronize(m->stream);
	delete m;
	return RPWA_OK;
}

extern "C" int rpwa_mlp_nmb_params(const rpwa_mlp* m, uint64_t* n)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_nmb_params") != RPWA_OK || !n) {
		set_error("rpwa_mlp_nmb_params: null pointer");
		return RPWA_ERR_INVALID;
	}
	*n = m->nParams;
	return RPWA_OK;
}

extern "C" int rpwa_mlp_get_params(const rpwa_mlp* m, float* params)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_get_params") != RPWA_OK || !params) {
		set_error("rpwa_mlp_get_params: null pointer");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(m->device));
	RPWA_HIP_CHECK(hipStreamSynchronize(m->stream));
	RPWA_HIP_CHECK(hipMemcpy(params, m->d_p, m->nParams * sizeof(float), hipMemcpyDeviceToHost));
	return RPWA_OK;
}

extern "C" int rpwa_mlp_set_params(rpwa_mlp* m, const float* params)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_set_params") != RPWA_OK || !params) {
		set_error("rpwa_mlp_set_params: null pointer");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(m->device));
	RPWA_HIP_CHECK(hipStreamSynchronize(m->stream));
	RPWA_HIP_CHECK(hipMemcpy(m->d_p, params, m->nParams * sizeof(float), hipMemcpyHostToDevice));
	RPWA_HIP_CHECK(hipMemset(m->d_m1, 0, m->nParams * sizeof(float)));
	RPWA_HIP_CHECK(hipMemset(m->d_m2, 0, m->nParams * sizeof(float)));
	m->adamStep = 0;
	return RPWA_OK;
}

extern "C" int rpwa_mlp_forward(rpwa_mlp* m, const float* x, uint32_t nRows, float* y)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_forward") != RPWA_OK || !x || !y || nRows == 0) {
		set_error("rpwa_mlp_forward: null pointer or no rows");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(m->device));
	const int rc = ensure_rows(m, nRows);
	if (rc != RPWA_OK)
		return rc;
	RPWA_HIP_CHECK(hipMemcpyAsync(m->d_x, x, (size_t)nRows * m->dims[0] * sizeof(float), hipMemcpyHostToDevice,
	                              m->stream));
	if (forward(m, nRows) != RPWA_OK)
		return RPWA_ERR_HIP;
	RPWA_HIP_CHECK(hipMemcpyAsync(y, m->d_Z.back(), (size_t)nRows * m->dims.back() * sizeof(float),
	                              hipMemcpyDeviceToHost, m->stream));
	RPWA_HIP_CHECK(hipStreamSynchronize(m->stream));
	return RPWA_OK;
}

extern "C" int rpwa_mlp_loss_grad(rpwa_mlp* m, const float* x, const float* target, uint32_t nRows, float* loss,
                                  float* grad)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_loss_grad") != RPWA_OK || !x || !target || !loss || nRows == 0) {
		set_error("rpwa_mlp_loss_grad: null pointer or no rows");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(m->device));
	int rc = ensure_rows(m, nRows);
	if (rc != RPWA_OK)
		return rc;
	RPWA_HIP_CHECK(hipMemcpyAsync(m->d_x, x, (size_t)nRows * m->dims[0] * sizeof(float), hipMemcpyHostToDevice,
	                              m->stream));
	RPWA_HIP_CHECK(hipMemcpyAsync(m->d_t, target, (size_t)nRows * m->dims.back() * sizeof(float),
	                              hipMemcpyHostToDevice, m->stream));
	if ((rc = forward(m, nRows)) != RPWA_OK || (rc = backward(m, nRows, nullptr)) != RPWA_OK)
		return rc;
	RPWA_HIP_CHECK(hipMemcpyAsync(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
	if (grad)
		RPWA_HIP_CHECK(hipMemcpyAsync(grad, m->d_g, m->nParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
	RPWA_HIP_CHECK(hipStreamSynchronize(m->stream));
	return RPWA_OK;
}

extern "C" int rpwa_mlp_train(rpwa_mlp* m, const float* x, const float* target, uint32_t nRows, uint32_t batchSize,
                              uint32_t nEpochs, float lr, uint64_t seed, float* lossPerEpoch)
{
	clear_error();
	if (check_mlp(m, "rpwa_mlp_train") != RPWA_OK || !x || !target || nRows == 0 || batchSize == 0) {
		set_error("rpwa_mlp_train: null pointer, no rows or zero batch size");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(m->device));
	const uint32_t B = std::min(batchSize, nRows);
	int rc = ensure_rows(m, B);
	if (rc != RPWA_OK)
		return rc;
	const int nIn = m->dims[0], nOut = m->dims.back();
	float *d_X = nullptr, *d_T = nullptr;
	uint32_t* d_idx = nullptr;
	auto cleanup = [&]() {
		if (d_X)
			(void)hipFree(d_X);
		if (d_T)
			(void)hipFree(d_T);
		if (d_idx)
			(void)hipFree(d_idx);
	};
	if (hipMalloc(&d_X, (size_t)nRows * nIn * sizeof(float)) != hipSuccess ||
	    hipMalloc(&d_T, (size_t)nRows * nOut * sizeof(float)) != hipSuccess ||
	    hipMalloc(&d_idx, (size_t)nRows * sizeof(uint32_t)) != hipSuccess) {
		cleanup();
		set_error("rpwa_mlp_train: device allocation failed");
		return RPWA_ERR_NOMEM;
	}
	if (hipMemcpy(d_X, x, (size_t)nRows * nIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(d_T, target, (size_t)nRows * nOut * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
		cleanup();
		set_error("rpwa_mlp_train: copy failed");
		return RPWA_ERR_HIP;
	}
	std::vector<uint32_t> perm(nRows);
	for (uint32_t epoch = 0; epoch < nEpochs && rc == RPWA_OK; ++epoch) {
		// Fisher-Yates shuffle from counter-based draws
		for (uint32_t i = 0; i < nRows; ++i)
			perm[i] = i;
		for (uint32_t i = nRows - 1; i > 0; --i) {
			const uint64_t h = mix64(seed ^ mix64(((uint64_t)epoch << 32) + i));
			std::swap(perm[i], perm[(uint32_t)(h % (i + 1))]);
		}
		if (hipMemcpyAsync(d_idx, perm.data(), nRows * sizeof(uint32_t), hipMemcpyHostToDevice, m->stream) !=
		        hipSuccess ||
		    hipMemsetAsync(m->d_epoch, 0, sizeof(double), m->stream) != hipSuccess) {
			rc = RPWA_ERR_HIP;
			break;
		}
		const uint32_t nBatches = nRows / B;  // full batches only
		for (uint32_t bi = 0; bi < nBatches && rc == RPWA_OK; ++bi) {
			const size_t ni = (size_t)B * nIn, no = (size_t)B * nOut;
			hipLaunchKernelGGL(gather_kernel, dim3((uint32_t)((ni + 255) / 256)), dim3(256), 0, m->stream, d_X,
			                   d_idx + (size_t)bi * B, (int)B, nIn, m->d_x);
			hipLaunchKernelGGL(gather_kernel, dim3((uint32_t)((no + 255) / 256)), dim3(256), 0, m->stream, d_T,
			                   d_idx + (size_t)bi * B, (int)B, nOut, m->d_t);
			if ((rc = forward(m, B)) != RPWA_OK || (rc = backward(m, B, m->d_epoch)) != RPWA_OK)
				break;
			++m->adamStep;
			const float c1 = (float)(1. / (1. - std::pow(0.9, (double)m->adamStep)));
			const float c2 = (float)(1. / (1. - std::pow(0.999, (double)m->adamStep)));
			hipLaunchKernelGGL(adam_kernel, dim3((uint32_t)((m->nParams + 255) / 256)), dim3(256), 0, m->stream,
			                   m->d_p, m->d_g, m->d_m1, m->d_m2, m->nParams, lr, c1, c2);
			if (hipGetLastError() != hipSuccess)
				rc = RPWA_ERR_HIP;
		}
		if (rc == RPWA_OK && lossPerEpoch) {
			double s = 0.;
			if (hipMemcpyAsync(&s, m->d_epoch, sizeof(double), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
			    hipStreamSynchronize(m->stream) != hipSuccess)
				rc = RPWA_ERR_HIP;
			lossPerEpoch[epoch] = (float)(s / std::max(nBatches, 1u));
		}
	}
	if (hipStreamSynchronize(m->stream) != hipSuccess && rc == RPWA_OK)
		rc = RPWA_ERR_HIP;
	cleanup();
	if (rc != RPWA_OK && rpwa_last_error()[0] == 0)
		set_error("rpwa_mlp_train: a kernel launch or copy failed");
	return rc;
}
