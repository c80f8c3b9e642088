// rng.h -- counter-based random numbers for the synthetic amplitude generator.
//
// Every draw is mix64(key(seed, stream) + golden * (counter + 1)), mix64 being
// the splitmix64 finaliser.  Uniforms are ((h >> 12) + 0.5) * 2^-52, exact in
// double and strictly inside (0, 1); normals come in Box-Muller pairs.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace rpwa_amd {
namespace rng {

constexpr uint32_t kStreamWaveScales = 0xFFFFFFF0u;

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t hash(uint64_t seed, uint32_t stream, uint64_t ctr)
{
	const uint64_t key = mix64(seed ^ (0xD1B54A32D192ED03ULL * ((uint64_t)stream + 1)));
	return mix64(key + 0x9E3779B97F4A7C15ULL * (ctr + 1));
}

__host__ __device__ inline double u01(uint64_t h)
{
	return ((double)(h >> 12) + 0.5) * 2.220446049250313080847263336181640625e-16;  // 2^-52
}

// Box-Muller pair of the slot counter q (uses counters 2q and 2q+1)
__host__ __device__ inline void normal_pair(uint64_t seed, uint32_t stream, uint64_t q, double& n0,
                                            double& n1)
{
	const double u1 = u01(hash(seed, stream, 2 * q));
	const double u2 = u01(hash(seed, stream, 2 * q + 1));
	const double r = sqrt(-2.0 * log(u1));
	const double th = 6.283185307179586476925286766559 * u2;
	n0 = r * cos(th);
	n1 = r * sin(th);
}

__host__ __device__ inline uint64_t slot_counter(uint64_t e, uint32_t W, uint32_t slot)
{
	return e * ((uint64_t)W + 2) + slot;
}

// (N(0,1) + i N(0,1)) of wave i of event e
__host__ __device__ inline double2 complex_normal(uint64_t seed, uint32_t stream, uint64_t e, uint32_t W,
                                                  uint32_t i)
{
	double a, b;
	normal_pair(seed, stream, slot_counter(e, W, i), a, b);
	return make_double2(a, b);
}

// per-event lognormal(0, 0.5) factor
__host__ __device__ inline double event_factor(uint64_t seed, uint32_t stream, uint64_t e, uint32_t W)
{
	double a, b;
	normal_pair(seed, stream, slot_counter(e, W, W), a, b);
	return exp(0.5 * a);
}

__host__ __device__ inline double accept_uniform(uint64_t seed, uint32_t stream, uint64_t e, uint32_t W)
{
	return u01(hash(seed, stream, 2 * slot_counter(e, W, W + 1)));
}

inline double wave_scale(uint64_t seed, uint32_t i)
{
	return 0.1 + 0.9 * u01(hash(seed, kStreamWaveScales, i));
}

inline void host_complex_normal(uint64_t seed, uint32_t stream, uint64_t k, double& re, double& im)
{
	normal_pair(seed, stream, k, re, im);
}

}  // namespace rng
}  // namespace rpwa_amd
