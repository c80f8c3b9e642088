/*
 * gen_oracle.c -- CPU restatement of the synthetic-amplitude generator.
 *
 * TEST INFRASTRUCTURE ONLY.  Restates the counter-based generator of the
 * product (splitmix64-keyed counters, Box-Muller normals, per-event lognormal
 * factor, accept-reject against the pool maximum) so that tests can produce
 * the benchmark workloads on the CPU and cross-check the GPU generator.  The
 * recipe itself is the benchmark's data recipe, not a reference function:
 *   A_i(e) = (N(0,1) + i N(0,1)) * s_i * exp(0.5 N(0,1)_e),  s_i ~ U[0.1, 1].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define STREAM_WAVE_SCALES 0xFFFFFFF0u

static uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

uint64_t orc_gen_hash(uint64_t seed, uint32_t stream, uint64_t ctr)
{
	const uint64_t key = mix64(seed ^ (0xD1B54A32D192ED03ULL * ((uint64_t)stream + 1)));
	return mix64(key + 0x9E3779B97F4A7C15ULL * (ctr + 1));
}

static double u01(uint64_t h) { return ((double)(h >> 12) + 0.5) * 2.220446049250313080847263336181640625e-16; }

static void normal_pair(uint64_t seed, uint32_t stream, uint64_t q, double* n0, double* n1)
{
	const double u1 = u01(orc_gen_hash(seed, stream, 2 * q));
	const double u2 = u01(orc_gen_hash(seed, stream, 2 * q + 1));
	const double r = sqrt(-2.0 * log(u1));
	const double th = 6.283185307179586476925286766559 * u2;
	*n0 = r * cos(th);
	*n1 = r * sin(th);
}

static uint64_t slot_counter(uint64_t e, uint32_t W, uint32_t slot) { return e * ((uint64_t)W + 2) + slot; }

void orc_gen_wave_scales(uint64_t seed, uint32_t W, double* scales)
{
	for (uint32_t i = 0; i < W; ++i)
		scales[i] = 0.1 + 0.9 * u01(orc_gen_hash(seed, STREAM_WAVE_SCALES, i));
}

void orc_gen_prod_amps(uint64_t seed, uint32_t stream, uint32_t rank, const uint32_t nW[2], double* prod)
{
	const uint32_t wmax = nW[0] > nW[1] ? nW[0] : nW[1];
	for (uint32_t r = 0; r < rank; ++r)
		for (uint32_t s = 0; s < 2; ++s)
			for (uint32_t j = 0; j < wmax; ++j) {
				const uint64_t k = ((uint64_t)r * 2 + s) * wmax + j;
				double re = 0, im = 0;
				if (j < nW[s] && j >= r) {
					normal_pair(seed, stream, k, &re, &im);
					if (j == r)
						im = 0;
				}
				prod[2 * k] = re;
				prod[2 * k + 1] = im;
			}
}

/* one event's W amplitudes */
static void event_amps(uint64_t seed, uint32_t stream, uint64_t e, uint32_t W, const double* scales, double* out)
{
	double a, b;
	normal_pair(seed, stream, slot_counter(e, W, W), &a, &b);
	const double lam = exp(0.5 * a);
	for (uint32_t i = 0; i < W; ++i) {
		double zr, zi;
		normal_pair(seed, stream, slot_counter(e, W, i), &zr, &zi);
		const double f = scales[i] * lam;
		out[2 * i] = zr * f;
		out[2 * i + 1] = zi * f;
	}
}

void orc_gen_amps(uint64_t seed, uint32_t stream, uint64_t e0, uint64_t n, uint32_t W, const double* scales,
                  double* out)
{
	for (uint64_t k = 0; k < n; ++k)
		event_amps(seed, stream, e0 + k, W, scales, out + 2 * k * W);
}

static double intensity(const double* A, const uint32_t nW[2], uint32_t rank, const double* prod)
{
	const uint32_t wmax = nW[0] > nW[1] ? nW[0] : nW[1];
	double sigma = 0;
	for (uint32_t r = 0; r < rank; ++r)
		for (uint32_t s = 0; s < 2; ++s) {
			double Sr = 0, Si = 0;
			for (uint32_t j = 0; j < nW[s]; ++j) {
				const uint32_t gw = s ? nW[0] + j : j;
				const double* t = prod + 2 * (((uint64_t)r * 2 + s) * wmax + j);
				const double ar = A[2 * gw], ai = A[2 * gw + 1];
				Sr += t[0] * ar - t[1] * ai;
				Si += t[0] * ai + t[1] * ar;
			}
			sigma += Sr * Sr + Si * Si;
		}
	return sigma;
}

/* accept-reject as the GPU generator does it; returns the pool size used
 * (0 on failure); idx receives the accepted pool indices */
uint64_t orc_gen_accept_reject(uint64_t seed, uint32_t stream, uint64_t nWanted, const uint32_t nW[2],
                               uint32_t rank, const double* prod, const double* scales, double* out,
                               uint64_t* idx, double* smaxOut)
{
	const uint32_t W = nW[0] + nW[1];
	uint64_t target = nWanted * 64 > 4096 ? nWanted * 64 : 4096;
	uint64_t pool = 0;
	double* sig = NULL;
	double smax = 0;
	double* A = (double*)malloc(sizeof(double) * 2 * W);
	for (int attempt = 0; attempt < 24; ++attempt) {
		sig = (double*)realloc(sig, sizeof(double) * target);
		for (uint64_t e = pool; e < target; ++e) {
			event_amps(seed, stream, e, W, scales, A);
			sig[e] = intensity(A, nW, rank, prod);
			if (sig[e] > smax)
				smax = sig[e];
		}
		pool = target;
		uint64_t n = 0;
		for (uint64_t e = 0; e < pool && n < nWanted; ++e) {
			const double u = u01(orc_gen_hash(seed, stream, 2 * slot_counter(e, W, W + 1)));
			if (u < sig[e] / smax)
				idx[n++] = e;
		}
		if (n >= nWanted) {
			for (uint64_t k = 0; k < nWanted; ++k)
				event_amps(seed, stream, idx[k], W, scales, out + 2 * k * W);
			free(sig);
			free(A);
			if (smaxOut)
				*smaxOut = smax;
			return pool;
		}
		target = pool * 2;
	}
	free(sig);
	free(A);
	return 0;
}
