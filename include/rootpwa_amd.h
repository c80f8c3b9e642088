/*
 * rootpwa_amd.h -- C ABI of the MI355X (gfx950) backend for ROOTPWA's unbinned
 * extended log-likelihood.
 *
 * This is the lower boundary of rpwa::pwaLikelihood: it replaces
 * rpwa::cuda::likelihoodInterface<cuda::complex<T>> (cuda/likelihoodInterface.cuh:72-166),
 * which the reference calls from partialWaveFit/pwaLikelihood.cc:352-357 (DoEval),
 * :533-541 (Gradient) and :1357-1377 (finishInit).  Differences from the
 * reference interface, all deliberate:
 *   - a context handle instead of a static singleton, so several likelihoods
 *     (and Clone()s sharing one context through rpwa_lik_retain) can coexist;
 *   - every entry point returns a status code and never throws or exits;
 *     rpwa_last_error() returns the message of the last failure on this thread;
 *   - 64-bit sizes (the reference's `unsigned int` byte counts overflow above
 *     4 GiB, likelihoodInterface.cu:236);
 *   - a fused value+gradient entry point (rpwa_lik_fdf) for pwaLikelihood::FdF,
 *     which has no GPU branch in the reference.
 * All calls are host-synchronous: results are in host memory on return.
 * Complex numbers are interleaved (re, im) doubles, the layout of
 * std::complex<double>.
 */
#ifndef ROOTPWA_AMD_H
#define ROOTPWA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RPWA_OK            0
#define RPWA_ERR_INVALID   1  /* bad argument (null pointer, size, rank ...) */
#define RPWA_ERR_STATE     2  /* call order violated (e.g. eval before amplitudes) */
#define RPWA_ERR_HIP       3  /* HIP runtime error (message in rpwa_last_error) */
#define RPWA_ERR_NOMEM     4  /* device or pinned host allocation failed */
#define RPWA_ERR_UNSUPPORTED 5 /* shape outside the built kernel variants */

/* amplitude storage / arithmetic of a likelihood context */
#define RPWA_PREC_FP64     0  /* complex128 amplitudes, FP64 arithmetic: the plain drop-in */
#define RPWA_PREC_MIXED    1  /* complex64 amplitudes (rounded after FP64 normalisation),
                                 FP64 arithmetic, FP64 compensated event sum */

/* host layouts accepted by rpwa_lik_load_decay_amps */
#define RPWA_LAYOUT_REFL_WAVE_EVT 0 /* [2][nmbWavesMax][nmbEvents], zero padded: the array
                                        pwaLikelihood::finishInit builds for the GPU
                                        (pwaLikelihood.cc:1357-1374) */
#define RPWA_LAYOUT_EVT_WAVE      1 /* [nmbEvents][nmbWavesRefl[0] + nmbWavesRefl[1]]:
                                        refl -1 waves first, then refl +1 waves */

typedef struct rpwa_lik_ctx rpwa_lik_ctx;

/* Text of the last error raised on the calling thread ("" if none). */
const char* rpwa_last_error(void);

/* Number of visible GPUs (0 when none); never fails. */
int rpwa_device_count(void);

/* Create a context on `device` (replaces likelihoodInterface::initCudaDevice,
 * likelihoodInterface.cu:141).  precision: RPWA_PREC_*. */
int rpwa_lik_create(rpwa_lik_ctx** ctx, int device, int precision);
/* Add a reference to a context (pwaLikelihood::Clone shares the device state). */
int rpwa_lik_retain(rpwa_lik_ctx* ctx);
/* Drop a reference; the last one frees the device memory
 * (replaces likelihoodInterface::closeCudaDevice, likelihoodInterface.cu:180). */
int rpwa_lik_release(rpwa_lik_ctx* ctx);

/* Upload the (already normalised) decay amplitudes once
 * (replaces likelihoodInterface::init / loadDecayAmps, likelihoodInterface.cu:119,195).
 * nmbDecayAmps = number of complex values in `decayAmps` for the given layout. */
int rpwa_lik_load_decay_amps(rpwa_lik_ctx* ctx, const double* decayAmps, uint64_t nmbDecayAmps,
                             uint64_t nmbEvents, const uint32_t nmbWavesRefl[2], int layout);

/* Same, from amplitudes already in device memory on the context's device, in
 * RPWA_LAYOUT_EVT_WAVE order.  waveNorm (host, [W0+W1], may be NULL) divides
 * each wave's column (real and imaginary parts) in FP64 before the amplitudes
 * are stored: with waveNorm = sqrt(Re I^norm_ii) it is the normalisation
 * amps /= sqrt(normInt.real()) of pwaLikelihood.cc:1253-1260, with the same
 * rounding as that host division. */
int rpwa_lik_load_decay_amps_device(rpwa_lik_ctx* ctx, const double* d_amps, uint64_t nmbEvents,
                                    const uint32_t nmbWavesRefl[2], const double* waveNorm);

/* Event part of -lnL: sum_e -ln( sum_r sum_refl |sum_i T_r,refl,i A_refl,i(e)|^2 + flat^2 )
 * (replaces likelihoodInterface::logLikelihood, likelihoodInterface.cuh:112-115).
 * prodAmps: [rank][2][nmbWavesMax] complex, nmbProdAmps = rank*2*nmbWavesMax.
 * Any rank: up to 4 in one fused pass, higher ranks in blocks of 4 through the
 * LDS-staged kernel (2 x blocks passes over the amplitudes). */
int rpwa_lik_log_likelihood(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                            double prodAmpFlat, uint32_t rank, double* logLikelihood);

/* Event part of the gradient: derivatives[r][refl][i] = sum_e -(2/sigma_e) S_r,refl(e)
 * conj(A_refl,i(e)), derivativeFlat = sum_e -(2/sigma_e) flat
 * (replaces likelihoodInterface::logLikelihoodDeriv, likelihoodInterface.cuh:117-122). */
int rpwa_lik_log_likelihood_deriv(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                                  double prodAmpFlat, uint32_t rank, double* derivatives,
                                  double* derivativeFlat);

/* Both event parts from one pass over the amplitudes.  No counterpart in the
 * reference (FdF has no GPU branch, pwaLikelihood.cc:172-315).  The value is
 * bitwise identical to rpwa_lik_log_likelihood's at the same prodAmps. */
int rpwa_lik_fdf(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                 double prodAmpFlat, uint32_t rank, double* logLikelihood,
                 double* derivatives, double* derivativeFlat);

/* Asynchronous form of rpwa_lik_fdf: rpwa_lik_submit starts the evaluation on
 * the context's stream and returns at once (the production amplitudes are
 * copied before it returns); rpwa_lik_collect waits for it and writes the
 * results (any output may be NULL).  At most one evaluation per context may be
 * pending; the results are bitwise those of rpwa_lik_fdf.  Lets the host work
 * on the normalisation term while the GPU streams the amplitudes. */
int rpwa_lik_submit(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                    double prodAmpFlat, uint32_t rank);
int rpwa_lik_collect(rpwa_lik_ctx* ctx, double* logLikelihood, double* derivatives,
                     double* derivativeFlat);

/* Event part of the Hessian (pwaLikelihood.cc:697-753): for a, b over the
 * [rank][2][nmbWavesMax] production amplitudes, hessian[a][b][3] holds the
 * real/real, real/imag and imag/imag event sums; flatTerms[a] (complex) the
 * mixed flat-wave terms and hessianFlat the flat/flat term.  The per-event rows
 * live in a device work area shared by the calls on one GPU (serialised),
 * grown to the largest call and kept until the process ends: (nEvents or 8 GiB
 * worth of rows) x 16 (rank W + W + 2) bytes, 8.5 GB at 2M events x 345
 * parameters. */
int rpwa_lik_hessian(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                     double prodAmpFlat, uint32_t rank, double* hessian, double* flatTerms,
                     double* hessianFlat);

/* Accumulated time spent in this context's kernels [s]
 * (likelihoodInterface::kernelTime / resetKernelTime, likelihoodInterface.cuh:124-125),
 * measured with HIP events bound to the kernel launches of the timed
 * evaluations while timing is on (the sum of the event kernels' and the
 * reduction's own device time, without launch gaps; rpwa_lik_set_timing, or
 * RPWA_KERNEL_TIMING=N in the environment; off by default: the events cost
 * about 15 us of host time per timed evaluation).  rpwa_lik_kernel_times
 * counts the timed evaluations.  Either way an evaluation's completion is read
 * from flags the reduction writes into host memory. */
int rpwa_lik_kernel_time(rpwa_lik_ctx* ctx, double* seconds);
int rpwa_lik_reset_kernel_time(rpwa_lik_ctx* ctx);
/* Kernel timing for the following evaluations: enable = N > 0 times every N-th
 * evaluation of the context (1: every one), 0 (or negative) turns it off. */
int rpwa_lik_set_timing(rpwa_lik_ctx* ctx, int enable);

/* Split of the accumulated kernel time: event kernel and fixed-order
 * reduction [s], and the number of evaluations they cover. */
int rpwa_lik_kernel_times(rpwa_lik_ctx* ctx, double* eventKernel, double* reduceKernel,
                          uint64_t* nmbEvaluations);

/* Shape and launch configuration of a context:
 * info[0] events, [1] W(refl -1), [2] W(refl +1), [3] precision, [4] lanes per event,
 * [5] slots per lane (events per tile for the staged kernel), [6] workgroups, [7] threads per
 * workgroup, [8] bytes of amplitudes; [4]-[6] for the lowest rank evaluated so far. */
int rpwa_lik_info(rpwa_lik_ctx* ctx, uint64_t info[9]);

/* Event tiles: before the first evaluation of a given rank the context times
 * every feasible kernel variant on its own amplitudes, together with the
 * reduction of its workgroup rows, and keeps the fastest for that rank;
 * results of a rank always come from one variant.  Variants:
 * lanes per event in {8,16,32,64} (register tiles: a group of lanes holds one
 * event's amplitude row and the production amplitudes of its columns) and
 * lanes per event 1 (LDS-staged: tiles of 64, 32 or 16 events, each timed,
 * pass through LDS, one lane contracts one event, production amplitudes come
 * through the scalar unit; rpwa_lik_set_tile(ctx, 1) takes the largest tile
 * that fits), and for complex64 rows of at most 128 waves the LDS-staged kernel
 * with its gradient phase on FP32 MFMA (tiles of 64 and 32 events; FP32
 * products of the amplitudes with u = -2S/sigma rounded to FP32, summed per
 * tile on the matrix cores and added in FP64; S, sigma and the log sum stay
 * FP64).  rpwa_lik_set_autotune(ctx, 0) (or RPWA_AUTOTUNE=0 in the
 * environment) keeps the modelled default instead; rpwa_lik_set_tile fixes the
 * variant of every rank up to 4 (RPWA_ERR_UNSUPPORTED if the row does not fit it);
 * ranks above 4 always run the LDS-staged kernel (the largest tile that fits, or
 * the tile fixed by rpwa_lik_set_stage_tile). */
int rpwa_lik_set_autotune(rpwa_lik_ctx* ctx, int enable);
int rpwa_lik_set_tile(rpwa_lik_ctx* ctx, uint32_t lanesPerEvent);
/* Fixes the LDS-staged kernel with tiles of eventsPerTile (16, 32 or 64)
 * events for every rank; RPWA_ERR_UNSUPPORTED if such a tile does not fit.
 * No counterpart in the reference: a tuning and testing control like
 * rpwa_lik_set_tile, for the tile sizes the autotuner chooses between. */
int rpwa_lik_set_stage_tile(rpwa_lik_ctx* ctx, uint32_t eventsPerTile);
/* As rpwa_lik_set_stage_tile, choosing also the variant: 0 = S and the
 * gradient on FP64 VALU; 1 = the gradient phase on FP32 MFMA (complex64
 * amplitudes, rows of at most 128 waves, 64 or 32 events per tile); 2 = the
 * wave-tile variant, each wave computing S of 16 events on FP64 MFMA, sigma and
 * u inside the wave and their gradient on FP32 MFMA (complex64, rows of at most
 * 96 waves, 64 events per tile; ranks 1 and 2, higher ranks keep variant 0).
 * No counterpart in the reference. */
int rpwa_lik_set_stage_variant(rpwa_lik_ctx* ctx, uint32_t eventsPerTile, int variant);
/* The event-kernel variant of a rank: variant[0] lanes per event (1: LDS-staged),
 * [1] slots per lane (events per tile for the staged kernel), [2] the staged
 * variant (0, 1 or 2 as in rpwa_lik_set_stage_variant), [3] workgroups.  No
 * counterpart in the reference. */
int rpwa_lik_kernel_variant(rpwa_lik_ctx* ctx, uint32_t rank, uint32_t variant[4]);

/* Stream on which the context launches its kernels (for event timing). */
int rpwa_lik_stream(rpwa_lik_ctx* ctx, void** hipStream);

/* ---------------------------------------------------------------------------
 * Device memory helpers, so that a caller needs no second GPU runtime to stage
 * amplitudes (tests, the benchmark, C++ drivers).  No counterpart in the
 * reference (its CUDA backend only copies host arrays, likelihoodInterface.cu:236-239).
 * --------------------------------------------------------------------------- */
#define RPWA_COPY_H2D 0
#define RPWA_COPY_D2H 1
#define RPWA_COPY_D2D 2
int rpwa_dev_alloc(int device, uint64_t bytes, void** ptr);
int rpwa_dev_free(void* ptr);
int rpwa_dev_copy(void* dst, const void* src, uint64_t bytes, int kind);
int rpwa_dev_synchronize(int device);
/* name (NUL-terminated, truncated to nameLen), compute units, global memory
 * bytes, peak engine clock [kHz] */
int rpwa_dev_info(int device, char* name, int nameLen, int* nCu, uint64_t* totalMem, int* clockKhz);

/* Whether the amplitudes of one wave (column `wave` of d_amps [nmbEvents][nmbWaves]
 * complex128, device memory) are non-zero for any event: *nonzero = 1, else 0.
 * The device-memory form of addAmplitude's zero-amplitude check
 * (partialWaveFit/pwaLikelihood.cc:1247-1264), which fixes the production
 * amplitudes of a wave whose acceptance integral and amplitudes are all zero
 * and rejects non-zero amplitudes of a wave with a zero normalisation integral. */
int rpwa_wave_nonzero_device(int device, const double* d_amps, uint64_t nmbEvents, uint32_t nmbWaves,
                             uint32_t wave, int* nonzero);

/* HBM bandwidth probe: GB/s of a streaming read and of a copy (read + write
 * bytes) over `bytes` of device memory, each the best median over several
 * variants (4 to 16 16-byte loads in flight per thread, temporal and
 * non-temporal loads, 4 to 16 workgroups per compute unit): the ceiling a
 * streaming kernel reaches on the device.  The counterpart of the reference's
 * bandwidth microbenchmarks (cuda/test/benchmarkMemoryBandwidth.cu:85-260). */
int rpwa_bandwidth_probe(int device, uint64_t bytes, int repetitions, double* readGBs, double* copyGBs);

/* ---------------------------------------------------------------------------
 * Normalisation integrals: ampIntegralMatrix::integrate
 * (decayAmplitude/ampIntegralMatrix.cc:300-488) for amplitudes with one
 * incoherent sub-amplitude, as an FP64-MFMA Hermitian rank-k update.
 * outSums[W][W] (complex): sum_e weight_e A_i(e) conj(A_j(e)), weight_e = 1/w_e
 * for importance weights w (or 1), renormalised by the mean weight when weights
 * are given (ampIntegralMatrix.cc:477-483).  The caller divides by the number
 * of events, as ampIntegralMatrix::element does.
 * --------------------------------------------------------------------------- */
/* amplitudes in host memory, [M][W] complex; streamed through the device */
int rpwa_integral(int device, const double* amps, uint64_t nmbEvents, uint32_t nmbWaves,
                  const double* weights, double* outSums);
/* amplitudes (and weights) already in device memory */
int rpwa_integral_device(int device, const double* d_amps, uint64_t nmbEvents, uint32_t nmbWaves,
                         const double* d_weights, double* outSums);
/* rpwa_integral_device without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results (bitwise equal to rpwa_integral_device on a
 * GPU with at most 256 compute units); the work area is one static device
 * array of 68.5 MiB that each GPU the library runs on holds until the process
 * ends.  At most 128 waves (RPWA_ERR_UNSUPPORTED above).  Calls share that
 * array per GPU: calls that overlap on one GPU (from several host threads)
 * race.  Runs on the device's null stream and returns after the result is in
 * outSums. */
int rpwa_integral_device_noalloc(int device, const double* d_amps, uint64_t nmbEvents, uint32_t nmbWaves,
                                 const double* d_weights, double* outSums);
/* Data moments for the learned initializer (no counterpart in the reference):
 * with a_i(e) = A_i(e) / sqrt(normDiag_i) (normDiag: the normalisation
 * integral's diagonal elements), outMoments [W][W] complex =
 * mean_e a_i(e) conj(a_j(e)), or with perEventNorm mean_e a_i conj(a_j) / |a(e)|^2,
 * which removes any overall per-event scale.  FP64 MFMA, d_amps [N][W] complex128. */
int rpwa_data_moments_device(int device, const double* d_amps, uint64_t nmbEvents, uint32_t nmbWaves,
                             const double* normDiag, int perEventNorm, double* outMoments);

/* ---------------------------------------------------------------------------
 * Model intensity of generated events: modelIntensity::getIntensity
 * (generators/modelIntensity.cc:170-187), with the decay amplitudes of
 * getDecayAmplitudes (:190-211) given as arrays,
 *   intensity[e] = sum_r sum_refl | sum_i T_r,refl,i A_refl,i(e) / waveNorm_refl,i |^2
 * amps: [nmbEvents][W0 + W1] complex128, refl -1 waves first; prodAmps (host):
 * [rank][2][nmbWavesMax] complex, nmbProdAmps = rank*2*nmbWavesMax (rank 1 with
 * the fit result's production amplitudes is the reference's quantity; the flat
 * wave is not included); waveNorm (host, [W0 + W1], may be NULL): the phase-space
 * integrals sqrt(Re I_ii) the amplitudes are divided by.  FP64.
 * --------------------------------------------------------------------------- */
/* amplitudes and result in device memory (d_intensity [nmbEvents] doubles) */
int rpwa_intensity_device(int device, const double* d_amps, uint64_t nmbEvents, const uint32_t nmbWavesRefl[2],
                          const double* prodAmps, uint32_t nmbProdAmps, uint32_t rank, const double* waveNorm,
                          double* d_intensity);
/* amplitudes and result in host memory (copied through the device) */
int rpwa_intensity(int device, const double* amps, uint64_t nmbEvents, const uint32_t nmbWavesRefl[2],
                   const double* prodAmps, uint32_t nmbProdAmps, uint32_t rank, const double* waveNorm,
                   double* intensity);

/* ---------------------------------------------------------------------------
 * Synthetic amplitudes (the generator behind the benchmark workloads).
 * A_i(e) = (N(0,1) + i N(0,1)) * s_i * exp(0.5 N(0,1)_e), drawn from a
 * counter-based hash of (seed, stream, event, slot), so any event of a pool
 * can be regenerated independently.
 * --------------------------------------------------------------------------- */
/* s_i ~ U[0.1, 1] for i < nmbWaves (host computation) */
int rpwa_gen_wave_scales(uint64_t seed, uint32_t nmbWaves, double* scales);
/* complex-normal production amplitudes [rank][2][Wmax] with the rank
 * restrictions of pwaLikelihood (T=0 for i<r, real for i=r) */
int rpwa_gen_prod_amps(uint64_t seed, uint32_t stream, uint32_t rank, const uint32_t nmbWavesRefl[2],
                       double* prodAmps);
/* events [eventOffset, eventOffset+nmbEvents) of pool `stream` into device
 * memory d_out [nmbEvents][nmbWaves] */
int rpwa_gen_amps_device(int device, uint64_t seed, uint32_t stream, uint64_t eventOffset,
                         uint64_t nmbEvents, uint32_t nmbWaves, const double* scales, double* d_out);
/* accept-reject: the first nmbWanted events of pool `stream` accepted with
 * probability sigma_e / max(sigma) under prodAmps ([rank][2][Wmax]), the
 * maximum taken over the pool; the pool grows (x2) until enough are accepted.
 * Writes the accepted amplitudes to d_out [nmbWanted][W0+W1] (device) and,
 * if not NULL, the pool size and the maximum intensity used. */
int rpwa_gen_accept_reject_device(int device, uint64_t seed, uint32_t stream, uint64_t nmbWanted,
                                  const uint32_t nmbWavesRefl[2], uint32_t rank,
                                  const double* prodAmps, const double* scales, double* d_out,
                                  uint64_t* poolSize, double* maxIntensity);

#ifdef __cplusplus
}
#endif
#endif
