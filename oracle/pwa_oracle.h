/*
 * pwa_oracle.h -- CPU oracle for the unbinned extended log-likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of ROOTPWA's
 * rpwa::pwaLikelihood<std::complex<double>> event loops, normalisation term,
 * prior, parameter mapping and of ampIntegralMatrix::integrate.  It exists so
 * that tests, the smoke check and the CPU-baseline leg of the benchmark can
 * check the HIP product path against the reference algorithm.  Nothing in the
 * product library links, loads or calls it.
 *
 * Complex numbers are interleaved (re, im) pairs of doubles, exactly the memory
 * layout of std::complex<double>.  All index conventions follow the reference:
 *   refl 0 = negative reflectivity, refl 1 = positive reflectivity
 *     (partialWaveFit/pwaLikelihood.h:279)
 *   production amplitudes: [rank][2][nWavesReflMax]   (pwaLikelihood.cc:1688)
 *   integral matrices:     [2][nWavesReflMax][2][nWavesReflMax]
 *                                                      (pwaLikelihood.cc:1566)
 *   decay amplitudes:      per refl, [nEvents][nWavesRefl[refl]]
 *                                                      (pwaLikelihood.cc:1226)
 */
#ifndef RPWA_ORACLE_H
#define RPWA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_lik orc_lik;

enum { ORC_PRIOR_FLAT = 0, ORC_PRIOR_HALF_CAUCHY = 1 };

/* pwaLikelihood::init -> buildParDataStruct (pwaLikelihood.cc:1489-1556).
 * thresholds: [2][nWavesReflMax] (may be NULL = all zero). */
orc_lik* orc_lik_new(unsigned rank, const unsigned nWavesRefl[2],
                     const double* thresholds, double massBinCenter);
void     orc_lik_free(orc_lik* L);

unsigned orc_lik_nmb_pars(const orc_lik* L);
unsigned orc_lik_nmb_pars_fixed(const orc_lik* L);
unsigned orc_lik_wmax(const orc_lik* L);
/* parameter description: rank, refl, wave index, fixed flag, realPart flag;
 * the flat wave has refl = -1, wave = -1. */
void     orc_lik_par_info(const orc_lik* L, unsigned i, int* rank, int* refl,
                          int* wave, int* fixed, int* realPart);
/* _prodAmpToFuncParMap: out[rank][2][Wmax][2] */
void     orc_lik_par_map(const orc_lik* L, int* out);

void orc_lik_use_normalized_amps(orc_lik* L, int flag);
void orc_lik_set_prior(orc_lik* L, int priorType, double cauchyWidth);

/* addNormIntegral + addAccIntegral + reorderIntegralMatrix
 * (pwaLikelihood.cc:1068-1107, 1562-1574).
 * normSums / accSums: the ampIntegralMatrix _integrals arrays (sums over MC
 * events) over the wave order [refl 0 waves..., refl 1 waves...], [W][W]
 * complex, with their nmbEvents.  element() = sum / nmbEvents
 * (ampIntegralMatrix.cc:238-252); the acceptance matrix is divided by the
 * (possibly overridden) number of generated events, as addAccIntegral does. */
int  orc_lik_add_integrals(orc_lik* L, const double* normSums, uint64_t normNmbEvents,
                           const double* accSums, uint64_t accNmbEvents,
                           uint64_t accEventsOverride);

/* addAmplitude for one wave (pwaLikelihood.cc:1112-1288): normalise by
 * sqrt(Re I^norm_ii), store, auto-fix an all-zero wave.  amps: [nEvents] complex. */
int  orc_lik_add_amplitude(orc_lik* L, unsigned refl, unsigned wave,
                           const double* amps, uint64_t nEvents);

/* finishInit (pwaLikelihood.cc:1293-1388): rescale integral matrices. */
int  orc_lik_finish_init(orc_lik* L);

/* read back the prepared state */
uint64_t orc_lik_nmb_events(const orc_lik* L);
void orc_lik_get_acc_matrix(const orc_lik* L, double* out);   /* [2][Wmax][2][Wmax] */
void orc_lik_get_norm_matrix(const orc_lik* L, double* out);  /* [2][Wmax][2][Wmax] */
void orc_lik_get_decay_amps(const orc_lik* L, unsigned refl, double* out); /* [N][W_refl] */
double orc_lik_tot_acc(const orc_lik* L);

/* parameter mapping (pwaLikelihood.cc:1682-1731) */
void orc_lik_copy_from_par(const orc_lik* L, const double* par, double* prodAmps,
                           double* flat);
void orc_lik_copy_to_par(const orc_lik* L, const double* prodAmps, double flat,
                         double* par);

/* the evaluation entry points */
double orc_lik_doeval  (orc_lik* L, const double* par);                 /* :320-435 */
void   orc_lik_gradient(orc_lik* L, const double* par, double* grad);   /* :477-636 */
void   orc_lik_fdf     (orc_lik* L, const double* par, double* f,
                        double* grad);                                  /* :172-315 */
double orc_lik_doderivative(orc_lik* L, const double* par, unsigned k); /* :440-471 */
void   orc_lik_hessian (orc_lik* L, const double* par, double* H);      /* :642-859 */
void   orc_lik_correct_param_signs(const orc_lik* L, const double* par,
                                   double* out);                        /* :986-1010 */

/* FdF with the event loop split over nThreads contiguous event ranges whose
 * partial Kahan sums are merged in range order (all-host-cores CPU baseline). */
void   orc_lik_fdf_mt(orc_lik* L, const double* par, double* f, double* grad,
                      int nThreads);

/* Event part only: the quantities cuda::likelihoodInterface returns
 * (cuda/likelihoodInterface.cuh:112-122): sum_e -ln sigma_e, and
 * deriv[r][refl][w] = sum_e -(2/sigma_e) S_r,refl(e) conj(A_refl,w(e)),
 * derivFlat = sum_e -(2/sigma_e) f.  prodAmps: [rank][2][Wmax] complex. */
double orc_event_loglik(const orc_lik* L, const double* prodAmps, double flat);
void   orc_event_deriv (const orc_lik* L, const double* prodAmps, double flat,
                        double* deriv, double* derivFlat);

/* ampIntegralMatrix::integrate (decayAmplitude/ampIntegralMatrix.cc:300-488)
 * for amplitudes with one incoherent sub-amplitude.
 * amps: [M][W] complex (event-major); weights: importance weights w_e (the
 * integral uses 1/w_e), or NULL.  out: [W][W] complex sums (the _integrals
 * array, renormalised to the weight integral when weights are given). */
void orc_integral_incoherent(const double* amps, unsigned S, uint64_t M, unsigned W, const double* weights,
                             double* out);
void orc_integral(const double* amps, uint64_t M, unsigned W, const double* weights,
                  double* out);

/* modelIntensity::getIntensity (generators/modelIntensity.cc:170-187) with the
 * decay amplitudes of getDecayAmplitudes (:190-211) given per event:
 * out[e] = sum over reflectivities (-1, then +1), then ranks, of
 * |sum_i T_r,refl,i (A_refl,i(e) / waveNorm_i)|^2 (plain std::complex
 * accumulation, as the reference's `amp +=`).  amps [N][W0+W1] complex (refl -1
 * first), prodAmps [rank][2][Wmax] complex, waveNorm [W0+W1] or NULL (no
 * division).  Rank 1 is the reference's quantity. */
void orc_intensity(const double* amps, uint64_t N, const unsigned nWavesRefl[2], const double* prodAmps,
                   unsigned rank, const double* waveNorm, double* out);

/* the testSumAccumulators scenario (test/utilities/testSumAccumulators.cc:46-60):
 * n additions of x in float, returned as {naive, compensated}. */
void orc_kahan_float_demo(uint64_t n, float x, float* naive, float* compensated);

#ifdef __cplusplus
}
#endif
#endif
