/*
 * rootpwa_amd_pwalik.h -- C ABI over the C++ likelihood class rpwa::pwaLikelihood
 * (include/rpwa/pwaLikelihood.h) and rpwa::ampIntegralMatrix.
 *
 * It exposes what ROOTPWA's Python binding of the likelihood exposes
 * (pyInterface/bindings/partialWaveFit/pwaLikelihood_py.cc:243-312: init,
 * addNormIntegral, addAccIntegral, addAmplitude, finishInit, Gradient, FdF,
 * DoEval, DoDerivative, Hessian, CovarianceMatrix, CorrectParamSigns, the
 * accessors and the prior settings), so a foreign-language caller (ctypes,
 * cgo, JNI) reaches the same object a C++ minimiser driver uses.  Arrays are
 * plain pointers with explicit sizes; complex values are interleaved doubles.
 * Every function returns a status code (RPWA_OK = 0, see rootpwa_amd.h);
 * rpwa_last_error() describes the last failure.
 */
#ifndef ROOTPWA_AMD_PWALIK_H
#define ROOTPWA_AMD_PWALIK_H

#include <stdint.h>

#include "rootpwa_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rpwa_intmat rpwa_intmat;   /* rpwa::ampIntegralMatrix */
typedef struct rpwa_pwalik rpwa_pwalik;   /* rpwa::pwaLikelihood<std::complex<T>> */

/* ---- ampIntegralMatrix (decayAmplitude/ampIntegralMatrix.h:67-172) ---- */
int rpwa_intmat_create(rpwa_intmat** m);
int rpwa_intmat_destroy(rpwa_intmat* m);
/* names: nmbWaves NUL-terminated strings */
int rpwa_intmat_set_wave_names(rpwa_intmat* m, uint32_t nmbWaves, const char* const* names);
int rpwa_intmat_nmb_waves(const rpwa_intmat* m, uint32_t* nmbWaves);
/* name of wave i, NUL-terminated, truncated to nameLen */
int rpwa_intmat_wave_name(const rpwa_intmat* m, uint32_t i, char* name, int nameLen);
/* sums [W][W] complex and the number of events they cover (W = rpwa_intmat_nmb_waves) */
int rpwa_intmat_set_sums(rpwa_intmat* m, const double* sums, uint64_t nmbEvents);
int rpwa_intmat_get_sums(const rpwa_intmat* m, double* sums, uint64_t* nmbEvents);
int rpwa_intmat_set_nmb_events(rpwa_intmat* m, uint64_t nmbEvents);
/* element(i, j) = sum_ij / nmbEvents (ampIntegralMatrix.cc:238-252) */
int rpwa_intmat_element(const rpwa_intmat* m, uint32_t i, uint32_t j, double* re, double* im);
/* ampIntegralMatrix::integrate on the GPU: amps [M][W] complex in host
 * (onDevice = 0) or device (onDevice = 1) memory, weights (importance weights,
 * same memory space) or NULL */
int rpwa_intmat_integrate(rpwa_intmat* m, const double* amps, uint64_t nmbEvents, const double* weights,
                          int device, int onDevice);
/* the same with nmbSubAmps incoherent sub-amplitudes per wave
 * (ampIntegralMatrix.cc:432-466): amps [nmbSubAmps][M][W] complex,
 * sum_ij = sum_e weight_e sum_s A_i^s(e) conj(A_j^s(e)), event count M */
int rpwa_intmat_integrate_incoherent(rpwa_intmat* m, const double* amps, uint32_t nmbSubAmps, uint64_t nmbEvents,
                                     const double* weights, int device, int onDevice);
int rpwa_intmat_add(rpwa_intmat* m, const rpwa_intmat* other);   /* operator+= */
int rpwa_intmat_renormalize(rpwa_intmat* m, uint64_t nmbEvents);
int rpwa_intmat_write_ascii(const rpwa_intmat* m, const char* fileName);
int rpwa_intmat_read_ascii(rpwa_intmat* m, const char* fileName);

/* ---- pwaLikelihood (partialWaveFit/pwaLikelihood.h:66-323) ---- */
/* precision: RPWA_PREC_FP64 -> pwaLikelihood<complex<double>>,
 *            RPWA_PREC_MIXED -> pwaLikelihood<complex<float>> (complex64 amplitudes) */
int rpwa_pwalik_create(rpwa_pwalik** L, int precision);
int rpwa_pwalik_clone(const rpwa_pwalik* L, rpwa_pwalik** out);   /* Clone(): shares the GPU context */
int rpwa_pwalik_destroy(rpwa_pwalik* L);
int rpwa_pwalik_set_device(rpwa_pwalik* L, int device);
/* init(waveDescThres, rank, massBinCenter): reflectivities are +1 / -1 */
int rpwa_pwalik_init(rpwa_pwalik* L, uint32_t nmbWaves, const char* const* names, const int* reflectivities,
                     const double* thresholds, uint32_t rank, double massBinCenter);
int rpwa_pwalik_use_normalized_amps(rpwa_pwalik* L, int flag);
int rpwa_pwalik_set_prior(rpwa_pwalik* L, int halfCauchy, double cauchyWidth);
/* normalizedAmpsUsed, priorType (1 = HALF_CAUCHY), cauchyWidth, cudaEnabled
 * (partialWaveFit/pwaLikelihood.h:208-215); any output may be NULL */
int rpwa_pwalik_settings(const rpwa_pwalik* L, int* useNormalized, int* halfCauchy, double* cauchyWidth,
                         int* gpuEnabled);
/* enableCuda (partialWaveFit/pwaLikelihood.h:208): the GPU is the only
 * event-sum backend, so with flag = 0 every evaluation fails with an error */
int rpwa_pwalik_enable_gpu(rpwa_pwalik* L, int flag);
/* setQuiet (partialWaveFit/pwaLikelihood.h:154), for every likelihood */
int rpwa_pwalik_set_quiet(int flag);
int rpwa_pwalik_add_norm_integral(rpwa_pwalik* L, const rpwa_intmat* m);
int rpwa_pwalik_add_acc_integral(rpwa_pwalik* L, rpwa_intmat* m, uint32_t accEventsOverride);
/* setOnTheFlyBinning (partialWaveFit/pwaLikelihood.h:230-231, .cc:1391-1436,
 * called from pyInterface/module/_likelihood.py:52): keep the entries with
 * lower[k] <= values[k][e] < upper[k] for every binning variable k; values is
 * [nmbVars][nmbEntries] over all entries of the event files, concatenated in
 * the order of the amplitude arrays (the variables' names are for messages and
 * must be distinct).  nmbVars = 0 keeps every entry.  After this call
 * add_amplitude takes arrays of nmbEntries amplitudes and keeps the selected
 * ones in entry order; add_amplitudes_device is refused.  Call before the
 * first add_amplitude. */
int rpwa_pwalik_set_on_the_fly_binning(rpwa_pwalik* L, uint32_t nmbVars, const char* const* varNames,
                                       const double* lower, const double* upper, uint64_t nmbEntries,
                                       const double* values);
/* number of entries that the on-the-fly binning selected */
int rpwa_pwalik_nmb_entries_in_bin(const rpwa_pwalik* L, uint64_t* nmbEntries);
int rpwa_pwalik_add_amplitude(rpwa_pwalik* L, const char* waveName, const double* amps, uint64_t nmbEvents);
/* all waves at once from device memory, [N][W] complex128: the refl -1 waves,
 * then the refl +1 waves, each group in init() wave-list order */
int rpwa_pwalik_add_amplitudes_device(rpwa_pwalik* L, const double* d_amps, uint64_t nmbEvents);
int rpwa_pwalik_finish_init(rpwa_pwalik* L);

int rpwa_pwalik_nmb_pars(const rpwa_pwalik* L, uint32_t* nmbPars, uint32_t* nmbParsFixed);
int rpwa_pwalik_nmb_events(const rpwa_pwalik* L, uint64_t* nmbEvents);
int rpwa_pwalik_nmb_waves(const rpwa_pwalik* L, int reflectivity, uint32_t* nmbWaves);
/* parameter i: name (parName(), truncated to nameLen), rank, fixed, realPart */
int rpwa_pwalik_parameter(const rpwa_pwalik* L, uint32_t i, char* name, int nameLen, uint32_t* rank,
                          int* fixed, int* realPart);

int rpwa_pwalik_doeval(const rpwa_pwalik* L, const double* par, double* value);
int rpwa_pwalik_gradient(const rpwa_pwalik* L, const double* par, double* gradient);
int rpwa_pwalik_fdf(const rpwa_pwalik* L, const double* par, double* value, double* gradient);
/* n calls of FdF at par from C++, as a native minimiser makes them; *seconds is
 * the wall time of the n calls, value/gradient the last result.  No counterpart
 * in the reference: it measures the native call path without the overhead of a
 * foreign-function binding. */
int rpwa_pwalik_time_fdf(const rpwa_pwalik* L, const double* par, uint32_t n, double* value, double* gradient,
                         double* seconds);
int rpwa_pwalik_doderivative(const rpwa_pwalik* L, const double* par, uint32_t k, double* value);
/* nmbPars x nmbPars row-major */
int rpwa_pwalik_hessian(const rpwa_pwalik* L, const double* par, double* hessian);
int rpwa_pwalik_covariance(const rpwa_pwalik* L, const double* par, double* covariance);
/* CovarianceMatrix(hessian) (partialWaveFit/pwaLikelihood.h:185, .cc:927-979):
 * inverse of the P x P Hessian reduced to the free parameters, zero rows and
 * columns at fixed parameters */
int rpwa_pwalik_covariance_from_hessian(const rpwa_pwalik* L, const double* hessian, double* covariance);
int rpwa_pwalik_correct_param_signs(const rpwa_pwalik* L, const double* par, double* out);
/* [rank][2][nmbWavesReflMax] complex and the flat amplitude */
int rpwa_pwalik_copy_from_par(const rpwa_pwalik* L, const double* par, double* prodAmps, double* flat);
int rpwa_pwalik_copy_to_par(const rpwa_pwalik* L, const double* prodAmps, double flat, double* par);
/* buildProdAmpArrays (partialWaveFit/pwaLikelihood.h:239-243, .cc:1624-1661):
 * *n production amplitudes (complex, at most nmbPars), their parameter indices
 * as (re, im) pairs with -1 for none, and their names joined by newlines into
 * names[namesLen]; prodAmps, parIndices and names may be NULL */
int rpwa_pwalik_build_prod_amp_arrays(const rpwa_pwalik* L, const double* par, int withFlat, double* prodAmps,
                                      int32_t* parIndices, uint32_t* n, char* names, uint64_t namesLen);
/* print() (funcInfo = 0) or printFuncInfo() (funcInfo = 1) as text
 * (partialWaveFit/pwaLikelihood.h:245-246); *needed = bytes with the
 * terminating zero; text may be NULL to ask for the size */
int rpwa_pwalik_print(const rpwa_pwalik* L, int funcInfo, char* text, uint64_t textLen, uint64_t* needed);
/* integralMatrices(withFlat): (W[+1])^2 complex each, phase-space integrals W[+1] */
int rpwa_pwalik_integral_matrices(const rpwa_pwalik* L, int withFlat, double* normMatrix, double* accMatrix,
                                  double* phaseSpaceIntegral);
/* the event-sum context of a finished likelihood (borrowed; shared with its
 * clones): kernel timing, tile selection (rpwa_lik_* of rootpwa_amd.h) */
int rpwa_pwalik_gpu_context(const rpwa_pwalik* L, rpwa_lik_ctx** ctx);
/* number of calls and accumulated total time [s] of FdF, Gradient, DoEval,
 * DoDerivative, Hessian (functionCallEnum order) */
int rpwa_pwalik_call_info(const rpwa_pwalik* L, uint32_t calls[5], double totalTime[5]);
/* HessianEigenVectors (partialWaveFit/pwaLikelihood.h:184, .cc:864-911) of a
 * P x P Hessian: nFree = P - nmbParsFixed eigenvalues in descending order and
 * the matching eigenvectors as rows [nFree][P] (zero at fixed parameters) */
int rpwa_pwalik_hessian_eigen(const rpwa_pwalik* L, const double* hessian, double* eigenValues,
                              double* eigenVectors);

/* ---- fit drivers (highLevelInterface/pwaNloptFit.cc:84-402, pwaFit.cc:64-433) ----
 * A fit runs the reference's flow on the likelihood: start values from the
 * seed, minimisation (L-BFGS on FdF, or DoEval + Gradient in Minuit's call
 * pattern), CorrectParamSigns with the bitwise DoEval check, Hessian
 * eigenvalues and covariance.  opts == NULL takes the reference's settings:
 * NLopt flow memory 10, maxEvaluations 50000, xtolRel 1e-4, ftolAbs 1e-6;
 * Minuit flow maxEvaluations 40000, edmTolerance 1e-10. */
typedef struct {
	uint32_t memory;          /* L-BFGS correction pairs */
	uint32_t maxEvaluations;
	double xtolRel;           /* 0: off */
	double ftolAbs;           /* 0: off */
	double edmTolerance;      /* Minuit flow: stop when EDM < 0.002 * edmTolerance */
} rpwa_fit_options;

typedef struct {
	double logLikelihood;     /* the minimum of -lnL the minimiser returned (bitwise DoEval of the result) */
	double edm;               /* Minuit flow: estimated distance to the minimum */
	int32_t status;           /* NLopt result codes: >0 success (1 success, 3 ftol, 4 xtol, 5 maxeval,
	                             7 EDM goal, 8 machine precision), <0 failure (-4 roundoff limited) */
	int32_t converged;        /* status > 0 and (if computed) a positive-definite Hessian */
	int32_t hasHessian;       /* the covariance is filled */
	uint32_t nmbEvaluations;  /* FdF calls (NLopt flow) or DoEval calls (Minuit flow) */
	uint32_t nmbIterations;
	uint32_t nmbNonPositiveEigenvalues;
} rpwa_fit_summary;

/* the start values both drivers draw from seed (pwaNloptFit.cc:183-212,
 * pwaFit.cc:235-257): for each parameter Uniform(0.01, sqrt(N)) and a sign
 * flip if a second draw exceeds 0.5, from MT19937 seeded by init_genrand(seed)
 * (the generator of ROOT's TRandom3); fixed parameters are 0.  out: [P] */
int rpwa_fit_start_values(const rpwa_pwalik* L, uint32_t seed, double* out);
/* pwaNloptFit: outPar [P] (after CorrectParamSigns); outCovariance [P][P] or
 * NULL (saveSpace); outEigenValues [P - nmbParsFixed] or NULL (filled when
 * checkHessian or a covariance is requested).  A changed likelihood after the
 * sign flip is an error (RPWA_ERR_INVALID), as the reference throws there. */
int rpwa_fit_nlopt(const rpwa_pwalik* L, uint32_t seed, int checkHessian, const rpwa_fit_options* opts,
                   double* outPar, double* outCovariance, double* outEigenValues, rpwa_fit_summary* out);
/* pwaFit: startPar [P] or NULL (a zero for a free parameter draws a random
 * start value, pwaFit.cc:250-257); otherwise as rpwa_fit_nlopt */
int rpwa_fit_minuit(const rpwa_pwalik* L, uint32_t seed, const double* startPar, int checkHessian,
                    const rpwa_fit_options* opts, double* outPar, double* outCovariance, double* outEigenValues,
                    rpwa_fit_summary* out);

#ifdef __cplusplus
}
#endif
#endif
