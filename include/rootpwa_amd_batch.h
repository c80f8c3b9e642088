/*
 * rootpwa_amd_batch.h -- batched fits: many bins, each with many fit starts,
 * minimised together on one GPU.
 *
 * ROOTPWA fits each mass/t bin separately, and each bin several times from
 * random start values (pyInterface/module/_fit.py:132-165, the attempts loop
 * around pwaFit / pwaNloptFit).  This interface runs all of those fits at once:
 * every bin's likelihood is the one of partialWaveFit/pwaLikelihood.cc
 * (rank 1, one reflectivity sector, normalised amplitudes, flat wave, flat
 * prior), every start runs an L-BFGS minimisation (the pwaNloptFit flow,
 * highLevelInterface/pwaNloptFit.cc:152-244), and one evaluation of all fits is
 * one pass of a persistent FP32-MFMA kernel over every bin's amplitudes.
 * It has no counterpart in the reference, which runs these fits one by one;
 * the per-fit results are those of pwaNloptFit's flow.
 *
 * Precision: amplitudes are stored as complex64 (rounded after the FP64
 * normalisation), the production amplitudes are rounded to FP32 for the
 * contraction (the norm term uses the same rounded values, in FP64), the
 * contraction and the per-tile gradient partials run on FP32 MFMA, and the
 * per-event log terms, the event sums, the norm term T^H I T, the gradient
 * accumulation over tiles and the minimiser state are FP64.
 *
 * Parameter layout per fit (as pwaLikelihood with rank 1 and the waves in one
 * sector): P = 2 W, [Re T_0, Re T_1, Im T_1, ..., Re T_{W-1}, Im T_{W-1}, flat].
 */
#ifndef ROOTPWA_AMD_BATCH_H
#define ROOTPWA_AMD_BATCH_H

#include <stdint.h>

#include "rootpwa_amd.h"
#include "rootpwa_amd_pwalik.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rpwa_batch rpwa_batch;

/* Largest number of waves per bin and of starts per bin. */
#define RPWA_BATCH_MAX_WAVES 96
#define RPWA_BATCH_MAX_STARTS 64

/* nmbWaves <= RPWA_BATCH_MAX_WAVES; nmbStarts <= RPWA_BATCH_MAX_STARTS */
int rpwa_batch_create(rpwa_batch** b, int device, uint32_t nmbWaves, uint32_t nmbStarts);
int rpwa_batch_destroy(rpwa_batch* b);

/* One bin: data amplitudes [nmbEvents][nmbWaves] complex128 in host memory
 * (onDevice = 0) or device memory (onDevice = 1); the normalisation integral
 * sums [W][W] complex over nmbNormEvents MC events and the acceptance integral
 * sums over nmbAccEvents (accSums = NULL: the normalisation integral,
 * acceptance 1).  The amplitudes are normalised with the integral's diagonal
 * and the acceptance matrix with the diagonal's square roots, as
 * pwaLikelihood::addAmplitude / finishInit do (pwaLikelihood.cc:1253-1260,
 * 1311-1334).  binIndex (or NULL) receives the bin's index. */
int rpwa_batch_add_bin(rpwa_batch* b, const double* amps, uint64_t nmbEvents, int onDevice,
                       const double* normSums, uint64_t nmbNormEvents, const double* accSums,
                       uint64_t nmbAccEvents, uint32_t* binIndex);
/* Plans the event tiles and tunes the kernel's grid; no bins can be added afterwards. */
int rpwa_batch_finish(rpwa_batch* b);

int rpwa_batch_nmb_pars(const rpwa_batch* b, uint32_t* nmbPars);
int rpwa_batch_nmb_bins(const rpwa_batch* b, uint32_t* nmbBins);

/* One evaluation of every fit: par [nBins][nStarts][P] -> -lnL [nBins][nStarts]
 * and its gradient [nBins][nStarts][P] (grad may be NULL). */
int rpwa_batch_eval(rpwa_batch* b, const double* par, double* value, double* grad);

/* The start values rpwa_batch_fit uses without startPar: for bin i and start
 * s the drivers' draws (rpwa_fit_start_values) from the seed
 * splitmix64(seed * 2^32 + i * 2^16 + s) truncated to 32 bits.  out: [nBins][nStarts][P] */
int rpwa_batch_start_values(const rpwa_batch* b, uint32_t seed, double* out);

/* Minimises every fit.  startPar [nBins][nStarts][P] or NULL (random starts
 * from seed); opts NULL: L-BFGS memory 10, maxEvaluations 50000,
 * xtolRel 1e-4, ftolAbs 1e-6 (pwaNloptFit.cc:105-107).  Outputs per fit:
 * outPar [P] after the sign convention (CorrectParamSigns), outValue the
 * minimum, outStatus (3 ftol reached, 4 xtol reached, 5 maxeval reached,
 * -4 no further decrease possible along the search direction, -1 non-finite
 * start), outNmbEval (evaluations used).  Any output may be NULL. */
int rpwa_batch_fit(rpwa_batch* b, const double* startPar, uint32_t seed, const rpwa_fit_options* opts,
                   double* outPar, double* outValue, int32_t* outStatus, uint32_t* outNmbEval);

/* Tile selection: wgPerCu workgroups per compute unit for the event kernel
 * (0: autotuned at rpwa_batch_finish, the default). */
int rpwa_batch_set_tile(rpwa_batch* b, uint32_t wgPerCu);
/* info: [0] bins, [1] event tiles (32 events), [2] partial slots, [3] workgroups,
 * [4] workgroups per CU, [5] waves per bin padded, [6] total events,
 * [7] batched evaluations of the last rpwa_batch_fit */
int rpwa_batch_info(const rpwa_batch* b, uint64_t info[8]);
/* average duration [ms] of the event kernel over reps launches at the current
 * evaluation point (after 5 untimed launches), timed with HIP events on the
 * batch's stream */
int rpwa_batch_time_event_kernel(rpwa_batch* b, uint32_t reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
