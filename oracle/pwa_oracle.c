/*
 * pwa_oracle.c -- CPU oracle for the unbinned extended log-likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pwa_oracle.h).  Restates, in plain C99, the
 * algorithm of partialWaveFit/pwaLikelihood.cc and of
 * decayAmplitude/ampIntegralMatrix.cc for std::complex<double>, keeping the
 * reference's loop order and its compensated (Kahan) accumulators
 * (utilities/sumAccumulators.hpp:153-184, applied component-wise to complex
 * samples), so that the result is the reference's result on the same inputs.
 *
 * Complex products are written out as the compiler expands
 * std::complex<double>::operator* for finite operands:
 *   (a + ib)(c + id) = (a*c - b*d) + i(a*d + b*c)
 * and norm(z) = re*re + im*im.  Build with -ffp-contract=off so no product is
 * fused into an FMA, as on the reference's x86-64 build.
 *
 * Each function names the reference lines it follows.
 */
#include "pwa_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static inline cplx c_make(double re, double im) { cplx z; z.re = re; z.im = im; return z; }
static inline cplx c_mul(cplx a, cplx b)
{
	return c_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx c_conj(cplx a) { return c_make(a.re, -a.im); }
static inline cplx c_scale(double s, cplx a) { return c_make(s * a.re, s * a.im); }
static inline cplx c_add(cplx a, cplx b) { return c_make(a.re + b.re, a.im + b.im); }
static inline cplx c_sub(cplx a, cplx b) { return c_make(a.re - b.re, a.im - b.im); }
static inline double c_norm(cplx a) { return a.re * a.re + a.im * a.im; }
static inline cplx c_ld(const double* p) { return c_make(p[0], p[1]); }
static inline void c_st(double* p, cplx z) { p[0] = z.re; p[1] = z.im; }

/* compensatedSumAccumulator (utilities/sumAccumulators.hpp:153-184) */
typedef struct { double s, c; } kahan;
static inline void k_add(kahan* k, double x)
{
	const double y = x - k->c;
	const double t = k->s + y;
	k->c = (t - k->s) - y;
	k->s = t;
}
typedef struct { kahan re, im; } kahanc;
static inline void kc_add(kahanc* k, cplx x) { k_add(&k->re, x.re); k_add(&k->im, x.im); }
static inline cplx kc_sum(const kahanc* k) { return c_make(k->re.s, k->im.s); }

struct orc_lik {
	unsigned rank;
	unsigned nW[2];
	unsigned wmax;
	unsigned nPars;
	unsigned nParsFixed;
	uint64_t nEvents;
	int useNormalizedAmps;
	int priorType;
	double cauchyWidth;
	uint64_t numbAccEvents;
	double totAcc;
	double* thr;      /* [2][wmax] */
	int* parMap;      /* [rank][2][wmax][2] */
	int* pRank; int* pRefl; int* pWave; int* pFixed; int* pReal;
	double* normMatrix;  /* [2][wmax][2][wmax] complex */
	double* accMatrix;
	double* psInt;       /* [2][wmax] */
	double* decayAmps[2];/* [N][nW[refl]] complex */
	unsigned char* waveAdded; /* [2][wmax] */
	double* parCache;
	double* derivCache;
};

#define MAT(L, m, r1, w1, r2, w2) \
	((m) + 2 * ((((size_t)(r1) * (L)->wmax + (w1)) * 2 + (r2)) * (L)->wmax + (w2)))
#define PA(L, r, e, w) (2 * ((size_t)(r) * 2 * (L)->wmax + (size_t)(e) * (L)->wmax + (w)))

static size_t prod_idx(const orc_lik* L, unsigned r, unsigned refl, unsigned w)
{
	return ((size_t)r * 2 + refl) * L->wmax + w;
}

/* ------------------------------------------------------------------------- */
/* buildParDataStruct (pwaLikelihood.cc:1489-1556)                            */
orc_lik* orc_lik_new(unsigned rank, const unsigned nWavesRefl[2],
                     const double* thresholds, double massBinCenter)
{
	if (nWavesRefl[0] + nWavesRefl[1] == 0 || rank == 0)
		return NULL;
	orc_lik* L = (orc_lik*)calloc(1, sizeof(orc_lik));
	L->rank = rank;
	L->nW[0] = nWavesRefl[0];
	L->nW[1] = nWavesRefl[1];
	L->wmax = nWavesRefl[0] > nWavesRefl[1] ? nWavesRefl[0] : nWavesRefl[1];
	L->useNormalizedAmps = 1;   /* pwaLikelihood.cc:147 */
	L->priorType = ORC_PRIOR_FLAT;
	L->cauchyWidth = 0.5;       /* pwaLikelihood.cc:149 */
	L->thr = (double*)calloc((size_t)2 * L->wmax, sizeof(double));
	if (thresholds)
		memcpy(L->thr, thresholds, sizeof(double) * 2 * L->wmax);

	unsigned nPars = 0;
	for (unsigned iRank = 0; iRank < rank; ++iRank) {
		const int nPos = (int)nWavesRefl[1] - (int)iRank;
		int nPosC = nPos - 1;
		if (nPosC < 0) nPosC = 0;
		const int nNeg = (int)nWavesRefl[0] - (int)iRank;
		int nNegC = nNeg - 1;
		if (nNegC < 0) nNegC = 0;
		nPars += 2 * (unsigned)(nPosC + nNegC);
		if (nPos > 0) ++nPars;
		if (nNeg > 0) ++nPars;
	}
	nPars += 1;
	L->nPars = nPars;
	L->nParsFixed = 0;
	L->pRank = (int*)calloc(nPars, sizeof(int));
	L->pRefl = (int*)calloc(nPars, sizeof(int));
	L->pWave = (int*)calloc(nPars, sizeof(int));
	L->pFixed = (int*)calloc(nPars, sizeof(int));
	L->pReal = (int*)calloc(nPars, sizeof(int));
	L->parCache = (double*)calloc(nPars, sizeof(double));
	L->derivCache = (double*)calloc(nPars, sizeof(double));
	L->parMap = (int*)malloc(sizeof(int) * 2 * (size_t)rank * 2 * L->wmax);
	for (size_t i = 0; i < (size_t)2 * rank * 2 * L->wmax; ++i)
		L->parMap[i] = -1;
	unsigned parIndex = 0;
	for (unsigned iRank = 0; iRank < rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const double t = L->thr[iRefl * L->wmax + iWave];
				const int fixed = (t != 0 && t >= massBinCenter);
				int* m = L->parMap + 2 * prod_idx(L, iRank, iRefl, iWave);
				if (iWave < iRank) {
					m[0] = -1; m[1] = -1;
				} else if (iWave == iRank) {
					L->pRank[parIndex] = (int)iRank; L->pRefl[parIndex] = (int)iRefl;
					L->pWave[parIndex] = (int)iWave; L->pFixed[parIndex] = fixed;
					L->pReal[parIndex] = 1;
					if (fixed) ++L->nParsFixed;
					m[0] = (int)parIndex; m[1] = -1;
					++parIndex;
				} else {
					for (int k = 0; k < 2; ++k) {
						L->pRank[parIndex + k] = (int)iRank; L->pRefl[parIndex + k] = (int)iRefl;
						L->pWave[parIndex + k] = (int)iWave; L->pFixed[parIndex + k] = fixed;
						L->pReal[parIndex + k] = (k == 0);
					}
					if (fixed) L->nParsFixed += 2;
					m[0] = (int)parIndex; m[1] = (int)parIndex + 1;
					parIndex += 2;
				}
			}
	/* flat wave */
	L->pRank[parIndex] = 0; L->pRefl[parIndex] = -1; L->pWave[parIndex] = -1;
	L->pFixed[parIndex] = 0; L->pReal[parIndex] = 1;

	const size_t matN = (size_t)2 * 2 * L->wmax * 2 * L->wmax;
	L->normMatrix = (double*)calloc(matN, sizeof(double));
	L->accMatrix = (double*)calloc(matN, sizeof(double));
	L->psInt = (double*)calloc((size_t)2 * L->wmax, sizeof(double));
	L->waveAdded = (unsigned char*)calloc((size_t)2 * L->wmax, 1);
	return L;
}

void orc_lik_free(orc_lik* L)
{
	if (!L) return;
	free(L->thr); free(L->parMap);
	free(L->pRank); free(L->pRefl); free(L->pWave); free(L->pFixed); free(L->pReal);
	free(L->normMatrix); free(L->accMatrix); free(L->psInt); free(L->waveAdded);
	free(L->decayAmps[0]); free(L->decayAmps[1]);
	free(L->parCache); free(L->derivCache);
	free(L);
}

unsigned orc_lik_nmb_pars(const orc_lik* L) { return L->nPars; }
unsigned orc_lik_nmb_pars_fixed(const orc_lik* L) { return L->nParsFixed; }
unsigned orc_lik_wmax(const orc_lik* L) { return L->wmax; }
uint64_t orc_lik_nmb_events(const orc_lik* L) { return L->nEvents; }
double orc_lik_tot_acc(const orc_lik* L) { return L->totAcc; }

void orc_lik_par_info(const orc_lik* L, unsigned i, int* rank, int* refl, int* wave,
                      int* fixed, int* realPart)
{
	*rank = L->pRank[i]; *refl = L->pRefl[i]; *wave = L->pWave[i];
	*fixed = L->pFixed[i]; *realPart = L->pReal[i];
}

void orc_lik_par_map(const orc_lik* L, int* out)
{
	memcpy(out, L->parMap, sizeof(int) * 2 * (size_t)L->rank * 2 * L->wmax);
}

void orc_lik_use_normalized_amps(orc_lik* L, int flag) { L->useNormalizedAmps = flag; }
void orc_lik_set_prior(orc_lik* L, int priorType, double cauchyWidth)
{
	L->priorType = priorType;
	L->cauchyWidth = cauchyWidth;
}

/* ------------------------------------------------------------------------- */
/* addNormIntegral / addAccIntegral / reorderIntegralMatrix                  */
/* (pwaLikelihood.cc:1068-1107, 1562-1574; element(): ampIntegralMatrix.cc:252) */
static void reorder(orc_lik* L, const double* sums, double nmbEvents, double* out)
{
	const unsigned W = L->nW[0] + L->nW[1];
	for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
		for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave)
			for (unsigned jWave = 0; jWave < L->nW[iRefl]; ++jWave) {
				const unsigned gi = (iRefl == 0 ? 0 : L->nW[0]) + iWave;
				const unsigned gj = (iRefl == 0 ? 0 : L->nW[0]) + jWave;
				const double* s = sums + 2 * ((size_t)gi * W + gj);
				double* o = MAT(L, out, iRefl, iWave, iRefl, jWave);
				o[0] = s[0] / nmbEvents;
				o[1] = s[1] / nmbEvents;
			}
}

int orc_lik_add_integrals(orc_lik* L, const double* normSums, uint64_t normNmbEvents,
                          const double* accSums, uint64_t accNmbEvents,
                          uint64_t accEventsOverride)
{
	L->numbAccEvents = normNmbEvents;
	reorder(L, normSums, (double)normNmbEvents, L->normMatrix);
	L->numbAccEvents = accEventsOverride == 0 ? L->numbAccEvents : accEventsOverride;
	L->totAcc = (double)accNmbEvents / (double)L->numbAccEvents;
	/* accMatrix.setNmbEvents(_numbAccEvents) before reordering */
	reorder(L, accSums, (double)L->numbAccEvents, L->accMatrix);
	return 0;
}

/* ------------------------------------------------------------------------- */
/* addAmplitude (pwaLikelihood.cc:1112-1288), one wave, no on-the-fly binning */
int orc_lik_add_amplitude(orc_lik* L, unsigned refl, unsigned wave, const double* amps,
                          uint64_t nEvents)
{
	if (refl > 1 || wave >= L->nW[refl] || nEvents == 0)
		return -1;
	if (L->nEvents == 0) {
		L->nEvents = nEvents;
		for (int r = 0; r < 2; ++r) {
			free(L->decayAmps[r]);
			L->decayAmps[r] = (double*)calloc((size_t)2 * nEvents * (L->nW[r] ? L->nW[r] : 1),
			                                  sizeof(double));
		}
	}
	if (nEvents != L->nEvents)
		return -2;
	const double* nI = MAT(L, L->normMatrix, refl, wave, refl, wave);
	const double* aI = MAT(L, L->accMatrix, refl, wave, refl, wave);
	int fixPar = (aI[0] == 0. && aI[1] == 0.);
	const int normZero = (nI[0] == 0. && nI[1] == 0.);
	for (uint64_t e = 0; e < nEvents; ++e) {
		cplx a = c_ld(amps + 2 * e);
		if (L->useNormalizedAmps) {
			if (normZero && (a.re != 0. || a.im != 0.))
				return -3;
			if (!normZero) {
				const double s = sqrt(nI[0]);
				a.re = a.re / s;
				a.im = a.im / s;
			}
		}
		if (a.re != 0. || a.im != 0.)
			fixPar = 0;
		c_st(L->decayAmps[refl] + 2 * ((size_t)e * L->nW[refl] + wave), a);
	}
	if (fixPar) {
		/* fixes every parameter of this wave; _nmbParsFixed is left unchanged,
		 * as in the reference (pwaLikelihood.cc:1268-1282) */
		for (unsigned p = 0; p + 1 < L->nPars; ++p)
			if (L->pRefl[p] == (int)refl && L->pWave[p] == (int)wave)
				L->pFixed[p] = 1;
	}
	L->waveAdded[refl * L->wmax + wave] = 1;
	return 0;
}

/* finishInit (pwaLikelihood.cc:1293-1353) */
int orc_lik_finish_init(orc_lik* L)
{
	for (unsigned r = 0; r < 2; ++r)
		for (unsigned w = 0; w < L->nW[r]; ++w)
			if (!L->waveAdded[r * L->wmax + w])
				return -1;
	for (unsigned r = 0; r < 2; ++r)
		for (unsigned w = 0; w < L->nW[r]; ++w)
			L->psInt[r * L->wmax + w] = sqrt(MAT(L, L->normMatrix, r, w, r, w)[0]);
	if (L->useNormalizedAmps) {
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				double norm_i = sqrt(MAT(L, L->normMatrix, iRefl, iWave, iRefl, iWave)[0]);
				if (norm_i == 0) norm_i = 1;
				for (unsigned jRefl = 0; jRefl < 2; ++jRefl)
					for (unsigned jWave = 0; jWave < L->nW[jRefl]; ++jWave) {
						double norm_j = sqrt(MAT(L, L->normMatrix, jRefl, jWave, jRefl, jWave)[0]);
						if (norm_j == 0) norm_j = 1;
						const double d = norm_i * norm_j;
						if (iRefl != jRefl || iWave != jWave) {
							double* n = MAT(L, L->normMatrix, iRefl, iWave, jRefl, jWave);
							n[0] = n[0] / d; n[1] = n[1] / d;
						}
						double* a = MAT(L, L->accMatrix, iRefl, iWave, jRefl, jWave);
						a[0] = a[0] / d; a[1] = a[1] / d;
					}
			}
		for (unsigned r = 0; r < 2; ++r)
			for (unsigned w = 0; w < L->nW[r]; ++w) {
				double* n = MAT(L, L->normMatrix, r, w, r, w);
				if (n[0] != 0. || n[1] != 0.) { n[0] = 1.; n[1] = 0.; }
			}
	}
	return 0;
}

void orc_lik_get_acc_matrix(const orc_lik* L, double* out)
{
	memcpy(out, L->accMatrix, sizeof(double) * 2 * 2 * L->wmax * 2 * L->wmax);
}
void orc_lik_get_norm_matrix(const orc_lik* L, double* out)
{
	memcpy(out, L->normMatrix, sizeof(double) * 2 * 2 * L->wmax * 2 * L->wmax);
}
void orc_lik_get_decay_amps(const orc_lik* L, unsigned refl, double* out)
{
	memcpy(out, L->decayAmps[refl], sizeof(double) * 2 * L->nEvents * L->nW[refl]);
}

/* ------------------------------------------------------------------------- */
/* copyFromParArray / copyToParArray (pwaLikelihood.cc:1682-1731)            */
void orc_lik_copy_from_par(const orc_lik* L, const double* par, double* prodAmps, double* flat)
{
	memset(prodAmps, 0, sizeof(double) * 2 * (size_t)L->rank * 2 * L->wmax);
	unsigned parIndex = 0;
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				double re, im;
				if (iWave < iRank) {
					re = im = 0;
				} else if (iWave == iRank) {
					re = par[parIndex]; im = 0; ++parIndex;
				} else {
					re = par[parIndex]; im = par[parIndex + 1]; parIndex += 2;
				}
				double* p = prodAmps + 2 * prod_idx(L, iRank, iRefl, iWave);
				p[0] = re; p[1] = im;
			}
	*flat = par[parIndex];
}

void orc_lik_copy_to_par(const orc_lik* L, const double* prodAmps, double flat, double* par)
{
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const int* m = L->parMap + 2 * prod_idx(L, iRank, iRefl, iWave);
				const double* p = prodAmps + 2 * prod_idx(L, iRank, iRefl, iWave);
				if (m[0] >= 0) par[m[0]] = p[0];
				if (m[1] >= 0) par[m[1]] = p[1];
			}
	par[L->nPars - 1] = flat;
}

/* ------------------------------------------------------------------------- */
/* event loop shared by DoEval (:361-375), Gradient (:547-577), FdF (:213-244) */
typedef struct {
	kahan lnl;       /* logLikelihoodAcc */
	kahan dflat;     /* derivativeFlatAcc */
	kahanc* dacc;    /* derivativesAcc [rank][2][wmax] */
} evacc;

static void event_loop(const orc_lik* L, const cplx* T, double flat, int wantLnl,
                       int wantDeriv, int allocPerEvent, uint64_t evBegin, uint64_t evEnd,
                       evacc* acc)
{
	const unsigned R = L->rank, Wm = L->wmax;
	const double flat2 = flat * flat;
	const size_t nd = (size_t)R * 2 * Wm;
	cplx* derivStatic = wantDeriv && !allocPerEvent ? (cplx*)calloc(nd, sizeof(cplx)) : NULL;
	for (uint64_t iEvt = evBegin; iEvt < evEnd; ++iEvt) {
		kahan likelihoodAcc = {0, 0};
		cplx* derivative = NULL;
		if (wantDeriv) {
			/* prodAmpsArrayType derivative(derivShape): a zeroed heap array per event */
			derivative = allocPerEvent ? (cplx*)calloc(nd, sizeof(cplx)) : derivStatic;
		}
		for (unsigned iRank = 0; iRank < R; ++iRank) {
			for (unsigned iRefl = 0; iRefl < 2; ++iRefl) {
				kahanc ampProdAcc = {{0, 0}, {0, 0}};
				const double* A = L->decayAmps[iRefl] + 2 * (size_t)iEvt * L->nW[iRefl];
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave)
					kc_add(&ampProdAcc, c_mul(T[prod_idx(L, iRank, iRefl, iWave)], c_ld(A + 2 * iWave)));
				const cplx ampProdSum = kc_sum(&ampProdAcc);
				k_add(&likelihoodAcc, c_norm(ampProdSum));
				if (wantDeriv)
					for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave)
						derivative[prod_idx(L, iRank, iRefl, iWave)] = ampProdSum;
			}
			if (wantDeriv)
				for (unsigned iRefl = 0; iRefl < 2; ++iRefl) {
					const double* A = L->decayAmps[iRefl] + 2 * (size_t)iEvt * L->nW[iRefl];
					for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
						cplx* d = &derivative[prod_idx(L, iRank, iRefl, iWave)];
						*d = c_mul(*d, c_conj(c_ld(A + 2 * iWave)));
					}
				}
		}
		k_add(&likelihoodAcc, flat2);
		if (wantLnl)
			k_add(&acc->lnl, -log(likelihoodAcc.s));
		if (wantDeriv) {
			const double factor = 2. / likelihoodAcc.s;
			for (unsigned iRank = 0; iRank < R; ++iRank)
				for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
					for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
						const size_t k = prod_idx(L, iRank, iRefl, iWave);
						kc_add(&acc->dacc[k], c_scale(-factor, derivative[k]));
					}
			k_add(&acc->dflat, -factor * flat);
			if (allocPerEvent)
				free(derivative);
		}
	}
	free(derivStatic);
}

static void load_prod_amps(const orc_lik* L, const double* par, cplx* T, double* flat)
{
	orc_lik_copy_from_par(L, par, (double*)T, flat);
}

/* norm term value (DoEval :384-399) and derivative (FdF :263-279) */
static double norm_term(const orc_lik* L, const cplx* T, double flat, cplx* deriv,
                        double* derivFlat)
{
	kahan normFactorAcc = {0, 0};
	const double nmbEvt = L->useNormalizedAmps ? 1 : (double)L->nEvents;
	const double twiceNmbEvt = 2 * nmbEvt;
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				kahanc normFactorDerivAcc = {{0, 0}, {0, 0}};
				for (unsigned jWave = 0; jWave < L->nW[iRefl]; ++jWave) {
					const cplx I = c_ld(MAT(L, L->accMatrix, iRefl, iWave, iRefl, jWave));
					const cplx Ti = T[prod_idx(L, iRank, iRefl, iWave)];
					const cplx Tj = T[prod_idx(L, iRank, iRefl, jWave)];
					k_add(&normFactorAcc, c_mul(c_mul(Ti, c_conj(Tj)), I).re);
					if (deriv)
						kc_add(&normFactorDerivAcc, c_mul(Tj, c_conj(I)));
				}
				if (deriv) {
					cplx* d = &deriv[prod_idx(L, iRank, iRefl, iWave)];
					*d = c_add(*d, c_scale(twiceNmbEvt, kc_sum(&normFactorDerivAcc)));
				}
			}
	k_add(&normFactorAcc, flat * flat * L->totAcc);
	if (derivFlat)
		*derivFlat += flat * twiceNmbEvt * L->totAcc;
	return normFactorAcc.s;
}

/* anonymous-namespace helpers of pwaLikelihood.cc:78-102 */
static double cauchyFunction(double x, double gamma)
{
	return 1. / (1. + ((x * x) / (gamma * gamma)));
}
static double cauchyFunctionDerivative(double x, double gamma)
{
	return (-2. * x * gamma * gamma) / ((gamma * gamma + x * x) * (gamma * gamma + x * x));
}
static double cauchyFunctionSecondDerivative(double x, double gamma)
{
	const double x2 = x * x;
	const double gamma2 = gamma * gamma;
	return gamma2 / ((gamma2 + x2) * (gamma2 + x2)) * (-2. + 8. * x2 / (gamma2 + x2));
}

/* prior value and derivative (FdF :281-297, Gradient :608-625) */
static double prior_term(const orc_lik* L, const cplx* T, cplx* deriv)
{
	double priorValue = 0.;
	if (L->priorType != ORC_PRIOR_HALF_CAUCHY)
		return priorValue;
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const cplx t = T[prod_idx(L, iRank, iRefl, iWave)];
				const double r = hypot(t.re, t.im);
				const double cv = cauchyFunction(r, L->cauchyWidth);
				priorValue -= log(cv);
				if (deriv) {
					const double factor = (1. / (r * cv)) * cauchyFunctionDerivative(r, L->cauchyWidth);
					cplx* d = &deriv[prod_idx(L, iRank, iRefl, iWave)];
					*d = c_sub(*d, c_make(factor * t.re, factor * t.im));
				}
			}
	return priorValue;
}

static void deriv_to_par(const orc_lik* L, const cplx* deriv, double derivFlat, double* grad)
{
	orc_lik_copy_to_par(L, (const double*)deriv, derivFlat, grad);
}

static cplx* alloc_prod(const orc_lik* L)
{
	return (cplx*)calloc((size_t)L->rank * 2 * L->wmax, sizeof(cplx));
}

/* FdF (pwaLikelihood.cc:172-315) */
static void fdf_impl(orc_lik* L, const double* par, double* f, double* grad, int nThreads);

void orc_lik_fdf(orc_lik* L, const double* par, double* f, double* grad)
{
	fdf_impl(L, par, f, grad, 1);
}

typedef struct {
	const orc_lik* L; const cplx* T; double flat; uint64_t b, e; evacc acc;
} mt_job;

static void* mt_worker(void* arg)
{
	mt_job* j = (mt_job*)arg;
	event_loop(j->L, j->T, j->flat, 1, 1, 1, j->b, j->e, &j->acc);
	return NULL;
}

static void fdf_impl(orc_lik* L, const double* par, double* f, double* grad, int nThreads)
{
	for (unsigned i = 0; i < L->nPars; ++i)
		L->parCache[i] = par[i];
	const size_t nd = (size_t)L->rank * 2 * L->wmax;
	cplx* T = alloc_prod(L);
	double flat;
	load_prod_amps(L, par, T, &flat);
	cplx* derivatives = alloc_prod(L);
	double derivativeFlat = 0;
	evacc acc;
	memset(&acc, 0, sizeof(acc));
	acc.dacc = (kahanc*)calloc(nd, sizeof(kahanc));
	if (nThreads <= 1) {
		event_loop(L, T, flat, 1, 1, 1, 0, L->nEvents, &acc);
	} else {
		mt_job* jobs = (mt_job*)calloc((size_t)nThreads, sizeof(mt_job));
		pthread_t* th = (pthread_t*)calloc((size_t)nThreads, sizeof(pthread_t));
		for (int t = 0; t < nThreads; ++t) {
			jobs[t].L = L; jobs[t].T = T; jobs[t].flat = flat;
			jobs[t].b = L->nEvents * (uint64_t)t / (uint64_t)nThreads;
			jobs[t].e = L->nEvents * (uint64_t)(t + 1) / (uint64_t)nThreads;
			jobs[t].acc.dacc = (kahanc*)calloc(nd, sizeof(kahanc));
			pthread_create(&th[t], NULL, mt_worker, &jobs[t]);
		}
		for (int t = 0; t < nThreads; ++t)
			pthread_join(th[t], NULL);
		/* merge the range partials in range order; a Kahan partial (s, c) stands
		 * for s - c, so both the sum and its compensation are added */
		for (int t = 0; t < nThreads; ++t) {
			k_add(&acc.lnl, jobs[t].acc.lnl.s);
			k_add(&acc.lnl, -jobs[t].acc.lnl.c);
			k_add(&acc.dflat, jobs[t].acc.dflat.s);
			k_add(&acc.dflat, -jobs[t].acc.dflat.c);
			for (size_t k = 0; k < nd; ++k) {
				const kahanc* p = &jobs[t].acc.dacc[k];
				kc_add(&acc.dacc[k], kc_sum(p));
				kc_add(&acc.dacc[k], c_make(-p->re.c, -p->im.c));
			}
			free(jobs[t].acc.dacc);
		}
		free(jobs);
		free(th);
	}
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const size_t k = prod_idx(L, iRank, iRefl, iWave);
				derivatives[k] = kc_sum(&acc.dacc[k]);
			}
	derivativeFlat = acc.dflat.s;
	const double nmbEvt = L->useNormalizedAmps ? 1 : (double)L->nEvents;
	const double normFactor = norm_term(L, T, flat, derivatives, &derivativeFlat);
	const double priorValue = prior_term(L, T, derivatives);
	deriv_to_par(L, derivatives, derivativeFlat, grad);
	deriv_to_par(L, derivatives, derivativeFlat, L->derivCache);
	*f = acc.lnl.s + nmbEvt * normFactor + priorValue;
	free(acc.dacc);
	free(derivatives);
	free(T);
}

void orc_lik_fdf_mt(orc_lik* L, const double* par, double* f, double* grad, int nThreads)
{
	fdf_impl(L, par, f, grad, nThreads);
}

/* DoEval (pwaLikelihood.cc:320-435, USE_FDF off) */
double orc_lik_doeval(orc_lik* L, const double* par)
{
	cplx* T = alloc_prod(L);
	double flat;
	load_prod_amps(L, par, T, &flat);
	evacc acc;
	memset(&acc, 0, sizeof(acc));
	event_loop(L, T, flat, 1, 0, 0, 0, L->nEvents, &acc);
	const double logLikelihood = acc.lnl.s;
	const double nmbEvt = L->useNormalizedAmps ? 1 : (double)L->nEvents;
	const double normFactor = norm_term(L, T, flat, NULL, NULL);
	const double priorValue = prior_term(L, T, NULL);
	free(T);
	return logLikelihood + nmbEvt * normFactor + priorValue;
}

/* Gradient (pwaLikelihood.cc:477-636, USE_FDF off) */
void orc_lik_gradient(orc_lik* L, const double* par, double* grad)
{
	for (unsigned i = 0; i < L->nPars; ++i)
		L->parCache[i] = par[i];
	const size_t nd = (size_t)L->rank * 2 * L->wmax;
	cplx* T = alloc_prod(L);
	double flat;
	load_prod_amps(L, par, T, &flat);
	cplx* derivatives = alloc_prod(L);
	evacc acc;
	memset(&acc, 0, sizeof(acc));
	acc.dacc = (kahanc*)calloc(nd, sizeof(kahanc));
	event_loop(L, T, flat, 0, 1, 1, 0, L->nEvents, &acc);
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const size_t k = prod_idx(L, iRank, iRefl, iWave);
				derivatives[k] = kc_sum(&acc.dacc[k]);
			}
	double derivativeFlat = acc.dflat.s;
	norm_term(L, T, flat, derivatives, &derivativeFlat);
	prior_term(L, T, derivatives);
	deriv_to_par(L, derivatives, derivativeFlat, grad);
	deriv_to_par(L, derivatives, derivativeFlat, L->derivCache);
	free(acc.dacc);
	free(derivatives);
	free(T);
}

/* DoDerivative (pwaLikelihood.cc:440-471) */
double orc_lik_doderivative(orc_lik* L, const double* par, unsigned k)
{
	int samePar = 1;
	for (unsigned i = 0; i < L->nPars; ++i)
		if (L->parCache[i] != par[i]) { samePar = 0; break; }
	if (samePar)
		return L->derivCache[k];
	double f;
	double* g = (double*)calloc(L->nPars, sizeof(double));
	orc_lik_fdf(L, par, &f, g);
	const double r = g[k];
	free(g);
	return r;
}

/* event part only, as cuda::likelihoodInterface returns it */
double orc_event_loglik(const orc_lik* L, const double* prodAmps, double flat)
{
	evacc acc;
	memset(&acc, 0, sizeof(acc));
	event_loop(L, (const cplx*)prodAmps, flat, 1, 0, 0, 0, L->nEvents, &acc);
	return acc.lnl.s;
}

void orc_event_deriv(const orc_lik* L, const double* prodAmps, double flat, double* deriv,
                     double* derivFlat)
{
	const size_t nd = (size_t)L->rank * 2 * L->wmax;
	evacc acc;
	memset(&acc, 0, sizeof(acc));
	acc.dacc = (kahanc*)calloc(nd, sizeof(kahanc));
	event_loop(L, (const cplx*)prodAmps, flat, 0, 1, 0, 0, L->nEvents, &acc);
	memset(deriv, 0, sizeof(double) * 2 * nd);
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const size_t k = prod_idx(L, iRank, iRefl, iWave);
				c_st(deriv + 2 * k, kc_sum(&acc.dacc[k]));
			}
	*derivFlat = acc.dflat.s;
	free(acc.dacc);
}

/* ------------------------------------------------------------------------- */
/* Hessian (pwaLikelihood.cc:642-859)                                          */
void orc_lik_hessian(orc_lik* L, const double* par, double* H)
{
	const unsigned R = L->rank, Wm = L->wmax, P = L->nPars;
	const size_t nd = (size_t)R * 2 * Wm;
	cplx* T = alloc_prod(L);
	double flat;
	load_prod_amps(L, par, T, &flat);
	const double flat2 = flat * flat;
	/* hessian[iRank][iRefl][iWave][jRank][jRefl][jWave][3] */
	kahan* hacc = (kahan*)calloc(nd * nd * 3, sizeof(kahan));
	kahanc* flatTermsAcc = (kahanc*)calloc(nd, sizeof(kahanc));
	kahan hessianFlatAcc = {0, 0};
	cplx* derivative = (cplx*)calloc(nd, sizeof(cplx));
#define HIX(a, b, c) (((a) * nd + (b)) * 3 + (c))
	for (uint64_t iEvt = 0; iEvt < L->nEvents; ++iEvt) {
		kahan likelihoodAcc = {0, 0};
		memset(derivative, 0, nd * sizeof(cplx));
		for (unsigned iRank = 0; iRank < R; ++iRank) {
			for (unsigned iRefl = 0; iRefl < 2; ++iRefl) {
				kahanc ampProdAcc = {{0, 0}, {0, 0}};
				const double* A = L->decayAmps[iRefl] + 2 * (size_t)iEvt * L->nW[iRefl];
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave)
					kc_add(&ampProdAcc, c_mul(T[prod_idx(L, iRank, iRefl, iWave)], c_ld(A + 2 * iWave)));
				const cplx ampProdSum = kc_sum(&ampProdAcc);
				k_add(&likelihoodAcc, c_norm(ampProdSum));
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave)
					derivative[prod_idx(L, iRank, iRefl, iWave)] = ampProdSum;
			}
			for (unsigned iRefl = 0; iRefl < 2; ++iRefl) {
				const double* A = L->decayAmps[iRefl] + 2 * (size_t)iEvt * L->nW[iRefl];
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
					cplx* d = &derivative[prod_idx(L, iRank, iRefl, iWave)];
					*d = c_mul(*d, c_conj(c_ld(A + 2 * iWave)));
				}
			}
		}
		k_add(&likelihoodAcc, flat2);
		const double factor = 2. / likelihoodAcc.s;
		const double factor2 = factor * factor;
		for (unsigned iRank = 0; iRank < R; ++iRank)
			for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
					const size_t a = prod_idx(L, iRank, iRefl, iWave);
					for (unsigned jRank = 0; jRank < R; ++jRank)
						for (unsigned jRefl = 0; jRefl < 2; ++jRefl)
							for (unsigned jWave = 0; jWave < L->nW[jRefl]; ++jWave) {
								const size_t b = prod_idx(L, jRank, jRefl, jWave);
								k_add(&hacc[HIX(a, b, 0)], factor2 * derivative[b].re * derivative[a].re);
								k_add(&hacc[HIX(a, b, 1)], factor2 * derivative[b].im * derivative[a].re);
								k_add(&hacc[HIX(a, b, 2)], factor2 * derivative[b].im * derivative[a].im);
								if (iRank == jRank && iRefl == jRefl) {
									const double* Aj = L->decayAmps[jRefl] + 2 * ((size_t)iEvt * L->nW[jRefl] + jWave);
									const double* Ai = L->decayAmps[iRefl] + 2 * ((size_t)iEvt * L->nW[iRefl] + iWave);
									const cplx uPrime = c_mul(c_conj(c_ld(Aj)), c_ld(Ai));
									k_add(&hacc[HIX(a, b, 0)], -factor * uPrime.re);
									k_add(&hacc[HIX(a, b, 1)], -factor * uPrime.im);
									k_add(&hacc[HIX(a, b, 2)], -factor * uPrime.re);
								}
							}
					kc_add(&flatTermsAcc[a], c_scale(factor2 * flat, derivative[a]));
				}
		k_add(&hessianFlatAcc, factor2 * flat2 - factor);
	}
	double* hes = (double*)calloc(nd * nd * 3, sizeof(double));
	cplx* flatTerms = (cplx*)calloc(nd, sizeof(cplx));
	for (size_t i = 0; i < nd * nd * 3; ++i)
		hes[i] = hacc[i].s;
	for (size_t a = 0; a < nd; ++a)
		flatTerms[a] = kc_sum(&flatTermsAcc[a]);
	double hessianFlat = hessianFlatAcc.s;
	const double nmbEvt = L->useNormalizedAmps ? 1 : (double)L->nEvents;
	const double twiceNmbEvt = 2 * nmbEvt;
	for (unsigned iRank = 0; iRank < R; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave)
				for (unsigned jWave = 0; jWave < L->nW[iRefl]; ++jWave) {
					const size_t a = prod_idx(L, iRank, iRefl, iWave);
					const size_t b = prod_idx(L, iRank, iRefl, jWave);
					const double* I = MAT(L, L->accMatrix, iRefl, iWave, iRefl, jWave);
					hes[HIX(a, b, 0)] += I[0] * twiceNmbEvt;
					hes[HIX(a, b, 1)] += I[1] * twiceNmbEvt;
					hes[HIX(a, b, 2)] += I[0] * twiceNmbEvt;
				}
	hessianFlat += twiceNmbEvt * L->totAcc;
	if (L->priorType == ORC_PRIOR_HALF_CAUCHY) {
		for (unsigned iRank = 0; iRank < R; ++iRank)
			for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
					const size_t a = prod_idx(L, iRank, iRefl, iWave);
					const double r = hypot(T[a].re, T[a].im);
					const double cv = cauchyFunction(r, L->cauchyWidth);
					const double cd = cauchyFunctionDerivative(r, L->cauchyWidth);
					const double cdd = cauchyFunctionSecondDerivative(r, L->cauchyWidth);
					const double a1 = cdd / (cv * r * r);
					const double a2 = cd * cd / (cv * cv * r * r);
					const double a3 = cd / (cv * r);
					const double a4 = cd / (cv * r * r * r);
					const double a124 = a1 - a2 - a4;
					const double re = T[a].re, im = T[a].im;
					hes[HIX(a, a, 0)] -= (a124) * re * re + a3;
					hes[HIX(a, a, 1)] -= (a124) * re * im;
					hes[HIX(a, a, 2)] -= (a124) * im * im + a3;
				}
	}
	memset(H, 0, sizeof(double) * (size_t)P * P);
	for (unsigned iRank = 0; iRank < R; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
				const size_t a = prod_idx(L, iRank, iRefl, iWave);
				const int r1 = L->parMap[2 * a], i1 = L->parMap[2 * a + 1];
				if (r1 < 0) continue;
				if (L->pFixed[r1]) continue;
				for (unsigned jRank = 0; jRank < R; ++jRank)
					for (unsigned jRefl = 0; jRefl < 2; ++jRefl)
						for (unsigned jWave = 0; jWave < L->nW[jRefl]; ++jWave) {
							const size_t b = prod_idx(L, jRank, jRefl, jWave);
							const int r2 = L->parMap[2 * b], i2 = L->parMap[2 * b + 1];
							if (r2 < 0) continue;
							if (L->pFixed[r2]) continue;
							H[(size_t)r1 * P + r2] = hes[HIX(a, b, 0)];
							if (i2 >= 0) H[(size_t)r1 * P + i2] = hes[HIX(a, b, 1)];
							if (i1 >= 0) H[(size_t)i1 * P + r2] = hes[HIX(b, a, 1)];
							if (i1 >= 0 && i2 >= 0) H[(size_t)i1 * P + i2] = hes[HIX(a, b, 2)];
						}
				H[(size_t)r1 * P + (P - 1)] = flatTerms[a].re;
				H[(size_t)(P - 1) * P + r1] = flatTerms[a].re;
				if (i1 >= 0) {
					H[(size_t)i1 * P + (P - 1)] = flatTerms[a].im;
					H[(size_t)(P - 1) * P + i1] = flatTerms[a].im;
				}
			}
	H[(size_t)(P - 1) * P + (P - 1)] = hessianFlat;
#undef HIX
	free(hacc); free(flatTermsAcc); free(derivative); free(hes); free(flatTerms); free(T);
}

/* CorrectParamSigns (pwaLikelihood.cc:986-1010) */
void orc_lik_correct_param_signs(const orc_lik* L, const double* par, double* out)
{
	cplx* T = alloc_prod(L);
	double flat;
	load_prod_amps(L, par, T, &flat);
	if (flat < 0)
		flat *= -1;
	for (unsigned iRank = 0; iRank < L->rank; ++iRank)
		for (unsigned iRefl = 0; iRefl < 2; ++iRefl) {
			if (iRank >= L->wmax)
				continue;
			const cplx anchor = T[prod_idx(L, iRank, iRefl, iRank)];
			if (anchor.re < 0 && anchor.im == 0)
				for (unsigned iWave = 0; iWave < L->nW[iRefl]; ++iWave) {
					cplx* t = &T[prod_idx(L, iRank, iRefl, iWave)];
					*t = c_make(-1 * t->re, -1 * t->im);
				}
		}
	orc_lik_copy_to_par(L, (const double*)T, flat, out);
	free(T);
}

/* ------------------------------------------------------------------------- */
/* ampIntegralMatrix::integrate (ampIntegralMatrix.cc:408-483) with S          */
/* incoherent sub-amplitudes per wave, amps [S][M][W]: the sub-amplitude sum    */
/* of one event (:462-464) enters the compensated sum of its matrix element     */
void orc_integral_incoherent(const double* amps, unsigned S, uint64_t M, unsigned W, const double* weights,
                             double* out)
{
	kahan weightAcc = {0, 0};
	kahanc* ampProdAcc = (kahanc*)calloc((size_t)W * W, sizeof(kahanc));
	const int useWeight = weights != NULL;
	for (uint64_t e = 0; e < M; ++e) {
		double w = 1;
		if (useWeight)
			w = weights[e];
		const double weight = 1 / w;
		k_add(&weightAcc, weight);
		for (unsigned i = 0; i < W; ++i)
			for (unsigned j = 0; j < W; ++j) {
				cplx val = c_make(0, 0);
				for (unsigned sub = 0; sub < S; ++sub) {
					const double* A = amps + 2 * ((size_t)sub * M + e) * W;
					val = c_add(val, c_mul(c_ld(A + 2 * i), c_conj(c_ld(A + 2 * j))));
				}
				if (useWeight)
					val = c_scale(weight, val);
				kc_add(&ampProdAcc[(size_t)i * W + j], val);
			}
	}
	const double weightNorm = weightAcc.s / (double)M;
	for (unsigned i = 0; i < W; ++i)
		for (unsigned j = 0; j < W; ++j) {
			cplx s = kc_sum(&ampProdAcc[(size_t)i * W + j]);
			if (useWeight) {
				const double inv = 1 / weightNorm;
				s = c_make(s.re * inv, s.im * inv);
			}
			c_st(out + 2 * ((size_t)i * W + j), s);
		}
	free(ampProdAcc);
}

/* one sub-amplitude, amps [M][W] */
void orc_integral(const double* amps, uint64_t M, unsigned W, const double* weights, double* out)
{
	orc_integral_incoherent(amps, 1, M, W, weights, out);
}

/* modelIntensity::getIntensity (generators/modelIntensity.cc:170-187); the
 * division of each decay amplitude by its phase-space integral is
 * getDecayAmplitudes' (:190-211), the loop over reflectivities the set
 * _allRefls in ascending order (-1, +1), extended by an inner loop over ranks */
void orc_intensity(const double* amps, uint64_t N, const unsigned nWavesRefl[2], const double* prodAmps,
                   unsigned rank, const double* waveNorm, double* out)
{
	const unsigned W = nWavesRefl[0] + nWavesRefl[1];
	const unsigned wmax = nWavesRefl[0] > nWavesRefl[1] ? nWavesRefl[0] : nWavesRefl[1];
	for (uint64_t e = 0; e < N; ++e) {
		double intensity = 0;
		for (unsigned refl = 0; refl < 2; ++refl) {
			if (nWavesRefl[refl] == 0)
				continue;  /* a reflectivity without waves is not in _allRefls */
			for (unsigned r = 0; r < rank; ++r) {
				cplx amp = c_make(0, 0);
				for (unsigned i = 0; i < nWavesRefl[refl]; ++i) {
					const unsigned col = (refl == 0 ? 0 : nWavesRefl[0]) + i;
					cplx a = c_ld(amps + 2 * ((size_t)e * W + col));
					if (waveNorm)
						a = c_make(a.re / waveNorm[col], a.im / waveNorm[col]);
					const cplx t = c_ld(prodAmps + 2 * (((size_t)r * 2 + refl) * wmax + i));
					amp = c_add(amp, c_mul(t, a));
				}
				intensity += c_norm(amp);
			}
		}
		out[e] = intensity;
	}
}

/* testSumAccumulators (test/utilities/testSumAccumulators.cc:46-60) */
void orc_kahan_float_demo(uint64_t n, float x, float* naive, float* compensated)
{
	float s = 0;
	float ks = 0, kc = 0;
	for (uint64_t i = 0; i < n; ++i) {
		s += x;
		const float y = x - kc;
		const float t = ks + y;
		kc = (t - ks) - y;
		ks = t;
	}
	*naive = s;
	*compensated = ks;
}
