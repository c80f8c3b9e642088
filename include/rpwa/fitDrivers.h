// fitDrivers.h -- the two fit flows ROOTPWA runs around the likelihood, using
// only its minimiser interface (DoEval / Gradient / FdF) and its post-fit
// methods (CorrectParamSigns, Hessian, HessianEigenVectors, CovarianceMatrix,
// buildProdAmpArrays, getIntegralMatrices).
//
// pwaNloptFit (highLevelInterface/pwaNloptFit.cc:84-402): start values
// Uniform(0.01, sqrt(N)) with a random sign per parameter, fixed parameters
// removed from the minimisation and inserted as zero (:45-80, :405-425),
// L-BFGS on FdF with xtol_rel 1e-4, ftol_abs 1e-6, maxeval 50000 (:105-107,
// :152-157), success when the result code is positive (:228-231), then
// CorrectParamSigns and the bitwise DoEval check (:236-244), then the analytic
// Hessian, its eigenvalues and its inverse (:246-268).
//
// pwaFit (highLevelInterface/pwaFit.cc:64-433): the same flow with Minuit's
// call pattern (DoEval, then Gradient at the accepted point), start values
// optionally taken from an earlier result (:235-257), tolerance 1e-10 and
// 40000 calls (:86-93), the bitwise check (:295-300) and the covariance from
// the Hessian (:330-354; Minuit's 0.5 * CovMatrix is the inverse Hessian of
// -lnL, which this flow takes from the analytic Hessian).
//
// NLopt and Minuit2 are not available here.  The minimiser is a limited-memory
// BFGS (Nocedal & Wright, Numerical Optimization, 2nd ed., algorithm 7.4) with
// a strong-Wolfe line search (algorithms 3.5 and 3.6) on FdF, or an Armijo
// backtracking search on DoEval in the Minuit pattern.  Stopping tests follow
// NLopt's documented semantics (ftol_abs: |df| < tol; xtol_rel: every
// |dx_i| < tol * (|x_i| + |x_i'|) / 2) and Minuit's (EDM = g'Vg/2 below
// 0.002 * tolerance, or below the machine-precision bound 4e-8 * |f|).
// Trajectories therefore differ from the reference's; minima agree.
#pragma once

#include <complex>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "IGradientFunctionMultiDim.h"
#include "pwaLikelihood.h"

namespace rpwa {

// NLopt's result codes (nlopt.h), kept so that "result > 0" means what it
// means in pwaNloptFit.cc:228-231 (MAXEVAL_REACHED counts as success there)
enum minimizerStatus {
	MIN_FAILURE = -1,
	MIN_INVALID_ARGS = -2,
	MIN_ROUNDOFF_LIMITED = -4,
	MIN_SUCCESS = 1,
	MIN_FTOL_REACHED = 3,
	MIN_XTOL_REACHED = 4,
	MIN_MAXEVAL_REACHED = 5,
	MIN_EDM_REACHED = 7,        // Minuit pattern: EDM below 0.002 * tolerance
	MIN_MACHINE_PRECISION = 8,  // Minuit pattern: EDM below 4e-8 * |f|
};

struct minimizerOptions {
	unsigned int memory = 10;        // L-BFGS correction pairs
	unsigned int maxEvaluations = 50000;
	double xtolRel = 1e-4;           // 0: off
	double ftolAbs = 1e-6;           // 0: off
	bool minuitCallPattern = false;  // DoEval (+ Gradient at accepted points) instead of FdF
	double edmTolerance = 1e-10;     // Minuit pattern only: stop at EDM < 0.002 * edmTolerance
};

struct minimizerResult {
	std::vector<double> x;  // all parameters (fixed ones at their start value)
	double fval = 0.;       // the value the function returned at x (bitwise)
	double edm = 0.;
	int status = MIN_FAILURE;
	unsigned int nmbEvaluations = 0;
	unsigned int nmbIterations = 0;
};

// minimise f over the coordinates with freePar[i] == true; the others stay at x0[i]
minimizerResult lbfgsMinimize(const Math::IGradientFunctionMultiDim& f, const std::vector<double>& x0,
                              const std::vector<bool>& freePar, const minimizerOptions& opts);

// The MT19937 generator (Matsumoto & Nishimura 1998, seeded by init_genrand)
// with the conversion ROOT's TRandom3 documents: a 32-bit draw times 2^-32,
// zero excluded; Uniform(a, b) = a + (b - a) * Rndm().
class mt19937Uniform {
public:
	explicit mt19937Uniform(uint32_t seed) : _gen(seed) {}
	double rndm()
	{
		for (;;) {
			const uint32_t y = (uint32_t)_gen();
			if (y)
				return (double)y * 2.3283064365386963e-10;
		}
	}
	double uniform(double a, double b) { return a + (b - a) * rndm(); }

private:
	std::mt19937 _gen;
};

namespace hli {

// what fitResult::fill receives in pwaFit.cc:415-428 / pwaNloptFit.cc:385-398,
// plus the minimiser bookkeeping
struct fitResultData {
	unsigned long nmbEvents = 0;
	int normNmbEvents = 0;
	double logLikelihood = 0.;               // the minimum (-lnL) as the minimiser returned it
	unsigned int rank = 0;
	std::vector<double> fitParameters;       // after CorrectParamSigns, all P parameters
	std::vector<std::complex<double> > prodAmps;
	std::vector<std::string> prodAmpNames;
	denseMatrix covMatrix;                   // P x P, empty with saveSpace
	std::vector<std::pair<int, int> > covMatrixIndices;
	std::vector<std::complex<double> > normIntegral;  // empty with saveSpace
	std::vector<std::complex<double> > accIntegral;
	std::vector<double> phaseSpaceIntegral;
	std::vector<double> hessianEigenValues;  // descending; empty unless computed
	bool converged = false;
	bool hasHessian = false;
	int status = MIN_FAILURE;
	double edm = 0.;
	unsigned int nmbEvaluations = 0;
	unsigned int nmbIterations = 0;
};

// thrown when CorrectParamSigns changes the likelihood value (the reference
// aborts there, pwaNloptFit.cc:238-241, pwaFit.cc:295-298)
class signConventionError : public std::runtime_error {
public:
	explicit signConventionError(const std::string& what) : std::runtime_error(what) {}
};

// pwaFit.cc:86-93: tolerance 1e-10, 40000 function calls
minimizerOptions minuitDefaults();

// start values of pwaNloptFit.cc:183-212 (and of pwaFit.cc:235-257 without a
// start-value file) for all P parameters, fixed ones zero
template <typename complexT>
std::vector<double> randomStartValues(const pwaLikelihood<complexT>& L, const unsigned int seed);

// pwaNloptFit.cc:84-402; opts defaults are the reference's settings
template <typename complexT>
fitResultData pwaNloptFit(const pwaLikelihood<complexT>& L, const unsigned int seed = 0,
                          const bool checkHessian = false, const bool saveSpace = false,
                          const bool verbose = false, const minimizerOptions& opts = minimizerOptions());

// pwaFit.cc:64-433; startValues (all P parameters, or null) replaces the
// start-value file: a zero for a free parameter draws a random start value
template <typename complexT>
fitResultData pwaFit(const pwaLikelihood<complexT>& L, const unsigned int seed = 0,
                     const std::vector<double>* startValues = nullptr, const bool checkHessian = false,
                     const bool saveSpace = false, const bool verbose = false,
                     const minimizerOptions& opts = minuitDefaults());

}  // namespace hli
}  // namespace rpwa
