// fitDrivers.cc -- L-BFGS minimiser and the pwaNloptFit / pwaFit flows
// (declarations and references in include/rpwa/fitDrivers.h).
#include "rpwa/fitDrivers.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace rpwa {

namespace {

double dot(const std::vector<double>& a, const std::vector<double>& b)
{
	double s = 0.;
	for (size_t i = 0; i < a.size(); ++i)
		s += a[i] * b[i];
	return s;
}

// the function restricted to the free coordinates, counting evaluations and
// remembering the best point seen (NLopt returns the best evaluated point)
class reducedFunction {
public:
	reducedFunction(const Math::IGradientFunctionMultiDim& f, const std::vector<double>& x0,
	                const std::vector<bool>& freePar, const bool minuitPattern)
		: _f(f), _full(x0), _gradFull(x0.size()), _minuit(minuitPattern)
	{
		for (size_t i = 0; i < freePar.size(); ++i)
			if (freePar[i])
				_idx.push_back((unsigned int)i);
	}

	size_t dim() const { return _idx.size(); }
	unsigned int nmbEvaluations() const { return _nEval; }

	std::vector<double> reduce(const std::vector<double>& full) const
	{
		std::vector<double> r(_idx.size());
		for (size_t k = 0; k < _idx.size(); ++k)
			r[k] = full[_idx[k]];
		return r;
	}

	std::vector<double> expand(const std::vector<double>& x) const
	{
		std::vector<double> full = _full;
		for (size_t k = 0; k < _idx.size(); ++k)
			full[_idx[k]] = x[k];
		return full;
	}

	// value and gradient: FdF, or DoEval followed by Gradient (Minuit pattern)
	double valueGradient(const std::vector<double>& x, std::vector<double>& g)
	{
		const std::vector<double> full = expand(x);
		double fv;
		if (_minuit) {
			fv = _f(full.data());
			_f.Gradient(full.data(), _gradFull.data());
		} else
			_f.FdF(full.data(), fv, _gradFull.data());
		++_nEval;
		g.resize(_idx.size());
		for (size_t k = 0; k < _idx.size(); ++k)
			g[k] = _gradFull[_idx[k]];
		record(x, fv);
		return fv;
	}

	// gradient only (Gradient), at a point whose value is known
	void gradient(const std::vector<double>& x, std::vector<double>& g)
	{
		const std::vector<double> full = expand(x);
		_f.Gradient(full.data(), _gradFull.data());
		g.resize(_idx.size());
		for (size_t k = 0; k < _idx.size(); ++k)
			g[k] = _gradFull[_idx[k]];
	}

	// value only (DoEval)
	double value(const std::vector<double>& x)
	{
		const std::vector<double> full = expand(x);
		const double fv = _f(full.data());
		++_nEval;
		record(x, fv);
		return fv;
	}

	bool haveBest() const { return _haveBest; }
	const std::vector<double>& bestX() const { return _bestX; }
	double bestF() const { return _bestF; }

private:
	void record(const std::vector<double>& x, const double fv)
	{
		if (std::isfinite(fv) and (not _haveBest or fv < _bestF)) {
			_bestF = fv;
			_bestX = x;
			_haveBest = true;
		}
	}

	const Math::IGradientFunctionMultiDim& _f;
	std::vector<unsigned int> _idx;
	std::vector<double> _full;
	std::vector<double> _gradFull;
	const bool _minuit;
	unsigned int _nEval = 0;
	bool _haveBest = false;
	std::vector<double> _bestX;
	double _bestF = 0.;
};

// minimiser of the cubic through (a, fa, da) and (b, fb, db) (Nocedal & Wright
// eq. 3.59), safeguarded into [lo + 0.1 w, hi - 0.1 w]; bisection when the cubic
// has no minimum there
double cubicStep(const double a, const double fa, const double da, const double b, const double fb,
                 const double db)
{
	const double lo = std::min(a, b), hi = std::max(a, b), w = hi - lo;
	const double d1 = da + db - 3. * (fa - fb) / (a - b);
	const double disc = d1 * d1 - da * db;
	double t = 0.5 * (a + b);
	if (disc >= 0.) {
		const double d2 = (b > a ? 1. : -1.) * std::sqrt(disc);
		const double den = db - da + 2. * d2;
		if (den != 0.)
			t = b - (b - a) * (db + d2 - d1) / den;
	}
	if (not std::isfinite(t) or t < lo + 0.1 * w or t > hi - 0.1 * w)
		t = 0.5 * (a + b);
	return t;
}

struct lineSearchPoint {
	double alpha;
	double f;
	double dphi;
	std::vector<double> x;
	std::vector<double> g;
};

// strong-Wolfe line search (Nocedal & Wright algorithms 3.5 and 3.6) along d
// from x with value f, directional derivative dphi0 < 0; c1 = 1e-4, c2 = 0.9
bool strongWolfe(reducedFunction& F, const std::vector<double>& x, const double f, const double dphi0,
                 const std::vector<double>& d, double alpha, const unsigned int maxEval, lineSearchPoint& out)
{
	const double c1 = 1e-4, c2 = 0.9;
	const size_t n = x.size();
	auto evalAt = [&](const double a, lineSearchPoint& p) {
		p.alpha = a;
		p.x.resize(n);
		for (size_t k = 0; k < n; ++k)
			p.x[k] = x[k] + a * d[k];
		p.f = F.valueGradient(p.x, p.g);
		p.dphi = std::isfinite(p.f) ? dot(p.g, d) : std::numeric_limits<double>::quiet_NaN();
	};
	lineSearchPoint prev;
	prev.alpha = 0.;
	prev.f = f;
	prev.dphi = dphi0;
	lineSearchPoint lo, hi, cur;
	bool bracketed = false;
	for (unsigned int i = 0; i < 40; ++i) {
		if (F.nmbEvaluations() >= maxEval)
			return false;
		evalAt(alpha, cur);
		if (not std::isfinite(cur.f) or not std::isfinite(cur.dphi)) {  // step too long: shrink
			alpha = 0.5 * (prev.alpha + alpha);
			continue;
		}
		if (cur.f > f + c1 * alpha * dphi0 or (i > 0 and cur.f >= prev.f)) {
			lo = prev;
			hi = cur;
			bracketed = true;
			break;
		}
		if (std::fabs(cur.dphi) <= -c2 * dphi0) {
			out = cur;
			return true;
		}
		if (cur.dphi >= 0.) {
			lo = cur;
			hi = prev;
			bracketed = true;
			break;
		}
		prev = cur;
		alpha *= 4.;
	}
	if (not bracketed)
		return false;
	// zoom (algorithm 3.6); lo always satisfies the sufficient-decrease condition
	for (unsigned int i = 0; i < 40; ++i) {
		if (F.nmbEvaluations() >= maxEval)
			break;
		const double a = cubicStep(lo.alpha, lo.f, lo.dphi, hi.alpha, hi.f, hi.dphi);
		if (a == lo.alpha or a == hi.alpha)
			break;
		evalAt(a, cur);
		if (not std::isfinite(cur.f) or cur.f > f + c1 * a * dphi0 or cur.f >= lo.f) {
			hi = cur;
			if (not std::isfinite(cur.f))
				hi.f = std::numeric_limits<double>::max(), hi.dphi = 0.;
		} else {
			if (std::fabs(cur.dphi) <= -c2 * dphi0) {
				out = cur;
				return true;
			}
			if (cur.dphi * (hi.alpha - lo.alpha) >= 0.)
				hi = lo;
			lo = cur;
		}
	}
	// no strong-Wolfe point: accept the best sufficient-decrease point if any
	if (lo.alpha > 0. and lo.f < f) {
		out = lo;
		return true;
	}
	return false;
}

// Armijo backtracking on DoEval with quadratic interpolation (the Minuit
// pattern evaluates only the value during its line search)
bool armijo(reducedFunction& F, const std::vector<double>& x, const double f, const double dphi0,
            const std::vector<double>& d, double alpha, const unsigned int maxEval, lineSearchPoint& out)
{
	const double c1 = 1e-4;
	const size_t n = x.size();
	std::vector<double> xt(n);
	for (unsigned int i = 0; i < 60; ++i) {
		if (F.nmbEvaluations() >= maxEval)
			return false;
		for (size_t k = 0; k < n; ++k)
			xt[k] = x[k] + alpha * d[k];
		const double ft = F.value(xt);
		if (std::isfinite(ft) and ft <= f + c1 * alpha * dphi0) {
			out.alpha = alpha;
			out.x = xt;
			out.f = ft;
			F.gradient(xt, out.g);  // Gradient at the accepted point
			out.dphi = dot(out.g, d);
			return true;
		}
		// minimiser of the parabola through phi(0), phi'(0), phi(alpha)
		double a = std::isfinite(ft) ? -dphi0 * alpha * alpha / (2. * (ft - f - dphi0 * alpha)) : 0.5 * alpha;
		if (not std::isfinite(a) or a < 0.1 * alpha or a > 0.5 * alpha)
			a = (a < 0.1 * alpha or not std::isfinite(a)) ? 0.1 * alpha : 0.5 * alpha;
		alpha = a;
		if (alpha == 0.)
			return false;
	}
	return false;
}

}  // namespace

minimizerResult lbfgsMinimize(const Math::IGradientFunctionMultiDim& f, const std::vector<double>& x0,
                              const std::vector<bool>& freePar, const minimizerOptions& opts)
{
	minimizerResult res;
	if (x0.size() != f.NDim() or freePar.size() != x0.size() or opts.memory == 0) {
		res.status = MIN_INVALID_ARGS;
		res.x = x0;
		return res;
	}
	reducedFunction F(f, x0, freePar, opts.minuitCallPattern);
	const size_t n = F.dim();
	std::vector<double> x = F.reduce(x0), g;
	double fx = F.valueGradient(x, g);
	if (not std::isfinite(fx)) {
		res.status = MIN_FAILURE;
		res.x = x0;
		res.fval = fx;
		res.nmbEvaluations = F.nmbEvaluations();
		return res;
	}

	std::deque<std::vector<double> > S, Y;
	std::deque<double> rho;
	std::vector<double> d(n), q(n), alphaHist;
	double gamma = 0.;  // H0 = gamma * I; 0 until the first update
	auto applyH = [&](const std::vector<double>& v, std::vector<double>& out) {
		// two-loop recursion (Nocedal & Wright algorithm 7.4): out = H v
		out = v;
		const size_t m = S.size();
		alphaHist.assign(m, 0.);
		for (size_t i = m; i-- > 0;) {
			alphaHist[i] = rho[i] * dot(S[i], out);
			for (size_t k = 0; k < n; ++k)
				out[k] -= alphaHist[i] * Y[i][k];
		}
		for (size_t k = 0; k < n; ++k)
			out[k] *= gamma;
		for (size_t i = 0; i < m; ++i) {
			const double beta = rho[i] * dot(Y[i], out);
			for (size_t k = 0; k < n; ++k)
				out[k] += S[i][k] * (alphaHist[i] - beta);
		}
	};

	int status = MIN_MAXEVAL_REACHED;
	unsigned int iter = 0;
	double edm = std::numeric_limits<double>::infinity();
	const double edmGoal = 0.002 * opts.edmTolerance;
	bool restarted = false;
	if (n == 0)
		status = MIN_SUCCESS;
	while (n > 0) {
		if (F.nmbEvaluations() >= opts.maxEvaluations) {
			status = MIN_MAXEVAL_REACHED;
			break;
		}
		double alpha0 = 1.;
		if (S.empty()) {
			// steepest descent with a step of unit length in parameter space
			const double gn = std::sqrt(dot(g, g));
			if (gn == 0.) {
				status = MIN_SUCCESS;
				edm = 0.;
				break;
			}
			for (size_t k = 0; k < n; ++k)
				d[k] = -g[k];
			alpha0 = 1. / gn;
		} else {
			applyH(g, q);
			for (size_t k = 0; k < n; ++k)
				d[k] = -q[k];
		}
		double dphi0 = dot(g, d);
		if (not (dphi0 < 0.)) {  // not a descent direction: restart from steepest descent
			S.clear(), Y.clear(), rho.clear();
			continue;
		}
		lineSearchPoint p;
		const bool ok = opts.minuitCallPattern
		                ? armijo(F, x, fx, dphi0, d, alpha0, opts.maxEvaluations, p)
		                : strongWolfe(F, x, fx, dphi0, d, alpha0, opts.maxEvaluations, p);
		if (not ok) {
			if (F.nmbEvaluations() >= opts.maxEvaluations) {
				status = MIN_MAXEVAL_REACHED;
				break;
			}
			if (not S.empty() and not restarted) {  // retry once along steepest descent
				S.clear(), Y.clear(), rho.clear();
				restarted = true;
				continue;
			}
			status = opts.minuitCallPattern and edm < 4e-8 * std::fabs(fx) ? MIN_MACHINE_PRECISION
			                                                                  : MIN_ROUNDOFF_LIMITED;
			break;
		}
		restarted = false;
		++iter;
		std::vector<double> s(n), y(n);
		for (size_t k = 0; k < n; ++k) {
			s[k] = p.x[k] - x[k];
			y[k] = p.g[k] - g[k];
		}
		const double fOld = fx;
		bool xConverged = opts.xtolRel > 0.;
		for (size_t k = 0; k < n and xConverged; ++k) {
			const double dx = std::fabs(p.x[k] - x[k]);
			if (not (dx < opts.xtolRel * 0.5 * (std::fabs(p.x[k]) + std::fabs(x[k])) or p.x[k] == x[k]))
				xConverged = false;
		}
		x = p.x;
		g = p.g;
		fx = p.f;
		const double sy = dot(s, y);
		if (sy > 1e-12 * std::sqrt(dot(s, s) * dot(y, y))) {
			if (S.size() == opts.memory) {
				S.pop_front(), Y.pop_front(), rho.pop_front();
			}
			S.push_back(s), Y.push_back(y), rho.push_back(1. / sy);
			gamma = sy / dot(y, y);
		}
		if (opts.minuitCallPattern) {
			if (S.empty())
				edm = std::numeric_limits<double>::infinity();
			else {
				applyH(g, q);
				edm = 0.5 * dot(g, q);
			}
			if (edm < edmGoal) {
				status = MIN_EDM_REACHED;
				break;
			}
			if (edm < 4e-8 * std::fabs(fx) and std::fabs(fOld - fx) <= 4e-16 * std::fabs(fx) * 8.) {
				status = MIN_MACHINE_PRECISION;
				break;
			}
		} else {
			if (opts.ftolAbs > 0. and std::fabs(fx - fOld) < opts.ftolAbs) {
				status = MIN_FTOL_REACHED;
				break;
			}
			if (xConverged) {
				status = MIN_XTOL_REACHED;
				break;
			}
		}
	}
	res.status = status;
	res.nmbIterations = iter;
	res.nmbEvaluations = F.nmbEvaluations();
	res.edm = edm;
	if (F.haveBest()) {
		res.x = F.expand(F.bestX());
		res.fval = F.bestF();
	} else {
		res.x = F.expand(x);
		res.fval = fx;
	}
	return res;
}

namespace hli {

minimizerOptions minuitDefaults()
{
	minimizerOptions o;
	o.minuitCallPattern = true;
	o.maxEvaluations = 40000;
	o.edmTolerance = 1e-10;
	o.xtolRel = 0.;
	o.ftolAbs = 0.;
	return o;
}

namespace {

// the post-minimisation part both drivers share: fixed parameters to zero,
// sign convention with the bitwise check, Hessian eigenvalues (eigenCheck),
// covariance (wantCovariance), result arrays
template <typename complexT>
void finishResult(const pwaLikelihood<complexT>& L, const minimizerResult& m, const bool eigenCheck,
                  const bool wantCovariance, const bool saveSpace, const bool verbose, fitResultData& out,
                  bool& covarianceOk)
{
	covarianceOk = false;
	const unsigned int nmbPars = L.nmbPars();
	out.status = m.status;
	out.nmbEvaluations = m.nmbEvaluations;
	out.nmbIterations = m.nmbIterations;
	out.edm = m.edm;
	out.logLikelihood = m.fval;
	out.converged = m.status > 0;

	// fixed parameters are zero (pwaNloptFit.cc:405-425; pwaFit.cc:267-271)
	std::vector<double> params = m.x;
	for (unsigned int i = 0; i < nmbPars; ++i)
		if (L.parameter(i).fixed())
			params[i] = 0.;
	out.fitParameters = L.CorrectParamSigns(params.data());
	const double newLikelihood = L.DoEval(out.fitParameters.data());
	if (m.fval != newLikelihood) {
		std::ostringstream msg;
		msg << std::setprecision(17) << "flipping signs according to sign conventions changed the likelihood (from "
		    << m.fval << " to " << newLikelihood << ")";
		throw signConventionError(msg.str());
	}
	if (verbose)
		std::cout << "likelihood unchanged at " << std::setprecision(17) << newLikelihood
		          << " by flipping signs according to conventions" << std::endl;

	if (eigenCheck or wantCovariance) {
		const denseMatrix hessian = L.Hessian(out.fitParameters.data());
		if (eigenCheck) {
			const auto eigen = L.HessianEigenVectors(hessian);
			for (size_t i = 0; i < eigen.size(); ++i) {
				out.hessianEigenValues.push_back(eigen[i].second);
				if (eigen[i].second <= 0.) {
					if (verbose)
						std::cout << "eigenvalue " << i << " of (analytic) Hessian is not positive ("
						          << eigen[i].second << ")" << std::endl;
					out.converged = false;
				}
			}
		}
		if (wantCovariance) {
			try {
				out.covMatrix = L.CovarianceMatrix(hessian);
				covarianceOk = true;
			} catch (const std::runtime_error&) {
				// singular Hessian: no covariance
			}
		}
	}

	L.buildProdAmpArrays(out.fitParameters.data(), out.prodAmps, out.covMatrixIndices, out.prodAmpNames, true);
	if (not saveSpace)
		L.getIntegralMatrices(out.normIntegral, out.accIntegral, out.phaseSpaceIntegral, true);
	out.nmbEvents = L.nmbEventsLong();
	out.normNmbEvents = L.normalizedAmpsUsed() ? 1 : (int)L.nmbEvents();
	out.rank = L.rank();
	if (verbose)
		L.printFuncInfo(std::cout);
}

}  // namespace

template <typename complexT>
std::vector<double> randomStartValues(const pwaLikelihood<complexT>& L, const unsigned int seed)
{
	const unsigned int nmbPars = L.NDim();
	const double sqrtNmbEvts = std::sqrt((double)L.nmbEvents());
	const double defaultStartValue = 0.01;
	// every parameter draws a value and a sign; fixed ones are then set to zero
	// (pwaNloptFit.cc:183-212)
	mt19937Uniform random(seed);
	std::vector<double> x0(nmbPars, 0.);
	for (unsigned int i = 0; i < nmbPars; ++i) {
		double startVal = random.uniform(defaultStartValue, sqrtNmbEvts);
		if (random.rndm() > 0.5)
			startVal *= -1.;
		if (not L.parameter(i).fixed())
			x0[i] = startVal;
	}
	return x0;
}

template <typename complexT>
fitResultData pwaNloptFit(const pwaLikelihood<complexT>& L, const unsigned int seed, const bool checkHessian,
                          const bool saveSpace, const bool verbose, const minimizerOptions& opts)
{
	const unsigned int nmbPars = L.NDim();
	const std::vector<double> x0 = randomStartValues(L, seed);
	std::vector<bool> freePar(nmbPars);
	for (unsigned int i = 0; i < nmbPars; ++i) {
		freePar[i] = not L.parameter(i).fixed();
		if (verbose)
			std::cout << (freePar[i] ? "    setting parameter [" : "    fixing parameter  [") << std::setw(3) << i
			          << "] " << L.parameter(i).parName() << " = " << x0[i] << std::endl;
	}
	minimizerOptions o = opts;
	o.minuitCallPattern = false;
	const minimizerResult m = lbfgsMinimize(L, x0, freePar, o);
	fitResultData out;
	// pwaNloptFit.cc:246-268: the Hessian's eigenvalues are checked whenever it
	// is computed; the covariance counts as a Hessian only for a converged fit
	bool covarianceOk;
	finishResult(L, m, checkHessian or not saveSpace, not saveSpace, saveSpace, verbose, out, covarianceOk);
	out.hasHessian = covarianceOk and out.converged;
	return out;
}

template <typename complexT>
fitResultData pwaFit(const pwaLikelihood<complexT>& L, const unsigned int seed,
                     const std::vector<double>* startValues, const bool checkHessian, const bool saveSpace,
                     const bool verbose, const minimizerOptions& opts)
{
	const unsigned int nmbPars = L.NDim();
	if (startValues and startValues->size() != nmbPars)
		throw std::invalid_argument("pwaFit: start values need one entry per parameter");
	const double sqrtNmbEvts = std::sqrt((double)L.nmbEvents());
	const double defaultStartValue = 0.01;
	// pwaFit.cc:230-271: start values from the earlier result, random ones
	// otherwise and for zeros of free parameters; fixed parameters sit at 0
	mt19937Uniform random(seed);
	auto draw = [&]() {
		double v = random.uniform(defaultStartValue, sqrtNmbEvts);
		if (random.rndm() > 0.5)
			v *= -1.;
		return v;
	};
	std::vector<double> x0(nmbPars, 0.);
	std::vector<bool> freePar(nmbPars);
	for (unsigned int i = 0; i < nmbPars; ++i) {
		double startVal = startValues ? (*startValues)[i] : draw();
		freePar[i] = not L.parameter(i).fixed();
		if (freePar[i]) {
			if (startVal == 0.)
				startVal = draw();
			x0[i] = startVal;
		}
	}
	minimizerOptions o = opts;
	o.minuitCallPattern = true;
	const minimizerResult m = lbfgsMinimize(L, x0, freePar, o);
	fitResultData out;
	// pwaFit.cc:304-354: eigenvalue check on request; the covariance (Hesse)
	// unless saveSpace, and a failed one marks the fit as not converged
	bool covarianceOk;
	finishResult(L, m, checkHessian, not saveSpace, saveSpace, verbose, out, covarianceOk);
	out.hasHessian = covarianceOk;
	if (not saveSpace and not covarianceOk)
		out.converged = false;
	return out;
}

template std::vector<double> randomStartValues(const pwaLikelihood<std::complex<double> >&, unsigned int);
template std::vector<double> randomStartValues(const pwaLikelihood<std::complex<float> >&, unsigned int);
template fitResultData pwaNloptFit(const pwaLikelihood<std::complex<double> >&, unsigned int, bool, bool, bool,
                                   const minimizerOptions&);
template fitResultData pwaNloptFit(const pwaLikelihood<std::complex<float> >&, unsigned int, bool, bool, bool,
                                   const minimizerOptions&);
template fitResultData pwaFit(const pwaLikelihood<std::complex<double> >&, unsigned int, const std::vector<double>*,
                              bool, bool, bool, const minimizerOptions&);
template fitResultData pwaFit(const pwaLikelihood<std::complex<float> >&, unsigned int, const std::vector<double>*,
                              bool, bool, bool, const minimizerOptions&);

}  // namespace hli
}  // namespace rpwa
