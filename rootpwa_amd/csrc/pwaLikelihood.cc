// pwaLikelihood.cc -- host side of the likelihood (see include/rpwa/pwaLikelihood.h).
//
// Event sums run on the GPU through the C ABI; everything below follows the
// reference's host arithmetic in partialWaveFit/pwaLikelihood.cc (line numbers
// cited per function) so that values agree with the reference to the
// precision of the event sums.
#include "rpwa/pwaLikelihood.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "rootpwa_amd.h"

namespace rpwa {

namespace {

// compensated (Kahan) accumulator, utilities/sumAccumulators.hpp:153-184;
// component-wise for complex samples, as Boost.Accumulators applies it
struct kahanSum {
	double s = 0., c = 0.;
	void operator()(const double x)
	{
		const double y = x - c;
		const double t = s + y;
		c = (t - s) - y;
		s = t;
	}
	double sum() const { return s; }
};
struct kahanSumC {
	kahanSum re, im;
	void operator()(const std::complex<double>& x)
	{
		re(x.real());
		im(x.imag());
	}
	std::complex<double> sum() const { return std::complex<double>(re.sum(), im.sum()); }
};

// pwaLikelihood.cc:78-102
double cauchyFunction(const double x, const double gamma)
{
	return 1. / (1. + ((x * x) / (gamma * gamma)));
}
double cauchyFunctionDerivative(const double x, const double gamma)
{
	return (-2. * x * gamma * gamma) / ((gamma * gamma + x * x) * (gamma * gamma + x * x));
}
double cauchyFunctionSecondDerivative(const double x, const double gamma)
{
	const double x2 = x * x;
	const double gamma2 = gamma * gamma;
	return gamma2 / ((gamma2 + x2) * (gamma2 + x2)) * (-2. + 8. * x2 / (gamma2 + x2));
}

// complex / real, component-wise: what GCC emits for std::complex<double>
// divided by a double in the reference build (clang would divide by d + 0i)
inline std::complex<double> cdiv(const std::complex<double>& z, const double d)
{
	return std::complex<double>(z.real() / d, z.imag() / d);
}

double seconds_since(const std::chrono::steady_clock::time_point& t0)
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void check_status(const int status, const char* what)
{
	if (status != RPWA_OK)
		throw std::runtime_error(std::string(what) + ": " + rpwa_last_error());
}

struct ctxDeleter {
	void operator()(rpwa_lik_ctx* c) const
	{
		if (c)
			rpwa_lik_release(c);
	}
};

// Eigen-decomposition of a symmetric n x n matrix: Householder reduction to
// tridiagonal form with the transformations accumulated, then the implicit QL
// iteration with Wilkinson-type shifts (the EISPACK routines tred2 and tql2,
// Wilkinson & Reinsch, Handbook for Automatic Computation II, 1971; the method
// ROOT's TMatrixDSymEigen implements).  On entry V holds the matrix (row-major);
// on exit column k of V is the eigenvector of eigenvalue d[k].
void symmetric_eigen(const int n, std::vector<double>& V, std::vector<double>& d)
{
	auto v = [&](int i, int j) -> double& { return V[(size_t)i * n + j]; };
	d.assign(n, 0.);
	if (n == 0)
		return;
	std::vector<double> e(n, 0.);
	// ---- tridiagonalisation (rows n-1 .. 1), V accumulates the reflections ----
	for (int j = 0; j < n; ++j)
		d[j] = v(n - 1, j);
	for (int i = n - 1; i > 0; --i) {
		double scale = 0., h = 0.;
		for (int k = 0; k < i; ++k)
			scale += std::fabs(d[k]);
		if (scale == 0.) {
			e[i] = d[i - 1];
			for (int j = 0; j < i; ++j) {
				d[j] = v(i - 1, j);
				v(i, j) = 0.;
				v(j, i) = 0.;
			}
		} else {
			for (int k = 0; k < i; ++k) {
				d[k] /= scale;
				h += d[k] * d[k];
			}
			double f = d[i - 1];
			double g = std::sqrt(h);
			if (f > 0.)
				g = -g;
			e[i] = scale * g;
			h -= f * g;
			d[i - 1] = f - g;
			for (int j = 0; j < i; ++j)
				e[j] = 0.;
			for (int j = 0; j < i; ++j) {
				f = d[j];
				v(j, i) = f;
				g = e[j] + v(j, j) * f;
				for (int k = j + 1; k <= i - 1; ++k) {
					g += v(k, j) * d[k];
					e[k] += v(k, j) * f;
				}
				e[j] = g;
			}
			f = 0.;
			for (int j = 0; j < i; ++j) {
				e[j] /= h;
				f += e[j] * d[j];
			}
			const double hh = f / (h + h);
			for (int j = 0; j < i; ++j)
				e[j] -= hh * d[j];
			for (int j = 0; j < i; ++j) {
				f = d[j];
				g = e[j];
				for (int k = j; k <= i - 1; ++k)
					v(k, j) -= (f * e[k] + g * d[k]);
				d[j] = v(i - 1, j);
				v(i, j) = 0.;
			}
		}
		d[i] = h;
	}
	for (int i = 0; i < n - 1; ++i) {
		v(n - 1, i) = v(i, i);
		v(i, i) = 1.;
		const double h = d[i + 1];
		if (h != 0.) {
			for (int k = 0; k <= i; ++k)
				d[k] = v(k, i + 1) / h;
			for (int j = 0; j <= i; ++j) {
				double g = 0.;
				for (int k = 0; k <= i; ++k)
					g += v(k, i + 1) * v(k, j);
				for (int k = 0; k <= i; ++k)
					v(k, j) -= g * d[k];
			}
		}
		for (int k = 0; k <= i; ++k)
			v(k, i + 1) = 0.;
	}
	for (int j = 0; j < n; ++j) {
		d[j] = v(n - 1, j);
		v(n - 1, j) = 0.;
	}
	v(n - 1, n - 1) = 1.;
	e[0] = 0.;
	// ---- implicit QL on the tridiagonal matrix (diagonal d, sub-diagonal e) ----
	for (int i = 1; i < n; ++i)
		e[i - 1] = e[i];
	e[n - 1] = 0.;
	double f = 0., tst1 = 0.;
	const double eps = std::ldexp(1., -52);
	for (int l = 0; l < n; ++l) {
		tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
		int m = l;
		while (m < n - 1 && std::fabs(e[m]) > eps * tst1)
			++m;
		if (m > l) {
			for (int iter = 0; iter < 60; ++iter) {
				double g = d[l];
				double p = (d[l + 1] - g) / (2. * e[l]);
				double r = std::hypot(p, 1.);
				if (p < 0.)
					r = -r;
				d[l] = e[l] / (p + r);
				d[l + 1] = e[l] * (p + r);
				const double dl1 = d[l + 1];
				double h = g - d[l];
				for (int i = l + 2; i < n; ++i)
					d[i] -= h;
				f += h;
				p = d[m];
				double c = 1., c2 = 1., c3 = 1., s = 0., s2 = 0.;
				const double el1 = e[l + 1];
				for (int i = m - 1; i >= l; --i) {
					c3 = c2;
					c2 = c;
					s2 = s;
					g = c * e[i];
					h = c * p;
					r = std::hypot(p, e[i]);
					e[i + 1] = s * r;
					s = e[i] / r;
					c = p / r;
					p = c * d[i] - s * g;
					d[i + 1] = h + s * (c * g + s * d[i]);
					for (int k = 0; k < n; ++k) {
						h = v(k, i + 1);
						v(k, i + 1) = s * v(k, i) + c * h;
						v(k, i) = c * v(k, i) - s * h;
					}
				}
				p = -s * s2 * c3 * el1 * e[l] / dl1;
				e[l] = s * p;
				d[l] = c * p;
				if (!(std::fabs(e[l]) > eps * tst1))
					break;
			}
		}
		d[l] += f;
		e[l] = 0.;
	}
}

template <typename complexT>
int gpu_precision() { return RPWA_PREC_FP64; }
template <>
int gpu_precision<std::complex<float> >() { return RPWA_PREC_MIXED; }

}  // namespace

template <typename complexT>
bool pwaLikelihood<complexT>::_debug = false;

// pwaLikelihood.cc:107-122
template <typename complexT>
std::string pwaLikelihood<complexT>::fitParameter::parName() const
{
	std::ostringstream parName;
	parName << "V";
	if (_waveName != "flat")
		parName << _rank;
	parName << "_" << _waveName;
	if (_realPart)
		parName << "_RE";
	else
		parName << "_IM";
	return parName.str();
}

template <typename complexT>
pwaLikelihood<complexT>::pwaLikelihood()
	: _nmbEvents(0), _rank(1), _nmbWaves(0), _nmbWavesReflMax(0), _nmbPars(0), _nmbParsFixed(0),
	  _initialized(false), _normIntAdded(false), _accIntAdded(false), _initFinished(false), _gpuEnabled(true),
	  _device(0), _useNormalizedAmps(true), _priorType(FLAT), _cauchyWidth(0.5), _numbAccEvents(0), _totAcc(0),
	  _onTheFlyBinning(false), _nmbEntriesTotal(0), _deviceAmps(nullptr), _evalCacheValid(false)
{
	_nmbWavesRefl[0] = 0;
	_nmbWavesRefl[1] = 0;
}

template <typename complexT>
pwaLikelihood<complexT>::pwaLikelihood(const pwaLikelihood& o)
	: Math::IGradientFunctionMultiDim(o), _nmbEvents(o._nmbEvents), _rank(o._rank), _nmbWaves(o._nmbWaves),
	  _nmbWavesReflMax(o._nmbWavesReflMax), _nmbPars(o._nmbPars), _nmbParsFixed(o._nmbParsFixed),
	  _initialized(o._initialized), _normIntAdded(o._normIntAdded), _accIntAdded(o._accIntAdded),
	  _initFinished(o._initFinished), _gpuEnabled(o._gpuEnabled), _device(o._device),
	  _useNormalizedAmps(o._useNormalizedAmps), _priorType(o._priorType), _cauchyWidth(o._cauchyWidth),
	  _numbAccEvents(o._numbAccEvents), _totAcc(o._totAcc), _parameters(o._parameters),
	  _waveParams(o._waveParams), _prodAmpToFuncParMap(o._prodAmpToFuncParMap),
	  _onTheFlyBinning(o._onTheFlyBinning), _nmbEntriesTotal(o._nmbEntriesTotal), _entriesInBin(o._entriesInBin),
	  _deviceAmps(o._deviceAmps),
	  _parCache(o._parCache), _derivCache(o._derivCache), _evalCacheValid(false), _normMatrix(o._normMatrix),
	  _accMatrix(o._accMatrix), _phaseSpaceIntegral(o._phaseSpaceIntegral), _ctx(o._ctx)
{
	for (int r = 0; r < 2; ++r) {
		_nmbWavesRefl[r] = o._nmbWavesRefl[r];
		_waveNames[r] = o._waveNames[r];
		_waveThresholds[r] = o._waveThresholds[r];
		_waveAmpAdded[r] = o._waveAmpAdded[r];
		// the host copy of the amplitudes is only needed until finishInit
		if (not _initFinished)
			_decayAmps[r] = o._decayAmps[r];
	}
	for (int i = 0; i < NMB_FUNCTIONCALLENUM; ++i)
		_funcCallInfo[i] = o._funcCallInfo[i];
}

template <typename complexT>
pwaLikelihood<complexT>::~pwaLikelihood()
{
}

template <typename complexT>
unsigned int pwaLikelihood<complexT>::nmbWaves(const int reflectivity) const
{
	if (reflectivity == 0)
		return _nmbWaves;
	else if (reflectivity > 0)
		return _nmbWavesRefl[1];
	return _nmbWavesRefl[0];
}

template <typename complexT>
void pwaLikelihood<complexT>::requireFinished() const
{
	if (not _initFinished)
		throw std::logic_error("pwaLikelihood::finishInit has not been called");
	if (not _gpuEnabled or not _ctx)
		throw std::runtime_error("pwaLikelihood: the GPU event-sum backend is disabled or not initialised");
}

// ---------------------------------------------------------------------------
// setup: init (pwaLikelihood.cc:1052-1062), readWaveList (:1445-1484),
// buildParDataStruct (:1489-1556)
template <typename complexT>
bool pwaLikelihood<complexT>::init(const std::vector<waveDescThresType>& waveDescThres, const unsigned int rank,
                                   const double massBinCenter)
{
	if (not readWaveList(waveDescThres))
		return false;
	if (not buildParDataStruct(rank, massBinCenter))
		return false;
	_initialized = true;
	return true;
}

template <typename complexT>
bool pwaLikelihood<complexT>::readWaveList(const std::vector<waveDescThresType>& waveDescThres)
{
	for (int r = 0; r < 2; ++r) {
		_nmbWavesRefl[r] = 0;
		_waveNames[r].clear();
		_waveThresholds[r].clear();
		_waveAmpAdded[r].clear();
	}
	_waveParams.clear();
	for (const waveDescThresType& w : waveDescThres) {
		if (w.reflectivity != -1 and w.reflectivity != +1) {
			std::cerr << "pwaLikelihood: wave '" << w.name << "' has reflectivity " << w.reflectivity
			          << " (must be +1 or -1)" << std::endl;
			return false;
		}
		const int r = w.reflectivity > 0 ? 1 : 0;
		++_nmbWavesRefl[r];
		_waveNames[r].push_back(w.name);
		_waveThresholds[r].push_back(w.threshold);
	}
	_nmbWaves = _nmbWavesRefl[0] + _nmbWavesRefl[1];
	_nmbWavesReflMax = std::max(_nmbWavesRefl[0], _nmbWavesRefl[1]);
	for (unsigned int r = 0; r < 2; ++r) {
		_waveAmpAdded[r].assign(_nmbWavesRefl[r], false);
		for (unsigned int w = 0; w < _nmbWavesRefl[r]; ++w)
			_waveParams.insert(std::make_pair(_waveNames[r][w], std::make_pair(r, w)));
	}
	return true;
}

template <typename complexT>
bool pwaLikelihood<complexT>::buildParDataStruct(const unsigned int rank, const double massBinCenter)
{
	if (_nmbWavesRefl[0] + _nmbWavesRefl[1] == 0) {
		std::cerr << "pwaLikelihood: no wave info" << std::endl;
		return false;
	}
	_rank = rank;
	_nmbPars = 0;
	for (unsigned int iRank = 0; iRank < _rank; ++iRank) {
		const int nmbProdAmpsPos = (int)_nmbWavesRefl[1] - (int)iRank;
		int nmbProdAmpsPosC = nmbProdAmpsPos - 1;
		if (nmbProdAmpsPosC < 0)
			nmbProdAmpsPosC = 0;
		const int nmbProdAmpsNeg = (int)_nmbWavesRefl[0] - (int)iRank;
		int nmbProdAmpsNegC = nmbProdAmpsNeg - 1;
		if (nmbProdAmpsNegC < 0)
			nmbProdAmpsNegC = 0;
		_nmbPars += 2 * (nmbProdAmpsPosC + nmbProdAmpsNegC);
		if (nmbProdAmpsPos > 0)
			++_nmbPars;
		if (nmbProdAmpsNeg > 0)
			++_nmbPars;
	}
	_nmbPars += 1;  // flat wave
	_nmbParsFixed = 0;
	_parameters.assign(_nmbPars, fitParameter());
	_parCache.assign(_nmbPars, 0);
	_derivCache.assign(_nmbPars, 0);
	_prodAmpToFuncParMap.assign((size_t)_rank * 2 * _nmbWavesReflMax, std::make_pair(-1, -1));
	unsigned int parIndex = 0;
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				const double thr = _waveThresholds[iRefl][iWave];
				const bool fixed = (thr != 0 and thr >= massBinCenter);
				std::pair<int, int>& m = _prodAmpToFuncParMap[prodIndex(iRank, iRefl, iWave)];
				if (iWave < iRank) {
					m = std::make_pair(-1, -1);
				} else if (iWave == iRank) {
					_parameters[parIndex] = fitParameter(_waveNames[iRefl][iWave], iRank, thr, fixed, true);
					if (fixed)
						++_nmbParsFixed;
					m = std::make_pair((int)parIndex, -1);
					++parIndex;
				} else {
					_parameters[parIndex] = fitParameter(_waveNames[iRefl][iWave], iRank, thr, fixed, true);
					_parameters[parIndex + 1] = fitParameter(_waveNames[iRefl][iWave], iRank, thr, fixed, false);
					if (fixed)
						_nmbParsFixed += 2;
					m = std::make_pair((int)parIndex, (int)parIndex + 1);
					parIndex += 2;
				}
			}
	_parameters[parIndex] = fitParameter("flat", 0, 0., false, true);
	return true;
}

// reorderIntegralMatrix (pwaLikelihood.cc:1562-1574): same-sector blocks only
template <typename complexT>
void pwaLikelihood<complexT>::reorderIntegralMatrix(const ampIntegralMatrix& integral,
                                                    std::vector<std::complex<double> >& reordered) const
{
	reordered.assign((size_t)4 * _nmbWavesReflMax * _nmbWavesReflMax, std::complex<double>(0, 0));
	for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
		for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave)
			for (unsigned int jWave = 0; jWave < _nmbWavesRefl[iRefl]; ++jWave)
				reordered[matIndex(iRefl, iWave, iRefl, jWave)] =
					integral.element(_waveNames[iRefl][iWave], _waveNames[iRefl][jWave]);
}

// pwaLikelihood.cc:1068-1083
template <typename complexT>
bool pwaLikelihood<complexT>::addNormIntegral(const ampIntegralMatrix& normMatrix)
{
	if (not _initialized) {
		std::cerr << "pwaLikelihood: call init() before addNormIntegral()" << std::endl;
		return false;
	}
	for (unsigned int r = 0; r < 2; ++r)
		for (const std::string& n : _waveNames[r])
			if (not normMatrix.containsWave(n)) {
				std::cerr << "pwaLikelihood: wave '" << n << "' missing in the normalisation integral" << std::endl;
				return false;
			}
	_numbAccEvents = normMatrix.nmbEvents();
	reorderIntegralMatrix(normMatrix, _normMatrix);
	_normIntAdded = true;
	return true;
}

// pwaLikelihood.cc:1088-1107
template <typename complexT>
bool pwaLikelihood<complexT>::addAccIntegral(ampIntegralMatrix& accMatrix, const unsigned int accEventsOverride)
{
	if (not _normIntAdded) {
		std::cerr << "pwaLikelihood: call addNormIntegral() before addAccIntegral()" << std::endl;
		return false;
	}
	for (unsigned int r = 0; r < 2; ++r)
		for (const std::string& n : _waveNames[r])
			if (not accMatrix.containsWave(n)) {
				std::cerr << "pwaLikelihood: wave '" << n << "' missing in the acceptance integral" << std::endl;
				return false;
			}
	_numbAccEvents = accEventsOverride == 0 ? _numbAccEvents : accEventsOverride;
	_totAcc = (double)accMatrix.nmbEvents() / _numbAccEvents;
	accMatrix.setNmbEvents(_numbAccEvents);
	reorderIntegralMatrix(accMatrix, _accMatrix);
	_accIntAdded = true;
	return true;
}

// setOnTheFlyBinning (pwaLikelihood.cc:1391-1436): the entries of the event
// files that lie in the bin, as half-open intervals [lower, upper)
template <typename complexT>
bool pwaLikelihood<complexT>::setOnTheFlyBinning(const multibinBoundariesType& multibinBoundaries,
                                                 const std::map<std::string, std::vector<double> >& eventVariables)
{
	if (_nmbEvents != 0 or _deviceAmps) {
		std::cerr << "pwaLikelihood: setOnTheFlyBinning() must come before the amplitudes are added" << std::endl;
		return false;
	}
	if (eventVariables.empty()) {
		std::cerr << "pwaLikelihood: no event variables given for on-the-fly binning" << std::endl;
		return false;
	}
	const unsigned long nmbEntries = eventVariables.begin()->second.size();
	for (const auto& v : eventVariables)
		if (v.second.size() != nmbEntries) {
			std::cerr << "pwaLikelihood: event variable '" << v.first << "' has " << v.second.size()
			          << " entries, expected " << nmbEntries << std::endl;
			return false;
		}
	std::vector<const double*> values;
	std::vector<double> lower, upper;
	for (const auto& b : multibinBoundaries) {
		const auto it = eventVariables.find(b.first);
		if (it == eventVariables.end()) {
			std::cerr << "pwaLikelihood: binning variable '" << b.first << "' is not among the event variables"
			          << std::endl;
			return false;
		}
		if (_debug)
			std::cout << "pwaLikelihood: use bin in '" << b.first << "' with range [" << b.second.first << ", "
			          << b.second.second << ")" << std::endl;
		values.push_back(it->second.data());
		lower.push_back(b.second.first);
		upper.push_back(b.second.second);
	}
	std::vector<unsigned long> entries;
	for (unsigned long e = 0; e < nmbEntries; ++e) {
		bool useEvent = true;
		for (size_t k = 0; k < values.size(); ++k) {
			const double v = values[k][e];
			// a NaN compares false both ways and is kept, as in the reference
			if (v < lower[k] or v >= upper[k]) {
				useEvent = false;
				break;
			}
		}
		if (useEvent)
			entries.push_back(e);
	}
	if (_debug)
		std::cout << "pwaLikelihood: found " << entries.size() << " of " << nmbEntries
		          << " entries to be in the given bin" << std::endl;
	_onTheFlyBinning = true;
	_nmbEntriesTotal = nmbEntries;
	_entriesInBin.swap(entries);
	return true;
}

// addAmplitude (pwaLikelihood.cc:1112-1288) for one wave's amplitude array
template <typename complexT>
bool pwaLikelihood<complexT>::addAmplitude(const std::string& waveName, const std::complex<double>* amps,
                                           const unsigned long nmbEntries)
{
	if (not _accIntAdded) {
		std::cerr << "pwaLikelihood: call addAccIntegral() before addAmplitude()" << std::endl;
		return false;
	}
	if (_deviceAmps) {
		std::cerr << "pwaLikelihood: amplitudes were already given in device memory" << std::endl;
		return false;
	}
	if (_onTheFlyBinning and nmbEntries != _nmbEntriesTotal) {
		std::cerr << "pwaLikelihood: amplitude array of wave '" << waveName << "' has " << nmbEntries
		          << " entries, the event variables of the on-the-fly binning have " << _nmbEntriesTotal << std::endl;
		return false;
	}
	const unsigned long nmbEvents = _onTheFlyBinning ? _entriesInBin.size() : nmbEntries;
	if (not amps or nmbEvents == 0) {
		std::cerr << "pwaLikelihood: no events to load" << std::endl;
		return false;
	}
	if (_waveParams.count(waveName) == 0) {
		std::cerr << "pwaLikelihood: wave '" << waveName << "' is not in the wave list" << std::endl;
		return false;
	}
	if (_nmbEvents == 0) {
		_nmbEvents = nmbEvents;
		for (int r = 0; r < 2; ++r)
			_decayAmps[r].assign((size_t)_nmbEvents * _nmbWavesRefl[r], std::complex<double>(0, 0));
	}
	if (nmbEvents != _nmbEvents) {
		std::cerr << "pwaLikelihood: size mismatch in amplitude arrays: " << nmbEvents << " events, previous "
		          << _nmbEvents << std::endl;
		return false;
	}
	const unsigned int refl = _waveParams[waveName].first;
	const unsigned int waveIndex = _waveParams[waveName].second;
	const std::complex<double> normInt = _normMatrix[matIndex(refl, waveIndex, refl, waveIndex)];
	bool fixPar = (_accMatrix[matIndex(refl, waveIndex, refl, waveIndex)] == 0.);
	const unsigned int W = _nmbWavesRefl[refl];
	for (unsigned long iEvt = 0; iEvt < _nmbEvents; ++iEvt) {
		std::complex<double> a = amps[_onTheFlyBinning ? _entriesInBin[iEvt] : iEvt];
		if (_useNormalizedAmps) {
			if (normInt == 0. and a != 0.) {
				std::cerr << "pwaLikelihood: normalization integral for wave '" << waveName
				          << "' is zero, but the amplitude is not" << std::endl;
				return false;
			}
			if (normInt != 0.)
				a = cdiv(a, std::sqrt(normInt.real()));
		}
		if (a != 0.)
			fixPar = false;
		_decayAmps[refl][(size_t)iEvt * W + waveIndex] = a;
	}
	if (fixPar) {
		// as the reference: the parameters are fixed, _nmbParsFixed is not changed
		for (fitParameter& par : _parameters)
			if (par.waveName() == waveName and not par.fixed()) {
				std::cerr << "pwaLikelihood: all amplitudes of wave '" << waveName
				          << "' are zero, fixing its production amplitudes to zero" << std::endl;
				par.fix(true);
			}
	}
	_waveAmpAdded[refl][waveIndex] = true;
	return true;
}

template <typename complexT>
bool pwaLikelihood<complexT>::addAmplitudesDevice(const double* d_amps, const unsigned long nmbEvents)
{
	if (not _accIntAdded) {
		std::cerr << "pwaLikelihood: call addAccIntegral() before addAmplitudesDevice()" << std::endl;
		return false;
	}
	if (not d_amps or nmbEvents == 0) {
		std::cerr << "pwaLikelihood: no events to load" << std::endl;
		return false;
	}
	if (_nmbEvents != 0) {
		std::cerr << "pwaLikelihood: amplitudes were already added" << std::endl;
		return false;
	}
	if (_onTheFlyBinning) {
		std::cerr << "pwaLikelihood: on-the-fly binning selects from host amplitude arrays; device amplitudes "
		          << "must hold the events of the bin only" << std::endl;
		return false;
	}
	// the zero-amplitude rules of addAmplitude (pwaLikelihood.cc:1247-1281), with the
	// columns inspected on the device only for waves with a zero integral diagonal
	std::vector<std::string> toFix;
	for (unsigned int r = 0; r < 2; ++r)
		for (unsigned int w = 0; w < _nmbWavesRefl[r]; ++w) {
			const bool accZero = _accMatrix[matIndex(r, w, r, w)] == 0.;
			const bool normZero = _useNormalizedAmps and _normMatrix[matIndex(r, w, r, w)] == 0.;
			if (not accZero and not normZero)
				continue;
			int nonzero = 0;
			const uint32_t col = (r == 0 ? 0 : _nmbWavesRefl[0]) + w;
			if (rpwa_wave_nonzero_device(_device, d_amps, nmbEvents, _nmbWaves, col, &nonzero) != RPWA_OK) {
				std::cerr << "pwaLikelihood: inspecting the device amplitudes of wave '" << _waveNames[r][w]
				          << "': " << rpwa_last_error() << std::endl;
				return false;
			}
			if (normZero and nonzero) {
				std::cerr << "pwaLikelihood: normalization integral for wave '" << _waveNames[r][w]
				          << "' is zero, but the amplitude is not" << std::endl;
				return false;
			}
			if (accZero and not nonzero)
				toFix.push_back(_waveNames[r][w]);
		}
	for (const std::string& waveName : toFix)
		for (fitParameter& par : _parameters)
			if (par.waveName() == waveName and not par.fixed()) {
				std::cerr << "pwaLikelihood: all amplitudes of wave '" << waveName
				          << "' are zero, fixing its production amplitudes to zero" << std::endl;
				par.fix(true);
			}
	_nmbEvents = nmbEvents;
	_deviceAmps = d_amps;
	for (unsigned int r = 0; r < 2; ++r)
		_waveAmpAdded[r].assign(_nmbWavesRefl[r], true);
	return true;
}

// finishInit (pwaLikelihood.cc:1293-1388)
template <typename complexT>
bool pwaLikelihood<complexT>::finishInit()
{
	for (unsigned int r = 0; r < 2; ++r)
		for (unsigned int w = 0; w < _nmbWavesRefl[r]; ++w)
			if (not _waveAmpAdded[r][w]) {
				std::cerr << "pwaLikelihood: amplitude of wave '" << _waveNames[r][w] << "' was not added"
				          << std::endl;
				return false;
			}
	_phaseSpaceIntegral.assign((size_t)2 * _nmbWavesReflMax, 0.);
	for (unsigned int r = 0; r < 2; ++r)
		for (unsigned int w = 0; w < _nmbWavesRefl[r]; ++w)
			_phaseSpaceIntegral[r * _nmbWavesReflMax + w] = std::sqrt(_normMatrix[matIndex(r, w, r, w)].real());
	// per-wave divisor of device-resident raw amplitudes (sqrt(Re I_ii), as addAmplitude divides)
	std::vector<double> waveNorm(_nmbWaves, 1.);
	if (_useNormalizedAmps) {
		for (unsigned int r = 0; r < 2; ++r)
			for (unsigned int w = 0; w < _nmbWavesRefl[r]; ++w) {
				const std::complex<double> n = _normMatrix[matIndex(r, w, r, w)];
				if (n != 0.)
					waveNorm[(r == 0 ? 0 : _nmbWavesRefl[0]) + w] = std::sqrt(n.real());
			}
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				double norm_i = std::sqrt(_normMatrix[matIndex(iRefl, iWave, iRefl, iWave)].real());
				if (norm_i == 0)
					norm_i = 1;
				for (unsigned int jRefl = 0; jRefl < 2; ++jRefl)
					for (unsigned int jWave = 0; jWave < _nmbWavesRefl[jRefl]; ++jWave) {
						double norm_j = std::sqrt(_normMatrix[matIndex(jRefl, jWave, jRefl, jWave)].real());
						if (norm_j == 0)
							norm_j = 1;
						const size_t ij = matIndex(iRefl, iWave, jRefl, jWave);
						if (iRefl != jRefl or iWave != jWave)
							_normMatrix[ij] = cdiv(_normMatrix[ij], norm_i * norm_j);
						_accMatrix[ij] = cdiv(_accMatrix[ij], norm_i * norm_j);
					}
			}
		for (unsigned int r = 0; r < 2; ++r)
			for (unsigned int w = 0; w < _nmbWavesRefl[r]; ++w)
				if (_normMatrix[matIndex(r, w, r, w)] != 0.)
					_normMatrix[matIndex(r, w, r, w)] = 1;
	}
	if (_gpuEnabled) {
		rpwa_lik_ctx* c = nullptr;
		const int st = rpwa_lik_create(&c, _device, gpu_precision<complexT>());
		if (st != RPWA_OK) {
			std::cerr << "pwaLikelihood: GPU context: " << rpwa_last_error() << std::endl;
			return false;
		}
		_ctx.reset(c, ctxDeleter());
		const uint32_t nW[2] = {_nmbWavesRefl[0], _nmbWavesRefl[1]};
		int rc;
		if (_deviceAmps) {
			rc = rpwa_lik_load_decay_amps_device(c, _deviceAmps, _nmbEvents, nW,
			                                     _useNormalizedAmps ? waveNorm.data() : nullptr);
		} else {
			// [N][W0 + W1], refl -1 waves first (already normalised on the host)
			std::vector<std::complex<double> > rows((size_t)_nmbEvents * _nmbWaves);
			for (unsigned long e = 0; e < _nmbEvents; ++e) {
				for (unsigned int w = 0; w < _nmbWavesRefl[0]; ++w)
					rows[(size_t)e * _nmbWaves + w] = _decayAmps[0][(size_t)e * _nmbWavesRefl[0] + w];
				for (unsigned int w = 0; w < _nmbWavesRefl[1]; ++w)
					rows[(size_t)e * _nmbWaves + _nmbWavesRefl[0] + w] =
						_decayAmps[1][(size_t)e * _nmbWavesRefl[1] + w];
			}
			rc = rpwa_lik_load_decay_amps(c, reinterpret_cast<const double*>(rows.data()), rows.size(),
			                              _nmbEvents, nW, RPWA_LAYOUT_EVT_WAVE);
		}
		if (rc != RPWA_OK) {
			std::cerr << "pwaLikelihood: loading amplitudes to the GPU: " << rpwa_last_error() << std::endl;
			_ctx.reset();
			return false;
		}
	}
	_initFinished = true;
	return true;
}

// ---------------------------------------------------------------------------
// parameter mapping (pwaLikelihood.cc:1682-1731)
template <typename complexT>
void pwaLikelihood<complexT>::copyFromParArray(const double* inPar, std::vector<std::complex<double> >& outVal,
                                               double& outFlatVal) const
{
	outVal.assign((size_t)_rank * 2 * _nmbWavesReflMax, std::complex<double>(0, 0));
	unsigned int parIndex = 0;
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				double re, im;
				if (iWave < iRank) {
					re = im = 0;
				} else if (iWave == iRank) {
					re = inPar[parIndex];
					im = 0;
					++parIndex;
				} else {
					re = inPar[parIndex];
					im = inPar[parIndex + 1];
					parIndex += 2;
				}
				outVal[prodIndex(iRank, iRefl, iWave)] = std::complex<double>(re, im);
			}
	outFlatVal = inPar[parIndex];
}

template <typename complexT>
void pwaLikelihood<complexT>::copyToParArray(const std::vector<std::complex<double> >& inVal,
                                             const double inFlatVal, double* outPar) const
{
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				const std::pair<int, int>& m = _prodAmpToFuncParMap[prodIndex(iRank, iRefl, iWave)];
				if (m.first >= 0)
					outPar[m.first] = inVal[prodIndex(iRank, iRefl, iWave)].real();
				if (m.second >= 0)
					outPar[m.second] = inVal[prodIndex(iRank, iRefl, iWave)].imag();
			}
	outPar[_nmbPars - 1] = inFlatVal;
}

// ---------------------------------------------------------------------------
// host terms

// normalisation term and its derivative (FdF :255-279, DoEval :384-399, Gradient :588-606).
// The derivative contributions (sum_j T_j conj(I_ij)) * 2nu and flat * 2nu * eta
// are written to normDeriv / normDerivFlat, to be added to the event sums in the
// reference's order once those arrive from the GPU.
template <typename complexT>
double pwaLikelihood<complexT>::normTerm(const std::vector<std::complex<double> >& prodAmps, const double prodAmpFlat,
                                         std::vector<std::complex<double> >* normDeriv,
                                         double* normDerivFlat) const
{
	kahanSum normFactorAcc;
	const double nmbEvt = _useNormalizedAmps ? 1 : (double)_nmbEvents;
	const double twiceNmbEvt = 2 * nmbEvt;
	if (normDeriv)
		normDeriv->assign(prodAmps.size(), std::complex<double>(0, 0));
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				kahanSumC normFactorDerivAcc;
				const std::complex<double> Ti = prodAmps[prodIndex(iRank, iRefl, iWave)];
				for (unsigned int jWave = 0; jWave < _nmbWavesRefl[iRefl]; ++jWave) {
					const std::complex<double> I = _accMatrix[matIndex(iRefl, iWave, iRefl, jWave)];
					const std::complex<double> Tj = prodAmps[prodIndex(iRank, iRefl, jWave)];
					normFactorAcc(std::real((Ti * std::conj(Tj)) * I));
					if (normDeriv)
						normFactorDerivAcc(Tj * std::conj(I));
				}
				if (normDeriv)
					(*normDeriv)[prodIndex(iRank, iRefl, iWave)] = normFactorDerivAcc.sum() * twiceNmbEvt;
			}
	const double prodAmpFlat2 = prodAmpFlat * prodAmpFlat;
	normFactorAcc(prodAmpFlat2 * _totAcc);
	if (normDerivFlat)
		*normDerivFlat = prodAmpFlat * twiceNmbEvt * _totAcc;
	return normFactorAcc.sum();
}

// half-Cauchy prior (FdF :281-297, DoEval :401-417, Gradient :608-625); the
// derivative contributions go to priorDeriv (subtracted later, as the reference does)
template <typename complexT>
double pwaLikelihood<complexT>::priorTerm(const std::vector<std::complex<double> >& prodAmps,
                                          std::vector<std::complex<double> >* priorDeriv) const
{
	double priorValue = 0.;
	if (priorDeriv)
		priorDeriv->clear();
	if (_priorType != HALF_CAUCHY)
		return priorValue;
	if (priorDeriv)
		priorDeriv->assign(prodAmps.size(), std::complex<double>(0, 0));
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				const std::complex<double> t = prodAmps[prodIndex(iRank, iRefl, iWave)];
				const double r = std::abs(t);
				const double cauchyFunctionValue = cauchyFunction(r, _cauchyWidth);
				priorValue -= std::log(cauchyFunctionValue);
				if (priorDeriv) {
					const double factor = (1. / (r * cauchyFunctionValue)) * cauchyFunctionDerivative(r, _cauchyWidth);
					(*priorDeriv)[prodIndex(iRank, iRefl, iWave)] =
						std::complex<double>(factor * t.real(), factor * t.imag());
				}
			}
	return priorValue;
}

// derivatives = event sums + norm term - prior, in the reference's order
template <typename complexT>
static void combine_derivatives(std::vector<std::complex<double> >& derivatives,
                                const std::vector<std::complex<double> >& normDeriv,
                                const std::vector<std::complex<double> >& priorDeriv)
{
	for (size_t i = 0; i < derivatives.size(); ++i)
		derivatives[i] += normDeriv[i];
	if (not priorDeriv.empty())
		for (size_t i = 0; i < derivatives.size(); ++i)
			derivatives[i] -= priorDeriv[i];
}

template <typename complexT>
void pwaLikelihood<complexT>::eventSums(const std::vector<std::complex<double> >& prodAmps, const double prodAmpFlat,
                                        double* lnl, std::vector<std::complex<double> >* deriv,
                                        double* derivFlat) const
{
	const uint32_t n = (uint32_t)prodAmps.size();
	const double* p = reinterpret_cast<const double*>(prodAmps.data());
	if (deriv) {
		deriv->assign(prodAmps.size(), std::complex<double>(0, 0));
		double* d = reinterpret_cast<double*>(deriv->data());
		check_status(rpwa_lik_fdf(_ctx.get(), p, n, prodAmpFlat, _rank, lnl, d, derivFlat), "rpwa_lik_fdf");
	} else {
		check_status(rpwa_lik_log_likelihood(_ctx.get(), p, n, prodAmpFlat, _rank, lnl), "rpwa_lik_log_likelihood");
	}
}

// FdF (pwaLikelihood.cc:172-315): the event sums run on the GPU while the host
// computes the normalisation term and the prior
template <typename complexT>
void pwaLikelihood<complexT>::fdfImpl(const double* par, double& funcVal, double* gradient,
                                      const functionCallEnum who) const
{
	const auto t0 = std::chrono::steady_clock::now();
	std::vector<std::complex<double> > prodAmps;
	double prodAmpFlat;
	copyFromParArray(par, prodAmps, prodAmpFlat);
	check_status(rpwa_lik_submit(_ctx.get(), reinterpret_cast<const double*>(prodAmps.data()),
	                             (uint32_t)prodAmps.size(), prodAmpFlat, _rank),
	             "rpwa_lik_submit");
	const auto t1 = std::chrono::steady_clock::now();
	const double nmbEvt = _useNormalizedAmps ? 1 : (double)_nmbEvents;
	std::vector<std::complex<double> > normDeriv, priorDeriv;
	double normDerivFlat = 0;
	const double normFactor = normTerm(prodAmps, prodAmpFlat, &normDeriv, &normDerivFlat);
	const double tNorm = seconds_since(t1);
	const double priorValue = priorTerm(prodAmps, &priorDeriv);
	std::vector<std::complex<double> > derivatives(prodAmps.size());
	double derivativeFlat = 0, logLikelihood = 0;
	check_status(rpwa_lik_collect(_ctx.get(), &logLikelihood, reinterpret_cast<double*>(derivatives.data()),
	                              &derivativeFlat),
	             "rpwa_lik_collect");
	combine_derivatives<complexT>(derivatives, normDeriv, priorDeriv);
	derivativeFlat += normDerivFlat;
	copyToParArray(derivatives, derivativeFlat, gradient);
	funcVal = logLikelihood + nmbEvt * normFactor + priorValue;
	const double tTot = seconds_since(t0);
	_funcCallInfo[who].funcTime += tTot - tNorm;
	_funcCallInfo[who].normTime += tNorm;
	_funcCallInfo[who].totalTime += tTot;
	if (_debug)
		std::cout << "raw log likelihood = " << std::setprecision(17) << logLikelihood
		          << ", normalization = " << normFactor << ", prior = " << priorValue
		          << ", normalized log likelihood = " << funcVal << std::endl;
}

template <typename complexT>
void pwaLikelihood<complexT>::FdF(const double* par, double& funcVal, double* gradient) const
{
	requireFinished();
	++_funcCallInfo[FDF].nmbCalls;
	for (unsigned int i = 0; i < _nmbPars; ++i)
		_parCache[i] = par[i];
	fdfImpl(par, funcVal, gradient, FDF);
	std::copy(gradient, gradient + _nmbPars, _derivCache.begin());
}

// DoEval (pwaLikelihood.cc:320-435).  One fused pass produces the value and the
// gradient; the gradient is kept for a Gradient() call at the same point, the
// call order of Minuit2 (DoEval then Gradient), so that costs one pass, not two.
// The value is bitwise that of FdF at the same point (same kernel, same host code).
template <typename complexT>
double pwaLikelihood<complexT>::DoEval(const double* par) const
{
	requireFinished();
	++_funcCallInfo[DOEVAL].nmbCalls;
	double funcVal;
	_evalGradCache.resize(_nmbPars);
	fdfImpl(par, funcVal, _evalGradCache.data(), DOEVAL);
	_evalParCache.assign(par, par + _nmbPars);
	_evalCacheValid = true;
	return funcVal;
}

// DoDerivative (pwaLikelihood.cc:440-471)
template <typename complexT>
double pwaLikelihood<complexT>::DoDerivative(const double* par, const unsigned int derivativeIndex) const
{
	requireFinished();
	++_funcCallInfo[DODERIVATIVE].nmbCalls;
	bool samePar = true;
	for (unsigned int i = 0; i < _nmbPars; ++i)
		if (_parCache[i] != par[i]) {
			samePar = false;
			break;
		}
	if (samePar)
		return _derivCache[derivativeIndex];
	double logLikelihood;
	std::vector<double> gradient(_nmbPars);
	FdF(par, logLikelihood, gradient.data());
	return gradient[derivativeIndex];
}

// Gradient (pwaLikelihood.cc:477-636)
template <typename complexT>
void pwaLikelihood<complexT>::Gradient(const double* par, double* gradient) const
{
	requireFinished();
	++_funcCallInfo[GRADIENT].nmbCalls;
	for (unsigned int i = 0; i < _nmbPars; ++i)
		_parCache[i] = par[i];
	bool cached = _evalCacheValid;
	if (cached)
		for (unsigned int i = 0; i < _nmbPars; ++i)
			if (_evalParCache[i] != par[i]) {
				cached = false;
				break;
			}
	if (cached) {
		std::copy(_evalGradCache.begin(), _evalGradCache.end(), gradient);
	} else {
		const auto t0 = std::chrono::steady_clock::now();
		std::vector<std::complex<double> > prodAmps;
		double prodAmpFlat;
		copyFromParArray(par, prodAmps, prodAmpFlat);
		check_status(rpwa_lik_submit(_ctx.get(), reinterpret_cast<const double*>(prodAmps.data()),
		                             (uint32_t)prodAmps.size(), prodAmpFlat, _rank),
		             "rpwa_lik_submit");
		const auto t1 = std::chrono::steady_clock::now();
		std::vector<std::complex<double> > normDeriv, priorDeriv;
		double normDerivFlat = 0;
		normTerm(prodAmps, prodAmpFlat, &normDeriv, &normDerivFlat);
		_funcCallInfo[GRADIENT].normTime += seconds_since(t1);
		priorTerm(prodAmps, &priorDeriv);
		std::vector<std::complex<double> > derivatives(prodAmps.size());
		double derivativeFlat = 0;
		check_status(rpwa_lik_collect(_ctx.get(), nullptr, reinterpret_cast<double*>(derivatives.data()),
		                              &derivativeFlat),
		             "rpwa_lik_collect");
		combine_derivatives<complexT>(derivatives, normDeriv, priorDeriv);
		derivativeFlat += normDerivFlat;
		copyToParArray(derivatives, derivativeFlat, gradient);
		_funcCallInfo[GRADIENT].totalTime += seconds_since(t0);
		_funcCallInfo[GRADIENT].funcTime += seconds_since(t0);
	}
	std::copy(gradient, gradient + _nmbPars, _derivCache.begin());
}

// Hessian (pwaLikelihood.cc:642-859): event sums on the GPU, host terms as the reference
template <typename complexT>
denseMatrix pwaLikelihood<complexT>::Hessian(const double* par) const
{
	requireFinished();
	++_funcCallInfo[HESSIAN].nmbCalls;
	const auto t0 = std::chrono::steady_clock::now();
	std::vector<std::complex<double> > prodAmps;
	double prodAmpFlat;
	copyFromParArray(par, prodAmps, prodAmpFlat);
	const size_t nd = prodAmps.size();
	std::vector<double> hessian(nd * nd * 3);
	std::vector<std::complex<double> > flatTerms(nd);
	double hessianFlat = 0;
	check_status(rpwa_lik_hessian(_ctx.get(), reinterpret_cast<const double*>(prodAmps.data()), (uint32_t)nd,
	                              prodAmpFlat, _rank, hessian.data(), reinterpret_cast<double*>(flatTerms.data()),
	                              &hessianFlat),
	             "rpwa_lik_hessian");
	_funcCallInfo[HESSIAN].funcTime += seconds_since(t0);
	auto H = [&](size_t a, size_t b, int c) -> double& { return hessian[(a * nd + b) * 3 + c]; };
	const double nmbEvt = _useNormalizedAmps ? 1 : (double)_nmbEvents;
	const double twiceNmbEvt = 2 * nmbEvt;
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave)
				for (unsigned int jWave = 0; jWave < _nmbWavesRefl[iRefl]; ++jWave) {
					const size_t a = prodIndex(iRank, iRefl, iWave), b = prodIndex(iRank, iRefl, jWave);
					const std::complex<double> I = _accMatrix[matIndex(iRefl, iWave, iRefl, jWave)];
					H(a, b, 0) += I.real() * twiceNmbEvt;
					H(a, b, 1) += I.imag() * twiceNmbEvt;
					H(a, b, 2) += I.real() * twiceNmbEvt;
				}
	hessianFlat += twiceNmbEvt * _totAcc;
	if (_priorType == HALF_CAUCHY)
		for (unsigned int iRank = 0; iRank < _rank; ++iRank)
			for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
				for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
					const size_t a = prodIndex(iRank, iRefl, iWave);
					const double r = std::abs(prodAmps[a]);
					const double cv = cauchyFunction(r, _cauchyWidth);
					const double cd = cauchyFunctionDerivative(r, _cauchyWidth);
					const double cdd = cauchyFunctionSecondDerivative(r, _cauchyWidth);
					const double a1 = cdd / (cv * r * r);
					const double a2 = cd * cd / (cv * cv * r * r);
					const double a3 = cd / (cv * r);
					const double a4 = cd / (cv * r * r * r);
					const double a124 = a1 - a2 - a4;
					const double re = prodAmps[a].real(), im = prodAmps[a].imag();
					H(a, a, 0) -= (a124) * re * re + a3;
					H(a, a, 1) -= (a124) * re * im;
					H(a, a, 2) -= (a124) * im * im + a3;
				}
	denseMatrix out(_nmbPars);
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				const size_t a = prodIndex(iRank, iRefl, iWave);
				const int r1 = _prodAmpToFuncParMap[a].first, i1 = _prodAmpToFuncParMap[a].second;
				if (r1 < 0 or _parameters[r1].fixed())
					continue;
				for (unsigned int jRank = 0; jRank < _rank; ++jRank)
					for (unsigned int jRefl = 0; jRefl < 2; ++jRefl)
						for (unsigned int jWave = 0; jWave < _nmbWavesRefl[jRefl]; ++jWave) {
							const size_t b = prodIndex(jRank, jRefl, jWave);
							const int r2 = _prodAmpToFuncParMap[b].first, i2 = _prodAmpToFuncParMap[b].second;
							if (r2 < 0 or _parameters[r2].fixed())
								continue;
							out(r1, r2) = H(a, b, 0);
							if (i2 >= 0)
								out(r1, i2) = H(a, b, 1);
							if (i1 >= 0)
								out(i1, r2) = H(b, a, 1);
							if (i1 >= 0 and i2 >= 0)
								out(i1, i2) = H(a, b, 2);
						}
				out(r1, _nmbPars - 1) = flatTerms[a].real();
				out(_nmbPars - 1, r1) = flatTerms[a].real();
				if (i1 >= 0) {
					out(i1, _nmbPars - 1) = flatTerms[a].imag();
					out(_nmbPars - 1, i1) = flatTerms[a].imag();
				}
			}
	out(_nmbPars - 1, _nmbPars - 1) = hessianFlat;
	_funcCallInfo[HESSIAN].totalTime += seconds_since(t0);
	return out;
}

// CovarianceMatrix (pwaLikelihood.cc:905-981): inverse of the Hessian reduced
// to the free parameters (Gauss-Jordan with partial pivoting)
template <typename complexT>
denseMatrix pwaLikelihood<complexT>::CovarianceMatrix(const denseMatrix& hessian) const
{
	std::vector<unsigned int> freeIdx;
	for (unsigned int i = 0; i < _nmbPars; ++i)
		if (not _parameters[i].fixed())
			freeIdx.push_back(i);
	const unsigned int n = (unsigned int)freeIdx.size();
	std::vector<double> a((size_t)n * 2 * n, 0.);
	for (unsigned int i = 0; i < n; ++i) {
		for (unsigned int j = 0; j < n; ++j)
			a[(size_t)i * 2 * n + j] = hessian(freeIdx[i], freeIdx[j]);
		a[(size_t)i * 2 * n + n + i] = 1.;
	}
	for (unsigned int c = 0; c < n; ++c) {
		unsigned int piv = c;
		for (unsigned int r = c + 1; r < n; ++r)
			if (std::fabs(a[(size_t)r * 2 * n + c]) > std::fabs(a[(size_t)piv * 2 * n + c]))
				piv = r;
		if (a[(size_t)piv * 2 * n + c] == 0.)
			throw std::runtime_error("pwaLikelihood::CovarianceMatrix: singular Hessian");
		if (piv != c)
			for (unsigned int k = 0; k < 2 * n; ++k)
				std::swap(a[(size_t)piv * 2 * n + k], a[(size_t)c * 2 * n + k]);
		const double d = a[(size_t)c * 2 * n + c];
		for (unsigned int k = 0; k < 2 * n; ++k)
			a[(size_t)c * 2 * n + k] /= d;
		for (unsigned int r = 0; r < n; ++r)
			if (r != c) {
				const double f = a[(size_t)r * 2 * n + c];
				if (f != 0.)
					for (unsigned int k = 0; k < 2 * n; ++k)
						a[(size_t)r * 2 * n + k] -= f * a[(size_t)c * 2 * n + k];
			}
	}
	denseMatrix cov(_nmbPars);
	for (unsigned int i = 0; i < n; ++i)
		for (unsigned int j = 0; j < n; ++j)
			cov(freeIdx[i], freeIdx[j]) = a[(size_t)i * 2 * n + n + j];
	return cov;
}

template <typename complexT>
denseMatrix pwaLikelihood<complexT>::CovarianceMatrix(const double* par) const
{
	return CovarianceMatrix(Hessian(par));
}

// HessianEigenVectors (pwaLikelihood.cc:864-911): eigen-decomposition of the
// Hessian reduced to the free parameters (symmetric_eigen above); eigenvectors
// are expanded back to all parameters with zeros at the fixed ones.  Order:
// descending eigenvalues.
template <typename complexT>
std::vector<std::pair<std::vector<double>, double> >
pwaLikelihood<complexT>::HessianEigenVectors(const denseMatrix& hessian) const
{
	std::vector<unsigned int> freeIdx;
	for (unsigned int i = 0; i < _nmbPars; ++i)
		if (not _parameters[i].fixed())
			freeIdx.push_back(i);
	const unsigned int n = (unsigned int)freeIdx.size();
	std::vector<double> v((size_t)n * n), lambda;
	for (unsigned int i = 0; i < n; ++i)
		for (unsigned int j = 0; j < n; ++j)
			// symmetrise: the reference's Hessian is symmetric by construction
			v[(size_t)i * n + j] = 0.5 * (hessian(freeIdx[i], freeIdx[j]) + hessian(freeIdx[j], freeIdx[i]));
	symmetric_eigen((int)n, v, lambda);
	std::vector<unsigned int> order(n);
	for (unsigned int i = 0; i < n; ++i)
		order[i] = i;
	std::stable_sort(order.begin(), order.end(), [&](unsigned int x, unsigned int y) {
		return lambda[x] > lambda[y];
	});
	std::vector<std::pair<std::vector<double>, double> > eigen;
	eigen.reserve(n);
	for (unsigned int k = 0; k < n; ++k) {
		std::vector<double> vec(_nmbPars, 0.);
		for (unsigned int j = 0; j < n; ++j)
			vec[freeIdx[j]] = v[(size_t)j * n + order[k]];
		eigen.push_back(std::make_pair(vec, lambda[order[k]]));
	}
	return eigen;
}

// CorrectParamSigns (pwaLikelihood.cc:986-1010)
template <typename complexT>
std::vector<double> pwaLikelihood<complexT>::CorrectParamSigns(const double* par) const
{
	std::vector<std::complex<double> > prodAmps;
	double prodAmpFlat;
	copyFromParArray(par, prodAmps, prodAmpFlat);
	if (prodAmpFlat < 0)
		prodAmpFlat *= -1;
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl) {
			if (iRank >= _nmbWavesReflMax)
				continue;
			const std::complex<double> anchor = prodAmps[prodIndex(iRank, iRefl, iRank)];
			if (anchor.real() < 0 and anchor.imag() == 0)
				for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave)
					prodAmps[prodIndex(iRank, iRefl, iWave)] *= -1;
		}
	std::vector<double> returnParams(_nmbPars);
	copyToParArray(prodAmps, prodAmpFlat, returnParams.data());
	return returnParams;
}

// getIntegralMatrices (pwaLikelihood.cc:1579-1619): row-major (W[+1])^2
template <typename complexT>
void pwaLikelihood<complexT>::getIntegralMatrices(std::vector<std::complex<double> >& normMatrix,
                                                  std::vector<std::complex<double> >& accMatrix,
                                                  std::vector<double>& phaseSpaceIntegral,
                                                  const bool withFlat) const
{
	const unsigned int size = _nmbWaves + (withFlat ? 1 : 0);
	normMatrix.assign((size_t)size * size, std::complex<double>(0, 0));
	accMatrix.assign((size_t)size * size, std::complex<double>(0, 0));
	phaseSpaceIntegral.assign(size, 0.);
	unsigned int iIndex = 0;
	for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
		for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
			phaseSpaceIntegral[iIndex] = _phaseSpaceIntegral[iRefl * _nmbWavesReflMax + iWave];
			unsigned int jIndex = 0;
			for (unsigned int jRefl = 0; jRefl < 2; ++jRefl)
				for (unsigned int jWave = 0; jWave < _nmbWavesRefl[jRefl]; ++jWave) {
					normMatrix[(size_t)iIndex * size + jIndex] = _normMatrix[matIndex(iRefl, iWave, jRefl, jWave)];
					accMatrix[(size_t)iIndex * size + jIndex] = _accMatrix[matIndex(iRefl, iWave, jRefl, jWave)];
					++jIndex;
				}
			++iIndex;
		}
	if (withFlat) {
		normMatrix[(size_t)_nmbWaves * size + _nmbWaves] = 1.;
		accMatrix[(size_t)_nmbWaves * size + _nmbWaves] = _totAcc;
		phaseSpaceIntegral[_nmbWaves] = 1.;
	}
}

// buildProdAmpArrays (pwaLikelihood.cc:1624-1661)
template <typename complexT>
void pwaLikelihood<complexT>::buildProdAmpArrays(const double* inPar, std::vector<std::complex<double> >& prodAmps,
                                                 std::vector<std::pair<int, int> >& parIndices,
                                                 std::vector<std::string>& prodAmpNames, const bool withFlat) const
{
	prodAmps.clear();
	parIndices.clear();
	prodAmpNames.clear();
	unsigned int parIndex = 0;
	for (unsigned int iRank = 0; iRank < _rank; ++iRank)
		for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
			for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave) {
				double re, im;
				if (iWave < iRank)
					continue;
				else if (iWave == iRank) {
					parIndices.push_back(std::make_pair((int)parIndex, -1));
					re = inPar[parIndex];
					im = 0;
					++parIndex;
				} else {
					parIndices.push_back(std::make_pair((int)parIndex, (int)parIndex + 1));
					re = inPar[parIndex];
					im = inPar[parIndex + 1];
					parIndex += 2;
				}
				prodAmps.push_back(std::complex<double>(re, im));
				std::ostringstream name;
				name << "V" << iRank << "_" << _waveNames[iRefl][iWave];
				prodAmpNames.push_back(name.str());
			}
	if (withFlat) {
		prodAmps.push_back(std::complex<double>(inPar[parIndex], 0));
		parIndices.push_back(std::make_pair((int)parIndex, -1));
		prodAmpNames.push_back("V_flat");
	}
}

template <typename complexT>
std::ostream& pwaLikelihood<complexT>::print(std::ostream& out) const
{
	out << "pwaLikelihood parameters:" << std::endl
	    << "number of events ........................ " << _nmbEvents << std::endl
	    << "rank .................................... " << _rank << std::endl
	    << "number of waves ......................... " << _nmbWaves << std::endl
	    << "number of positive reflectivity waves ... " << _nmbWavesRefl[1] << std::endl
	    << "number of negative reflectivity waves ... " << _nmbWavesRefl[0] << std::endl
	    << "number of function parameters ........... " << _nmbPars << std::endl
	    << "number of fixed function parameters ..... " << _nmbParsFixed << std::endl
	    << "use GPU event sums ...................... " << _gpuEnabled << std::endl
	    << "use normalized amplitudes ............... " << _useNormalizedAmps << std::endl
	    << "list of waves: " << std::endl;
	for (unsigned int iRefl = 0; iRefl < 2; ++iRefl)
		for (unsigned int iWave = 0; iWave < _nmbWavesRefl[iRefl]; ++iWave)
			out << "        [" << std::setw(2) << ((int)iRefl * 2 - 1) << " " << std::setw(3) << iWave << "] "
			    << _waveNames[iRefl][iWave] << "    threshold = " << _waveThresholds[iRefl][iWave] << " GeV/c^2"
			    << std::endl;
	out << "list of function parameters: " << std::endl;
	for (unsigned int iPar = 0; iPar < _nmbPars; ++iPar)
		out << "        [" << std::setw(3) << iPar << "] " << _parameters[iPar].parName() << "    "
		    << "threshold = " << _parameters[iPar].threshold() << " GeV/c^2" << std::endl;
	return out;
}

template <typename complexT>
std::ostream& pwaLikelihood<complexT>::printFuncInfo(std::ostream& out) const
{
	const std::string funcNames[NMB_FUNCTIONCALLENUM] = {"FdF", "Gradient", "DoEval", "DoDerivative", "Hessian"};
	for (unsigned int i = 0; i < NMB_FUNCTIONCALLENUM; ++i)
		if (_funcCallInfo[i].nmbCalls > 0)
			out << "    " << _funcCallInfo[i].nmbCalls << " calls to pwaLikelihood<complexT>::" << funcNames[i]
			    << "()" << std::endl
			    << "    time spent in pwaLikelihood<complexT>::" << funcNames[i] << "(): " << std::endl
			    << "        total time ................. " << _funcCallInfo[i].totalTime << " sec" << std::endl
			    << "        return value calculation ... " << _funcCallInfo[i].funcTime << " sec" << std::endl
			    << "        normalization .............. " << _funcCallInfo[i].normTime << " sec" << std::endl;
	return out;
}

template class pwaLikelihood<std::complex<double> >;
template class pwaLikelihood<std::complex<float> >;

}  // namespace rpwa
