// ampIntegralMatrix.cc -- see include/rpwa/ampIntegralMatrix.h.
#include "rpwa/ampIntegralMatrix.h"

#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "rootpwa_amd.h"

namespace rpwa {

namespace {
// complex / real, component-wise (as GCC divides std::complex<double> by a double)
inline std::complex<double> cdiv(const std::complex<double>& z, const double d)
{
	return std::complex<double>(z.real() / d, z.imag() / d);
}
}  // namespace

ampIntegralMatrix::ampIntegralMatrix() : _nmbEvents(0) {}

// ampIntegralMatrix::setWaveNames (ampIntegralMatrix.cc:255-277)
bool ampIntegralMatrix::setWaveNames(const std::vector<std::string>& waveNames)
{
	if (not _waveNames.empty()) {
		if (waveNames != _waveNames) {
			std::cerr << "ampIntegralMatrix: already initialised with a different wave set" << std::endl;
			return false;
		}
		return true;
	}
	_waveNames = waveNames;
	_waveIndex.clear();
	for (size_t i = 0; i < _waveNames.size(); ++i)
		_waveIndex[_waveNames[i]] = (unsigned int)i;
	if (_waveIndex.size() != _waveNames.size()) {
		std::cerr << "ampIntegralMatrix: duplicate wave names" << std::endl;
		_waveNames.clear();
		_waveIndex.clear();
		return false;
	}
	_integrals.assign(_waveNames.size() * _waveNames.size(), std::complex<double>(0, 0));
	return true;
}

bool ampIntegralMatrix::containsWave(const std::string& waveName) const
{
	return _waveIndex.count(waveName) > 0;
}

unsigned int ampIntegralMatrix::waveIndex(const std::string& waveName) const
{
	std::map<std::string, unsigned int>::const_iterator it = _waveIndex.find(waveName);
	if (it == _waveIndex.end())
		throw std::out_of_range("ampIntegralMatrix: unknown wave '" + waveName + "'");
	return it->second;
}

const std::string& ampIntegralMatrix::waveName(const unsigned int waveIndex) const
{
	if (waveIndex >= _waveNames.size())
		throw std::out_of_range("ampIntegralMatrix: wave index out of range");
	return _waveNames[waveIndex];
}

// element (ampIntegralMatrix.cc:238-252)
std::complex<double> ampIntegralMatrix::element(const unsigned int waveIndexI, const unsigned int waveIndexJ) const
{
	const unsigned int W = nmbWaves();
	if (waveIndexI >= W or waveIndexJ >= W)
		throw std::out_of_range("ampIntegralMatrix::element: wave index out of range");
	return cdiv(_integrals[(size_t)waveIndexI * W + waveIndexJ], (double)_nmbEvents);
}

// integrate (ampIntegralMatrix.cc:300-488): I_ij = sum_e weight_e sum_s A_i^s(e) conj(A_j^s(e)).
// Each incoherent sub-amplitude s is one rank-k update over the same events and
// weights (so one mean-weight renormalisation), added on the host in order s = 0, 1, ...
bool ampIntegralMatrix::integrateImpl(const double* amps, const unsigned long nmbEvents, const double* weights,
                                      const int device, const unsigned int nmbSubAmps, const bool onDevice)
{
	const char* fn = onDevice ? "ampIntegralMatrix::integrateDevice" : "ampIntegralMatrix::integrate";
	if (_waveNames.empty()) {
		std::cerr << fn << ": set the wave names first" << std::endl;
		return false;
	}
	if (nmbSubAmps < 1) {
		std::cerr << fn << ": amplitudes need at least one incoherent sub-amplitude" << std::endl;
		return false;
	}
	const size_t W = nmbWaves();
	std::vector<std::complex<double> > sums(W * W, std::complex<double>(0, 0));
	std::vector<std::complex<double> > part(nmbSubAmps > 1 ? W * W : 0);
	for (unsigned int s = 0; s < nmbSubAmps; ++s) {
		const double* a = amps + (size_t)s * nmbEvents * W * 2;
		double* out = reinterpret_cast<double*>(s == 0 ? sums.data() : part.data());
		const int st = onDevice ? rpwa_integral_device(device, a, nmbEvents, (uint32_t)W, weights, out)
		                        : rpwa_integral(device, a, nmbEvents, (uint32_t)W, weights, out);
		if (st != RPWA_OK) {
			std::cerr << fn << ": " << rpwa_last_error() << std::endl;
			return false;
		}
		if (s > 0)
			for (size_t k = 0; k < W * W; ++k)
				sums[k] += part[k];
	}
	_integrals.swap(sums);
	_nmbEvents = nmbEvents;
	return true;
}

bool ampIntegralMatrix::integrate(const std::complex<double>* amps, const unsigned long nmbEvents,
                                  const double* weights, const int device, const unsigned int nmbSubAmps)
{
	return integrateImpl(reinterpret_cast<const double*>(amps), nmbEvents, weights, device, nmbSubAmps, false);
}

bool ampIntegralMatrix::integrateDevice(const double* d_amps, const unsigned long nmbEvents,
                                        const double* d_weights, const int device, const unsigned int nmbSubAmps)
{
	return integrateImpl(d_amps, nmbEvents, d_weights, device, nmbSubAmps, true);
}

// addEvent (ampIntegralMatrix.cc:280-297)
bool ampIntegralMatrix::addEvent(std::map<std::string, std::complex<double> >& amplitudes)
{
	const unsigned int W = nmbWaves();
	for (unsigned int i = 0; i < W; ++i)
		if (not amplitudes.count(_waveNames[i])) {
			std::cerr << "ampIntegralMatrix::addEvent: wave '" << _waveNames[i] << "' not in amplitudes" << std::endl;
			return false;
		}
	for (unsigned int i = 0; i < W; ++i) {
		const std::complex<double> iAmp = amplitudes[_waveNames[i]];
		for (unsigned int j = 0; j < W; ++j)
			_integrals[(size_t)i * W + j] += iAmp * std::conj(amplitudes[_waveNames[j]]);
	}
	++_nmbEvents;
	return true;
}

// renormalize (ampIntegralMatrix.cc:492-500)
void ampIntegralMatrix::renormalize(const unsigned long nmbEventsRenorm)
{
	*this *= (double)nmbEventsRenorm / _nmbEvents;
	_nmbEvents = nmbEventsRenorm;
}

// operator+= (ampIntegralMatrix.cc:126-143): elements matched by wave name
ampIntegralMatrix& ampIntegralMatrix::operator+=(const ampIntegralMatrix& integral)
{
	if (not hasIdenticalWaveSet(integral))
		throw std::invalid_argument("ampIntegralMatrix: cannot add integrals of different wave sets");
	const unsigned int W = nmbWaves();
	for (unsigned int i = 0; i < W; ++i) {
		const unsigned int oi = integral.waveIndex(_waveNames[i]);
		for (unsigned int j = 0; j < W; ++j)
			_integrals[(size_t)i * W + j] += integral._integrals[(size_t)oi * W + integral.waveIndex(_waveNames[j])];
	}
	_nmbEvents += integral._nmbEvents;
	return *this;
}

ampIntegralMatrix& ampIntegralMatrix::operator-=(const ampIntegralMatrix& integral)
{
	if (not hasIdenticalWaveSet(integral))
		throw std::invalid_argument("ampIntegralMatrix: cannot subtract integrals of different wave sets");
	const unsigned int W = nmbWaves();
	for (unsigned int i = 0; i < W; ++i) {
		const unsigned int oi = integral.waveIndex(_waveNames[i]);
		for (unsigned int j = 0; j < W; ++j)
			_integrals[(size_t)i * W + j] -= integral._integrals[(size_t)oi * W + integral.waveIndex(_waveNames[j])];
	}
	_nmbEvents -= integral._nmbEvents;
	return *this;
}

ampIntegralMatrix& ampIntegralMatrix::operator*=(const double factor)
{
	for (std::complex<double>& v : _integrals)
		v *= factor;
	return *this;
}

ampIntegralMatrix& ampIntegralMatrix::operator/=(const double factor)
{
	for (std::complex<double>& v : _integrals)
		v = cdiv(v, factor);
	return *this;
}

bool ampIntegralMatrix::hasIdenticalWaveSet(const ampIntegralMatrix& integral) const
{
	if (nmbWaves() != integral.nmbWaves())
		return false;
	for (const std::string& n : _waveNames)
		if (not integral.containsWave(n))
			return false;
	return true;
}

bool operator==(const ampIntegralMatrix& lhs, const ampIntegralMatrix& rhs)
{
	if (not lhs.hasIdenticalWaveSet(rhs) or lhs.nmbEvents() != rhs.nmbEvents())
		return false;
	const unsigned int W = lhs.nmbWaves();
	for (unsigned int i = 0; i < W; ++i)
		for (unsigned int j = 0; j < W; ++j)
			if (lhs.matrix()[(size_t)i * W + j] !=
			    rhs.matrix()[(size_t)rhs.waveIndex(lhs.waveName(i)) * W + rhs.waveIndex(lhs.waveName(j))])
				return false;
	return true;
}

// writeAscii / readAscii (ampIntegralMatrix.cc:503-571)
bool ampIntegralMatrix::writeAscii(std::ostream& out) const
{
	if (not out)
		return false;
	const unsigned int W = nmbWaves();
	out << W << std::endl << _nmbEvents << std::endl;
	out << W << " " << W << std::endl;
	out << std::setprecision(std::numeric_limits<double>::max_digits10);
	for (unsigned int i = 0; i < W; ++i) {
		for (unsigned int j = 0; j < W; ++j)
			out << _integrals[(size_t)i * W + j] << "\t";
		out << std::endl;
	}
	out << _waveNames.size() << std::endl;
	for (unsigned int i = 0; i < W; ++i)
		out << _waveNames[i] << " " << i << std::endl;
	return (bool)out;
}

bool ampIntegralMatrix::readAscii(std::istream& in)
{
	unsigned int W;
	unsigned long nEv;
	if (not (in >> W >> nEv))
		return false;
	unsigned int nmbRows, nmbCols;
	if (not (in >> nmbRows >> nmbCols) or nmbRows != W or nmbCols != W)
		return false;
	std::vector<std::complex<double> > vals((size_t)W * W);
	for (size_t k = 0; k < vals.size(); ++k)
		if (not (in >> vals[k]))
			return false;
	unsigned int mapSize;
	if (not (in >> mapSize) or mapSize != W)
		return false;
	std::vector<std::string> names(W);
	for (unsigned int k = 0; k < W; ++k) {
		std::string name;
		unsigned int idx;
		if (not (in >> name >> idx) or idx >= W)
			return false;
		names[idx] = name;
	}
	_waveNames.clear();
	if (not setWaveNames(names))
		return false;
	_integrals = vals;
	_nmbEvents = nEv;
	return true;
}

bool ampIntegralMatrix::writeAscii(const std::string& outFileName) const
{
	std::ofstream f(outFileName.c_str());
	return f.good() and writeAscii(f);
}

bool ampIntegralMatrix::readAscii(const std::string& inFileName)
{
	std::ifstream f(inFileName.c_str());
	return f.good() and readAscii(f);
}

std::ostream& ampIntegralMatrix::print(std::ostream& out, const bool printIntegralValues) const
{
	out << "amplitude integral matrix:" << std::endl
	    << "    number of waves .... " << nmbWaves() << std::endl
	    << "    number of events ... " << _nmbEvents << std::endl
	    << "    wave names:" << std::endl;
	for (size_t i = 0; i < _waveNames.size(); ++i)
		out << "        " << i << " -> " << _waveNames[i] << std::endl;
	if (printIntegralValues) {
		const unsigned int W = nmbWaves();
		for (unsigned int i = 0; i < W; ++i)
			for (unsigned int j = 0; j < W; ++j)
				out << "        [" << std::setw(4) << i << ", " << std::setw(4) << j
				    << "] = " << _integrals[(size_t)i * W + j] << std::endl;
	}
	return out;
}

}  // namespace rpwa
