// ampIntegralMatrix.h -- normalisation integral matrix I_ij = sum_e A_i(e) conj(A_j(e)).
//
// Mirrors the interface of rpwa::ampIntegralMatrix (decayAmplitude/ampIntegralMatrix.h:67-172)
// for amplitudes given as arrays: the sums are accumulated on the GPU
// (FP64 MFMA, rpwa_integral*) instead of event by event from ROOT trees.
// element(i, j) returns the sum divided by the number of events
// (ampIntegralMatrix.cc:238-252).
#pragma once

#include <complex>
#include <iosfwd>
#include <map>
#include <string>
#include <vector>

namespace rpwa {

class ampIntegralMatrix {
public:
	ampIntegralMatrix();

	bool setWaveNames(const std::vector<std::string>& waveNames);

	unsigned int nmbWaves() const { return (unsigned int)_waveNames.size(); }
	unsigned long nmbEvents() const { return _nmbEvents; }
	void setNmbEvents(const unsigned long nmbEvents) { _nmbEvents = nmbEvents; }

	bool containsWave(const std::string& waveName) const;
	unsigned int waveIndex(const std::string& waveName) const;  // throws if absent
	const std::string& waveName(const unsigned int waveIndex) const;
	const std::vector<std::string>& waveNames() const { return _waveNames; }

	// sums, row-major [W][W]
	const std::vector<std::complex<double> >& matrix() const { return _integrals; }
	std::vector<std::complex<double> >& matrix() { return _integrals; }

	std::complex<double> element(const unsigned int waveIndexI, const unsigned int waveIndexJ) const;
	std::complex<double> element(const std::string& waveNameI, const std::string& waveNameJ) const
	{ return element(waveIndex(waveNameI), waveIndex(waveNameJ)); }

	// ampIntegralMatrix::integrate: amps [nmbSubAmps][nmbEvents][nmbWaves]
	// complex (host or device memory), the incoherent sub-amplitudes of every
	// wave (ampIntegralMatrix.cc:432-466), importance weights w per event (the
	// integral uses 1/w) or null.  Replaces the sums and the event count.
	bool integrate(const std::complex<double>* amps, const unsigned long nmbEvents,
	               const double* weights = 0, const int device = 0, const unsigned int nmbSubAmps = 1);
	bool integrateDevice(const double* d_amps, const unsigned long nmbEvents,
	                     const double* d_weights = 0, const int device = 0, const unsigned int nmbSubAmps = 1);

	// ampIntegralMatrix::addEvent (ampIntegralMatrix.cc:280-297), on the host
	bool addEvent(std::map<std::string, std::complex<double> >& amplitudes);

	void renormalize(const unsigned long nmbEventsRenorm);

	ampIntegralMatrix& operator+=(const ampIntegralMatrix& integral);
	ampIntegralMatrix& operator-=(const ampIntegralMatrix& integral);
	ampIntegralMatrix& operator*=(const double factor);
	ampIntegralMatrix& operator/=(const double factor);

	bool hasIdenticalWaveSet(const ampIntegralMatrix& integral) const;

	// ASCII format of ampIntegralMatrix::writeAscii/readAscii
	// (ampIntegralMatrix.cc:503-571); the wave-name map is written to the
	// stream itself.
	bool writeAscii(std::ostream& out) const;
	bool readAscii(std::istream& in);
	bool writeAscii(const std::string& outFileName) const;
	bool readAscii(const std::string& inFileName);

	std::ostream& print(std::ostream& out, const bool printIntegralValues = false) const;

private:
	bool integrateImpl(const double* amps, const unsigned long nmbEvents, const double* weights, const int device,
	                   const unsigned int nmbSubAmps, const bool onDevice);
	std::vector<std::string> _waveNames;
	std::map<std::string, unsigned int> _waveIndex;
	unsigned long _nmbEvents;
	std::vector<std::complex<double> > _integrals;
};

bool operator==(const ampIntegralMatrix& lhs, const ampIntegralMatrix& rhs);
inline bool operator!=(const ampIntegralMatrix& lhs, const ampIntegralMatrix& rhs) { return not (lhs == rhs); }

}  // namespace rpwa
