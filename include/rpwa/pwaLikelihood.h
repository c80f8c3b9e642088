// pwaLikelihood.h -- unbinned extended log-likelihood of a partial-wave fit,
// evaluated on an MI355X through the rootpwa_amd C ABI.
//
// Drop-in for rpwa::pwaLikelihood<complexT> (partialWaveFit/pwaLikelihood.h:66-323):
// same parameter layout, same method names and call order
// (init -> addNormIntegral -> addAccIntegral -> addAmplitude... -> finishInit ->
// DoEval / Gradient / FdF / DoDerivative / Hessian), same results.  The event
// sums run on the GPU (rpwa_lik_* entry points); the normalisation term, the
// prior, the parameter mapping and the sign convention run on the host with
// the reference's arithmetic.  Differences, all deliberate:
//   - amplitudes arrive as arrays (addAmplitude(name, amps, n) or, for data
//     already in device memory, addAmplitudesDevice) instead of ROOT trees;
//   - the wave list is (name, reflectivity, threshold) instead of a
//     waveDescription whose decay topology yields the reflectivity;
//   - the GPU is the only event-sum backend: there is no CPU event loop, and an
//     evaluation without a usable GPU throws std::runtime_error;
//   - errors in evaluation throw std::logic_error / std::runtime_error where
//     the reference executes `throw;` without an active exception;
//   - complexT selects the amplitude storage on the GPU: std::complex<double>
//     stores complex128, std::complex<float> stores complex64 (rounded after the
//     FP64 normalisation); all arithmetic is FP64 in both;
//   - Clone() shares the device context (reference counted) instead of
//     copying the amplitudes.
#pragma once

#include <chrono>
#include <complex>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ampIntegralMatrix.h"
#include "IGradientFunctionMultiDim.h"

struct rpwa_lik_ctx;

namespace rpwa {

// row-major square matrix (the TMatrixT<double> the reference returns)
struct denseMatrix {
	unsigned int n = 0;
	std::vector<double> data;
	denseMatrix() {}
	explicit denseMatrix(unsigned int size) : n(size), data((size_t)size * size, 0.) {}
	double& operator()(unsigned int i, unsigned int j) { return data[(size_t)i * n + j]; }
	double operator()(unsigned int i, unsigned int j) const { return data[(size_t)i * n + j]; }
};

template <typename complexT>
class pwaLikelihood : public Math::IGradientFunctionMultiDim {
public:
	typedef typename complexT::value_type value_type;

	enum functionCallEnum {
		FDF = 0,
		GRADIENT = 1,
		DOEVAL = 2,
		DODERIVATIVE = 3,
		HESSIAN = 4,
		NMB_FUNCTIONCALLENUM = 5
	};

	struct functionCallInfo {
		unsigned int nmbCalls = 0;
		double funcTime = 0;   // time for the event sums [s]
		double normTime = 0;   // time for the normalisation term [s]
		double totalTime = 0;  // total [s]
	};

	enum priorEnum { FLAT, HALF_CAUCHY };

	// wave name, reflectivity (+1 / -1), mass threshold
	struct waveDescThresType {
		std::string name;
		int reflectivity;
		double threshold;
	};

	class fitParameter {
	public:
		fitParameter() : _waveName(""), _rank(0), _threshold(0.), _fixed(true), _realPart(true) {}
		fitParameter(const std::string& waveName, const unsigned int rank, const double threshold,
		             const bool fixed, const bool realPart)
			: _waveName(waveName), _rank(rank), _threshold(threshold), _fixed(fixed), _realPart(realPart) {}
		bool operator==(const fitParameter& rhs) const
		{
			return _waveName == rhs._waveName and _rank == rhs._rank and _threshold == rhs._threshold
			       and _fixed == rhs._fixed and _realPart == rhs._realPart;
		}
		void fix(const bool fixed) { _fixed = fixed; }
		const std::string& waveName() const { return _waveName; }
		unsigned int rank() const { return _rank; }
		double threshold() const { return _threshold; }
		bool fixed() const { return _fixed; }
		bool realPart() const { return _realPart; }
		std::string parName() const;  // pwaLikelihood.cc:107-122

	private:
		std::string _waveName;
		unsigned int _rank;
		double _threshold;
		bool _fixed;
		bool _realPart;
	};

	static bool _debug;
	static void setQuiet(const bool flag = true) { _debug = not flag; }

	pwaLikelihood();
	pwaLikelihood(const pwaLikelihood& other);
	~pwaLikelihood();

	// IGradientFunctionMultiDim
	pwaLikelihood* Clone() const override { return new pwaLikelihood(*this); }
	unsigned int NDim() const override { return nmbPars(); }
	void Gradient(const double* par, double* gradient) const override;
	void FdF(const double* par, double& funcVal, double* gradient) const override;
	double DoEval(const double* par) const override;
	double DoDerivative(const double* par, const unsigned int derivativeIndex) const override;

	denseMatrix Hessian(const double* par) const;
	denseMatrix CovarianceMatrix(const double* par) const;
	denseMatrix CovarianceMatrix(const denseMatrix& hessian) const;
	// (eigenvector over all P parameters, eigenvalue), descending eigenvalues
	std::vector<std::pair<std::vector<double>, double> > HessianEigenVectors(const denseMatrix& hessian) const;
	std::vector<double> CorrectParamSigns(const double* par) const;

	unsigned int nmbEvents() const { return (unsigned int)_nmbEvents; }
	unsigned long nmbEventsLong() const { return _nmbEvents; }
	unsigned int rank() const { return _rank; }
	unsigned int nmbWaves(const int reflectivity = 0) const;
	unsigned int nmbPars() const { return _nmbPars; }
	unsigned int nmbParsFixed() const { return _nmbParsFixed; }
	const fitParameter& parameter(const unsigned int parIndex) const { return _parameters[parIndex]; }
	const std::vector<fitParameter>& parameters() const { return _parameters; }

	double dLcache(const unsigned int i) const { return _derivCache[i]; }
	unsigned int ncalls(const functionCallEnum func = FDF) const { return _funcCallInfo[func].nmbCalls; }
	double Ltime(const functionCallEnum func = FDF) const { return _funcCallInfo[func].funcTime; }
	double Ntime(const functionCallEnum func = FDF) const { return _funcCallInfo[func].normTime; }
	const functionCallInfo& callInfo(const functionCallEnum func) const { return _funcCallInfo[func]; }

	// modifiers
	void enableCuda(const bool enable = true) { _gpuEnabled = enable; }  // name kept for callers
	bool cudaEnabled() const { return _gpuEnabled; }
	void enableGpu(const bool enable = true) { _gpuEnabled = enable; }
	bool gpuEnabled() const { return _gpuEnabled; }
	void setGpuDevice(const int device) { _device = device; }
	void useNormalizedAmps(const bool useNorm = true) { _useNormalizedAmps = useNorm; }
	bool normalizedAmpsUsed() const { return _useNormalizedAmps; }
	// both change the objective: a value+gradient kept from an earlier DoEval no longer applies
	void setPriorType(const priorEnum priorType = FLAT)
	{
		_priorType = priorType;
		_evalCacheValid = false;
	}
	priorEnum priorType() const { return _priorType; }
	void setCauchyWidth(const double cauchyWidth)
	{
		_cauchyWidth = cauchyWidth;
		_evalCacheValid = false;
	}
	double cauchyWidth() const { return _cauchyWidth; }

	// operations
	bool init(const std::vector<waveDescThresType>& waveDescThres, const unsigned int rank = 1,
	          const double massBinCenter = 0.);
	bool addNormIntegral(const ampIntegralMatrix& normMatrix);
	bool addAccIntegral(ampIntegralMatrix& accMatrix, const unsigned int accEventsOverride = 0);
	// lower and upper edge of a bin in each binning variable
	typedef std::map<std::string, std::pair<double, double> > multibinBoundariesType;
	// on-the-fly binning (pwaLikelihood.cc:1391-1436): keep the events whose
	// binning variables v satisfy lower <= v < upper for every variable of the
	// bin.  eventVariables holds one array per variable over all entries of the
	// event files, concatenated in the order of the amplitude arrays; after this
	// call addAmplitude() takes arrays over all those entries and keeps the
	// selected ones in entry order.  Call before the first addAmplitude().
	bool setOnTheFlyBinning(const multibinBoundariesType& multibinBoundaries,
	                        const std::map<std::string, std::vector<double> >& eventVariables);
	unsigned long nmbEntriesInBin() const { return _entriesInBin.size(); }
	// amplitudes of one wave for all events (host memory)
	bool addAmplitude(const std::string& waveName, const std::complex<double>* amps, const unsigned long nmbEvents);
	// amplitudes of all waves, already in device memory on the GPU device, as
	// [nmbEvents][nmbWaves] complex128: the refl -1 waves, then the refl +1
	// waves, each group in the order of the wave list given to init(); they are
	// normalised on the device at finishInit
	bool addAmplitudesDevice(const double* d_amps, const unsigned long nmbEvents);
	bool finishInit();

	void getIntegralMatrices(std::vector<std::complex<double> >& normMatrix,
	                         std::vector<std::complex<double> >& accMatrix,
	                         std::vector<double>& phaseSpaceIntegral, const bool withFlat = false) const;

	void buildProdAmpArrays(const double* inPar, std::vector<std::complex<double> >& prodAmps,
	                        std::vector<std::pair<int, int> >& parIndices, std::vector<std::string>& prodAmpNames,
	                        const bool withFlat = false) const;

	std::ostream& print(std::ostream& out = std::cout) const;
	std::ostream& printFuncInfo(std::ostream& out = std::cout) const;

	// [rank][2][nmbWavesReflMax] complex, the reference's prodAmpsArrayType layout
	void copyFromParArray(const double* inPar, std::vector<std::complex<double> >& outVal, double& outFlatVal) const;
	void copyToParArray(const std::vector<std::complex<double> >& inVal, const double inFlatVal, double* outPar) const;

	// the device context (shared by clones)
	rpwa_lik_ctx* gpuContext() const { return _ctx.get(); }
	// normalised amplitudes of one reflectivity on the host ([N][W_refl]); only
	// for amplitudes added with addAmplitude
	const std::vector<std::complex<double> >& decayAmps(const unsigned int refl) const { return _decayAmps[refl]; }

private:
	bool readWaveList(const std::vector<waveDescThresType>& waveDescThres);
	bool buildParDataStruct(const unsigned int rank, const double massBinCenter);
	void reorderIntegralMatrix(const ampIntegralMatrix& integral, std::vector<std::complex<double> >& reordered) const;
	void requireFinished() const;
	size_t prodIndex(unsigned int r, unsigned int refl, unsigned int w) const
	{ return ((size_t)r * 2 + refl) * _nmbWavesReflMax + w; }
	size_t matIndex(unsigned int r1, unsigned int w1, unsigned int r2, unsigned int w2) const
	{ return (((size_t)r1 * _nmbWavesReflMax + w1) * 2 + r2) * _nmbWavesReflMax + w2; }

	// host parts of the evaluation (reference arithmetic)
	double normTerm(const std::vector<std::complex<double> >& prodAmps, const double prodAmpFlat,
	                std::vector<std::complex<double> >* normDeriv, double* normDerivFlat) const;
	double priorTerm(const std::vector<std::complex<double> >& prodAmps,
	                 std::vector<std::complex<double> >* priorDeriv) const;
	// event sums on the GPU; deriv may be null (value only)
	void eventSums(const std::vector<std::complex<double> >& prodAmps, const double prodAmpFlat, double* lnl,
	               std::vector<std::complex<double> >* deriv, double* derivFlat) const;
	void fdfImpl(const double* par, double& funcVal, double* gradient, const functionCallEnum who) const;

	unsigned long _nmbEvents;
	unsigned int _rank;
	unsigned int _nmbWaves;
	unsigned int _nmbWavesRefl[2];
	unsigned int _nmbWavesReflMax;
	unsigned int _nmbPars;
	unsigned int _nmbParsFixed;
	bool _initialized;
	bool _normIntAdded;
	bool _accIntAdded;
	bool _initFinished;

	bool _gpuEnabled;
	int _device;
	bool _useNormalizedAmps;
	priorEnum _priorType;
	double _cauchyWidth;

	unsigned long _numbAccEvents;
	double _totAcc;

	std::vector<std::string> _waveNames[2];
	std::vector<double> _waveThresholds[2];
	std::vector<bool> _waveAmpAdded[2];
	std::vector<fitParameter> _parameters;
	std::map<std::string, std::pair<unsigned int, unsigned int> > _waveParams;
	std::vector<std::pair<int, int> > _prodAmpToFuncParMap;  // [rank][2][Wmax]

	bool _onTheFlyBinning;
	unsigned long _nmbEntriesTotal;             // entries of the event files (on-the-fly binning)
	std::vector<unsigned long> _entriesInBin;   // selected entries, ascending

	std::vector<std::complex<double> > _decayAmps[2];  // host copy [N][W_refl] (addAmplitude path)
	const double* _deviceAmps;                         // addAmplitudesDevice path (not owned)

	mutable std::vector<double> _parCache;
	mutable std::vector<double> _derivCache;
	// value+gradient of the last DoEval, reused by a Gradient call at the same point
	mutable std::vector<double> _evalParCache;
	mutable std::vector<double> _evalGradCache;
	mutable bool _evalCacheValid;

	std::vector<std::complex<double> > _normMatrix;  // [2][Wmax][2][Wmax]
	std::vector<std::complex<double> > _accMatrix;
	std::vector<double> _phaseSpaceIntegral;         // [2][Wmax]

	mutable functionCallInfo _funcCallInfo[NMB_FUNCTIONCALLENUM];

	std::shared_ptr<rpwa_lik_ctx> _ctx;
};

}  // namespace rpwa
