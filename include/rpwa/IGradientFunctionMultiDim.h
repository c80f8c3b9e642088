// IGradientFunctionMultiDim.h -- the minimiser-facing interface of a likelihood.
//
// Same virtual interface as ROOT::Math::IGradientFunctionMultiDim (the base of
// rpwa::pwaLikelihood, partialWaveFit/pwaLikelihood.h:67), so that a
// Minuit2-style driver (DoEval + Gradient) or an NLopt-style driver (FdF)
// calls the likelihood exactly as it calls the reference.  When ROOT is
// available, a one-line adapter deriving from the ROOT class and forwarding to
// these methods makes the likelihood usable by ROOT's minimisers unchanged.
#pragma once

namespace rpwa {
namespace Math {

class IGradientFunctionMultiDim {
public:
	virtual ~IGradientFunctionMultiDim() {}
	virtual IGradientFunctionMultiDim* Clone() const = 0;
	virtual unsigned int NDim() const = 0;
	double operator()(const double* x) const { return DoEval(x); }
	virtual void Gradient(const double* x, double* grad) const = 0;
	virtual void FdF(const double* x, double& f, double* df) const = 0;
	double Derivative(const double* x, unsigned int icoord) const { return DoDerivative(x, icoord); }

	// public here (private virtuals in ROOT) so that the C ABI can reach them
	virtual double DoEval(const double* x) const = 0;
	virtual double DoDerivative(const double* x, unsigned int icoord) const = 0;
};

}  // namespace Math
}  // namespace rpwa
