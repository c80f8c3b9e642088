// pwalik_dropin.cc -- a C++ consumer of the drop-in rpwa::pwaLikelihood, used
// the way ROOTPWA's drivers use the reference class (pwaFit.cc:150-300,
// pwaNloptFit.cc:150-260): init, addNormIntegral, addAccIntegral,
// addAmplitude per wave, finishInit, then FdF / DoEval / Gradient /
// DoDerivative through the IGradientFunctionMultiDim interface and a Clone().
//
// usage: pwalik_dropin <input file>      evaluate the points of the file
//        pwalik_dropin --host-only       set-up checks that need no GPU
//
// Input (whitespace-separated text): nW0 nW1 rank nEvents nMC complex64,
// the integral sums [W][W] (re im), the amplitudes [nEvents][W] (re im, refl -1
// waves first), the number of points, then each point's parameters.  Output:
// for each point one line "value g_0 g_1 ..." (17 significant digits) and a
// line "bitwise <0|1>" for the DoEval == FdF value, Gradient == FdF gradient
// and Clone contracts.
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rpwa/pwaLikelihood.h"

namespace {

template <typename complexT>
std::vector<typename rpwa::pwaLikelihood<complexT>::waveDescThresType> waves(unsigned nW0, unsigned nW1)
{
	std::vector<typename rpwa::pwaLikelihood<complexT>::waveDescThresType> w;
	char name[32];
	for (unsigned i = 0; i < nW0; ++i) {
		std::snprintf(name, sizeof(name), "neg%02u", i);
		w.push_back({name, -1, 0.});
	}
	for (unsigned i = 0; i < nW1; ++i) {
		std::snprintf(name, sizeof(name), "pos%02u", i);
		w.push_back({name, +1, 0.});
	}
	return w;
}

int host_only()
{
	typedef std::complex<double> cd;
	rpwa::pwaLikelihood<cd>::setQuiet();
	rpwa::pwaLikelihood<cd> L;
	if (!L.init(waves<cd>(1, 2), 2, 0.))
		return 2;
	for (unsigned i = 0; i < L.nmbPars(); ++i)
		std::cout << L.parameter(i).parName() << (i + 1 < L.nmbPars() ? " " : "\n");
	std::vector<double> par(L.nmbPars(), 1.);
	try {
		L.DoEval(par.data());
		std::cout << "evaluation before finishInit did not throw\n";
		return 3;
	} catch (const std::logic_error&) {
		std::cout << "evaluation before finishInit throws\n";
	}
	return 0;
}

template <typename complexT>
int evaluate(std::istream& in, unsigned nW0, unsigned nW1, unsigned rank, unsigned long nEvents, unsigned long nMC)
{
	const unsigned W = nW0 + nW1;
	rpwa::pwaLikelihood<complexT>::setQuiet();
	rpwa::pwaLikelihood<complexT> L;
	const auto wl = waves<complexT>(nW0, nW1);
	std::vector<std::string> names;
	for (const auto& w : wl)
		names.push_back(w.name);
	rpwa::ampIntegralMatrix I;
	I.setWaveNames(names);
	for (unsigned k = 0; k < W * W; ++k) {
		double re, im;
		in >> re >> im;
		I.matrix()[k] = std::complex<double>(re, im);
	}
	I.setNmbEvents(nMC);
	std::vector<std::complex<double> > amps((size_t)nEvents * W);
	for (auto& a : amps) {
		double re, im;
		in >> re >> im;
		a = std::complex<double>(re, im);
	}
	if (!L.init(wl, rank, 0.) || !L.addNormIntegral(I) || !L.addAccIntegral(I))
		return 2;
	std::vector<std::complex<double> > column(nEvents);
	for (unsigned k = 0; k < W; ++k) {
		for (unsigned long e = 0; e < nEvents; ++e)
			column[e] = amps[(size_t)e * W + k];
		if (!L.addAmplitude(names[k], column.data(), nEvents))
			return 2;
	}
	if (!L.finishInit())
		return 2;
	std::unique_ptr<rpwa::Math::IGradientFunctionMultiDim> f(L.Clone());
	const unsigned P = f->NDim();
	unsigned nPoints = 0;
	if (!(in >> nPoints)) {
		std::cerr << "pwalik_dropin: malformed input\n";
		return 1;
	}
	std::vector<double> par(P), g(P), g2(P);
	for (unsigned n = 0; n < nPoints; ++n) {
		for (auto& x : par)
			in >> x;
		if (!in) {
			std::cerr << "pwalik_dropin: malformed input\n";
			return 1;
		}
		double v;
		f->FdF(par.data(), v, g.data());
		bool bitwise = f->DoEval(par.data()) == v && L.DoEval(par.data()) == v;
		f->Gradient(par.data(), g2.data());
		bitwise = bitwise && g2 == g && f->DoDerivative(par.data(), P - 1) == g[P - 1];
		std::printf("%.17g", v);
		for (double x : g)
			std::printf(" %.17g", x);
		std::printf("\nbitwise %d\n", bitwise ? 1 : 0);
	}
	return 0;
}

}  // namespace

int main(int argc, char** argv)
{
	if (argc < 2) {
		std::cerr << "usage: pwalik_dropin <input file> | --host-only\n";
		return 1;
	}
	if (std::string(argv[1]) == "--host-only")
		return host_only();
	std::ifstream in(argv[1]);
	unsigned nW0, nW1, rank, c64;
	unsigned long nEvents, nMC;
	if (!(in >> nW0 >> nW1 >> rank >> nEvents >> nMC >> c64))
		return 1;
	try {
		return c64 ? evaluate<std::complex<float> >(in, nW0, nW1, rank, nEvents, nMC)
		           : evaluate<std::complex<double> >(in, nW0, nW1, rank, nEvents, nMC);
	} catch (const std::exception& e) {
		std::cerr << "pwalik_dropin: " << e.what() << "\n";
		return 4;
	}
}
