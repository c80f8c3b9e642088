// pwalik_capi.cc -- C ABI over rpwa::pwaLikelihood / rpwa::ampIntegralMatrix
// (declared in include/rootpwa_amd_pwalik.h).
#include <complex>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "rootpwa_amd_pwalik.h"
#include "rpwa/fitDrivers.h"
#include "rpwa/pwaLikelihood.h"

namespace rpwa_amd {
void set_error(const std::string& msg);
void clear_error();
}  // namespace rpwa_amd

using rpwa_amd::clear_error;
using rpwa_amd::set_error;

struct rpwa_intmat {
	rpwa::ampIntegralMatrix m;
};

struct rpwa_pwalik {
	std::unique_ptr<rpwa::pwaLikelihood<std::complex<double> > > d;
	std::unique_ptr<rpwa::pwaLikelihood<std::complex<float> > > f;
};

#define GUARD_BEGIN                \
	clear_error();                 \
	try {
#define GUARD_END                                        \
	}                                                    \
	catch (const std::exception& ex) {                   \
		set_error(ex.what());                            \
		return RPWA_ERR_INVALID;                         \
	}                                                    \
	catch (...) {                                        \
		set_error("unknown exception");                  \
		return RPWA_ERR_INVALID;                         \
	}

// run `body` with `lik` bound to whichever instantiation the handle holds
#define WITH_LIK(L, ...)               \
	do {                               \
		if ((L)->d) {                  \
			auto& lik = *(L)->d;       \
			__VA_ARGS__;               \
		} else {                       \
			auto& lik = *(L)->f;       \
			__VA_ARGS__;               \
		}                              \
	} while (0)

static int need(const void* p, const char* fn)
{
	if (!p) {
		set_error(std::string(fn) + ": null pointer");
		return RPWA_ERR_INVALID;
	}
	return RPWA_OK;
}

static int fail(const char* fn, const char* what)
{
	set_error(std::string(fn) + ": " + what);
	return RPWA_ERR_INVALID;
}

// ---- ampIntegralMatrix ------------------------------------------------------
extern "C" int rpwa_intmat_create(rpwa_intmat** m)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_create"))
		return st;
	*m = new rpwa_intmat();
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_destroy(rpwa_intmat* m)
{
	clear_error();
	delete m;
	return RPWA_OK;
}

extern "C" int rpwa_intmat_set_wave_names(rpwa_intmat* m, uint32_t nmbWaves, const char* const* names)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_set_wave_names"))
		return st;
	if (int st = need(names, "rpwa_intmat_set_wave_names"))
		return st;
	std::vector<std::string> n(names, names + nmbWaves);
	if (not m->m.setWaveNames(n))
		return fail("rpwa_intmat_set_wave_names", "wave names rejected (duplicates or a different set)");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_nmb_waves(const rpwa_intmat* m, uint32_t* nmbWaves)
{
	GUARD_BEGIN
	if (!m || !nmbWaves)
		return fail("rpwa_intmat_nmb_waves", "null pointer");
	*nmbWaves = m->m.nmbWaves();
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_wave_name(const rpwa_intmat* m, uint32_t i, char* name, int nameLen)
{
	GUARD_BEGIN
	if (!m || !name || nameLen <= 0)
		return fail("rpwa_intmat_wave_name", "null pointer");
	if (i >= m->m.nmbWaves())
		return fail("rpwa_intmat_wave_name", "index out of range");
	std::strncpy(name, m->m.waveName(i).c_str(), (size_t)nameLen - 1);
	name[nameLen - 1] = 0;
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_set_sums(rpwa_intmat* m, const double* sums, uint64_t nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_set_sums"))
		return st;
	if (int st = need(sums, "rpwa_intmat_set_sums"))
		return st;
	std::vector<std::complex<double> >& v = m->m.matrix();
	std::memcpy(v.data(), sums, v.size() * sizeof(std::complex<double>));
	m->m.setNmbEvents(nmbEvents);
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_get_sums(const rpwa_intmat* m, double* sums, uint64_t* nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_get_sums"))
		return st;
	if (sums)
		std::memcpy(sums, m->m.matrix().data(), m->m.matrix().size() * sizeof(std::complex<double>));
	if (nmbEvents)
		*nmbEvents = m->m.nmbEvents();
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_set_nmb_events(rpwa_intmat* m, uint64_t nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_set_nmb_events"))
		return st;
	m->m.setNmbEvents(nmbEvents);
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_element(const rpwa_intmat* m, uint32_t i, uint32_t j, double* re, double* im)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_element"))
		return st;
	const std::complex<double> v = m->m.element(i, j);
	if (re)
		*re = v.real();
	if (im)
		*im = v.imag();
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_integrate(rpwa_intmat* m, const double* amps, uint64_t nmbEvents, const double* weights,
                                     int device, int onDevice)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_integrate"))
		return st;
	const bool ok = onDevice ? m->m.integrateDevice(amps, nmbEvents, weights, device)
	                         : m->m.integrate(reinterpret_cast<const std::complex<double>*>(amps), nmbEvents,
	                                          weights, device);
	if (not ok) {
		const std::string prev = rpwa_last_error();
		set_error("rpwa_intmat_integrate failed" + (prev.empty() ? std::string() : ": " + prev));
		return RPWA_ERR_INVALID;
	}
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_integrate_incoherent(rpwa_intmat* m, const double* amps, uint32_t nmbSubAmps,
                                                uint64_t nmbEvents, const double* weights, int device, int onDevice)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_integrate_incoherent"))
		return st;
	if (nmbSubAmps < 1)
		return fail("rpwa_intmat_integrate_incoherent", "at least one sub-amplitude is needed");
	const bool ok = onDevice ? m->m.integrateDevice(amps, nmbEvents, weights, device, nmbSubAmps)
	                         : m->m.integrate(reinterpret_cast<const std::complex<double>*>(amps), nmbEvents,
	                                          weights, device, nmbSubAmps);
	if (not ok) {
		const std::string prev = rpwa_last_error();
		set_error("rpwa_intmat_integrate_incoherent failed" + (prev.empty() ? std::string() : ": " + prev));
		return RPWA_ERR_INVALID;
	}
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_add(rpwa_intmat* m, const rpwa_intmat* other)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_add"))
		return st;
	if (int st = need(other, "rpwa_intmat_add"))
		return st;
	m->m += other->m;
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_renormalize(rpwa_intmat* m, uint64_t nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_renormalize"))
		return st;
	m->m.renormalize(nmbEvents);
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_write_ascii(const rpwa_intmat* m, const char* fileName)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_write_ascii"))
		return st;
	if (not m->m.writeAscii(std::string(fileName ? fileName : "")))
		return fail("rpwa_intmat_write_ascii", "cannot write file");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_intmat_read_ascii(rpwa_intmat* m, const char* fileName)
{
	GUARD_BEGIN
	if (int st = need(m, "rpwa_intmat_read_ascii"))
		return st;
	if (not m->m.readAscii(std::string(fileName ? fileName : "")))
		return fail("rpwa_intmat_read_ascii", "cannot read file or malformed content");
	return RPWA_OK;
	GUARD_END
}

// ---- pwaLikelihood ------------------------------------------------------------
extern "C" int rpwa_pwalik_create(rpwa_pwalik** L, int precision)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_create"))
		return st;
	rpwa_pwalik* h = new rpwa_pwalik();
	if (precision == RPWA_PREC_MIXED)
		h->f.reset(new rpwa::pwaLikelihood<std::complex<float> >());
	else if (precision == RPWA_PREC_FP64)
		h->d.reset(new rpwa::pwaLikelihood<std::complex<double> >());
	else {
		delete h;
		return fail("rpwa_pwalik_create", "unknown precision");
	}
	*L = h;
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_clone(const rpwa_pwalik* L, rpwa_pwalik** out)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_clone"))
		return st;
	if (int st = need(out, "rpwa_pwalik_clone"))
		return st;
	rpwa_pwalik* h = new rpwa_pwalik();
	if (L->d)
		h->d.reset(L->d->Clone());
	else
		h->f.reset(L->f->Clone());
	*out = h;
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_destroy(rpwa_pwalik* L)
{
	clear_error();
	delete L;
	return RPWA_OK;
}

extern "C" int rpwa_pwalik_set_device(rpwa_pwalik* L, int device)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_set_device"))
		return st;
	WITH_LIK(L, lik.setGpuDevice(device));
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_init(rpwa_pwalik* L, uint32_t nmbWaves, const char* const* names,
                                const int* reflectivities, const double* thresholds, uint32_t rank,
                                double massBinCenter)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_init"))
		return st;
	if (!names || !reflectivities)
		return fail("rpwa_pwalik_init", "null wave list");
	bool ok = false;
	WITH_LIK(L, {
		typedef typename std::remove_reference<decltype(lik)>::type likT;
		std::vector<typename likT::waveDescThresType> waves;
		for (uint32_t i = 0; i < nmbWaves; ++i)
			waves.push_back({std::string(names[i]), reflectivities[i], thresholds ? thresholds[i] : 0.});
		ok = lik.init(waves, rank, massBinCenter);
	});
	if (!ok)
		return fail("rpwa_pwalik_init", "invalid wave list or rank");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_use_normalized_amps(rpwa_pwalik* L, int flag)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_use_normalized_amps"))
		return st;
	WITH_LIK(L, lik.useNormalizedAmps(flag != 0));
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_set_prior(rpwa_pwalik* L, int halfCauchy, double cauchyWidth)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_set_prior"))
		return st;
	WITH_LIK(L, {
		typedef typename std::remove_reference<decltype(lik)>::type likT;
		lik.setPriorType(halfCauchy ? likT::HALF_CAUCHY : likT::FLAT);
		lik.setCauchyWidth(cauchyWidth);
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_add_norm_integral(rpwa_pwalik* L, const rpwa_intmat* m)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_add_norm_integral"))
		return st;
	if (int st = need(m, "rpwa_pwalik_add_norm_integral"))
		return st;
	bool ok = false;
	WITH_LIK(L, ok = lik.addNormIntegral(m->m));
	if (!ok)
		return fail("rpwa_pwalik_add_norm_integral", "rejected (call order or wave set)");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_add_acc_integral(rpwa_pwalik* L, rpwa_intmat* m, uint32_t accEventsOverride)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_add_acc_integral"))
		return st;
	if (int st = need(m, "rpwa_pwalik_add_acc_integral"))
		return st;
	bool ok = false;
	WITH_LIK(L, ok = lik.addAccIntegral(m->m, accEventsOverride));
	if (!ok)
		return fail("rpwa_pwalik_add_acc_integral", "rejected (call order or wave set)");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_add_amplitude(rpwa_pwalik* L, const char* waveName, const double* amps,
                                         uint64_t nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_add_amplitude"))
		return st;
	if (!waveName || !amps)
		return fail("rpwa_pwalik_add_amplitude", "null input");
	bool ok = false;
	WITH_LIK(L, ok = lik.addAmplitude(waveName, reinterpret_cast<const std::complex<double>*>(amps), nmbEvents));
	if (!ok)
		return fail("rpwa_pwalik_add_amplitude", "rejected (call order, unknown wave, event count or zero integral)");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_add_amplitudes_device(rpwa_pwalik* L, const double* d_amps, uint64_t nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_add_amplitudes_device"))
		return st;
	bool ok = false;
	WITH_LIK(L, ok = lik.addAmplitudesDevice(d_amps, nmbEvents));
	if (!ok)
		return fail("rpwa_pwalik_add_amplitudes_device", "rejected (call order, empty input or zero integral)");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_finish_init(rpwa_pwalik* L)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_finish_init"))
		return st;
	bool ok = false;
	WITH_LIK(L, ok = lik.finishInit());
	if (!ok) {
		const std::string prev = rpwa_last_error();
		return fail("rpwa_pwalik_finish_init", prev.empty() ? "missing amplitudes" : prev.c_str());
	}
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_nmb_pars(const rpwa_pwalik* L, uint32_t* nmbPars, uint32_t* nmbParsFixed)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_nmb_pars"))
		return st;
	WITH_LIK(L, {
		if (nmbPars)
			*nmbPars = lik.nmbPars();
		if (nmbParsFixed)
			*nmbParsFixed = lik.nmbParsFixed();
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_nmb_events(const rpwa_pwalik* L, uint64_t* nmbEvents)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_nmb_events"))
		return st;
	WITH_LIK(L, *nmbEvents = lik.nmbEventsLong());
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_nmb_waves(const rpwa_pwalik* L, int reflectivity, uint32_t* nmbWaves)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_nmb_waves"))
		return st;
	WITH_LIK(L, *nmbWaves = lik.nmbWaves(reflectivity));
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_parameter(const rpwa_pwalik* L, uint32_t i, char* name, int nameLen, uint32_t* rank,
                                     int* fixed, int* realPart)
{
	GUARD_BEGIN
	if (int st = need(L, "rpwa_pwalik_parameter"))
		return st;
	int st = RPWA_OK;
	WITH_LIK(L, {
		if (i >= lik.nmbPars()) {
			st = fail("rpwa_pwalik_parameter", "index out of range");
		} else {
			const auto& p = lik.parameter(i);
			if (name && nameLen > 0) {
				std::strncpy(name, p.parName().c_str(), (size_t)nameLen - 1);
				name[nameLen - 1] = 0;
			}
			if (rank)
				*rank = p.rank();
			if (fixed)
				*fixed = p.fixed() ? 1 : 0;
			if (realPart)
				*realPart = p.realPart() ? 1 : 0;
		}
	});
	return st;
	GUARD_END
}

#define EVAL_GUARD_END                                   \
	}                                                    \
	catch (const std::logic_error& ex) {                 \
		set_error(ex.what());                            \
		return RPWA_ERR_STATE;                           \
	}                                                    \
	catch (const std::exception& ex) {                   \
		set_error(ex.what());                            \
		return RPWA_ERR_HIP;                             \
	}

extern "C" int rpwa_pwalik_doeval(const rpwa_pwalik* L, const double* par, double* value)
{
	GUARD_BEGIN
	if (!L || !par || !value)
		return fail("rpwa_pwalik_doeval", "null pointer");
	WITH_LIK(L, *value = lik.DoEval(par));
	return RPWA_OK;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_gradient(const rpwa_pwalik* L, const double* par, double* gradient)
{
	GUARD_BEGIN
	if (!L || !par || !gradient)
		return fail("rpwa_pwalik_gradient", "null pointer");
	WITH_LIK(L, lik.Gradient(par, gradient));
	return RPWA_OK;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_fdf(const rpwa_pwalik* L, const double* par, double* value, double* gradient)
{
	GUARD_BEGIN
	if (!L || !par || !value || !gradient)
		return fail("rpwa_pwalik_fdf", "null pointer");
	WITH_LIK(L, lik.FdF(par, *value, gradient));
	return RPWA_OK;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_time_fdf(const rpwa_pwalik* L, const double* par, uint32_t n, double* value,
                                    double* gradient, double* seconds)
{
	GUARD_BEGIN
	if (!L || !par || !value || !gradient || !seconds || n == 0)
		return fail("rpwa_pwalik_time_fdf", "null pointer or no call");
	WITH_LIK(L, {
		const auto t0 = std::chrono::steady_clock::now();
		for (uint32_t i = 0; i < n; ++i)
			lik.FdF(par, *value, gradient);
		*seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	});
	return RPWA_OK;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_doderivative(const rpwa_pwalik* L, const double* par, uint32_t k, double* value)
{
	GUARD_BEGIN
	if (!L || !par || !value)
		return fail("rpwa_pwalik_doderivative", "null pointer");
	int st = RPWA_OK;
	WITH_LIK(L, {
		if (k >= lik.nmbPars())
			st = fail("rpwa_pwalik_doderivative", "index out of range");
		else
			*value = lik.DoDerivative(par, k);
	});
	return st;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_hessian(const rpwa_pwalik* L, const double* par, double* hessian)
{
	GUARD_BEGIN
	if (!L || !par || !hessian)
		return fail("rpwa_pwalik_hessian", "null pointer");
	WITH_LIK(L, {
		const rpwa::denseMatrix H = lik.Hessian(par);
		std::memcpy(hessian, H.data.data(), H.data.size() * sizeof(double));
	});
	return RPWA_OK;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_covariance(const rpwa_pwalik* L, const double* par, double* covariance)
{
	GUARD_BEGIN
	if (!L || !par || !covariance)
		return fail("rpwa_pwalik_covariance", "null pointer");
	WITH_LIK(L, {
		const rpwa::denseMatrix C = lik.CovarianceMatrix(par);
		std::memcpy(covariance, C.data.data(), C.data.size() * sizeof(double));
	});
	return RPWA_OK;
	EVAL_GUARD_END
}

extern "C" int rpwa_pwalik_correct_param_signs(const rpwa_pwalik* L, const double* par, double* out)
{
	GUARD_BEGIN
	if (!L || !par || !out)
		return fail("rpwa_pwalik_correct_param_signs", "null pointer");
	WITH_LIK(L, {
		const std::vector<double> v = lik.CorrectParamSigns(par);
		std::memcpy(out, v.data(), v.size() * sizeof(double));
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_copy_from_par(const rpwa_pwalik* L, const double* par, double* prodAmps, double* flat)
{
	GUARD_BEGIN
	if (!L || !par || !prodAmps || !flat)
		return fail("rpwa_pwalik_copy_from_par", "null pointer");
	WITH_LIK(L, {
		std::vector<std::complex<double> > v;
		lik.copyFromParArray(par, v, *flat);
		std::memcpy(prodAmps, v.data(), v.size() * sizeof(std::complex<double>));
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_copy_to_par(const rpwa_pwalik* L, const double* prodAmps, double flat, double* par)
{
	GUARD_BEGIN
	if (!L || !par || !prodAmps)
		return fail("rpwa_pwalik_copy_to_par", "null pointer");
	WITH_LIK(L, {
		const size_t n = (size_t)lik.rank() * 2 * std::max(lik.nmbWaves(-1), lik.nmbWaves(+1));
		const std::complex<double>* p = reinterpret_cast<const std::complex<double>*>(prodAmps);
		std::vector<std::complex<double> > v(p, p + n);
		lik.copyToParArray(v, flat, par);
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_integral_matrices(const rpwa_pwalik* L, int withFlat, double* normMatrix,
                                             double* accMatrix, double* phaseSpaceIntegral)
{
	GUARD_BEGIN
	if (!L)
		return fail("rpwa_pwalik_integral_matrices", "null pointer");
	WITH_LIK(L, {
		std::vector<std::complex<double> > n, a;
		std::vector<double> ps;
		lik.getIntegralMatrices(n, a, ps, withFlat != 0);
		if (normMatrix)
			std::memcpy(normMatrix, n.data(), n.size() * sizeof(std::complex<double>));
		if (accMatrix)
			std::memcpy(accMatrix, a.data(), a.size() * sizeof(std::complex<double>));
		if (phaseSpaceIntegral)
			std::memcpy(phaseSpaceIntegral, ps.data(), ps.size() * sizeof(double));
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_set_quiet(int flag)
{
	clear_error();
	rpwa::pwaLikelihood<std::complex<double> >::setQuiet(flag != 0);
	rpwa::pwaLikelihood<std::complex<float> >::setQuiet(flag != 0);
	return RPWA_OK;
}

extern "C" int rpwa_pwalik_enable_gpu(rpwa_pwalik* L, int flag)
{
	GUARD_BEGIN
	if (!L)
		return fail("rpwa_pwalik_enable_gpu", "null pointer");
	WITH_LIK(L, lik.enableCuda(flag != 0));
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_settings(const rpwa_pwalik* L, int* useNormalized, int* halfCauchy, double* cauchyWidth,
                                    int* gpuEnabled)
{
	GUARD_BEGIN
	if (!L)
		return fail("rpwa_pwalik_settings", "null pointer");
	WITH_LIK(L, {
		typedef typename std::remove_reference<decltype(lik)>::type likT;
		if (useNormalized)
			*useNormalized = lik.normalizedAmpsUsed() ? 1 : 0;
		if (halfCauchy)
			*halfCauchy = lik.priorType() == likT::HALF_CAUCHY ? 1 : 0;
		if (cauchyWidth)
			*cauchyWidth = lik.cauchyWidth();
		if (gpuEnabled)
			*gpuEnabled = lik.cudaEnabled() ? 1 : 0;
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_set_on_the_fly_binning(rpwa_pwalik* L, uint32_t nmbVars, const char* const* varNames,
                                                  const double* lower, const double* upper, uint64_t nmbEntries,
                                                  const double* values)
{
	GUARD_BEGIN
	if (!L || (nmbVars > 0 && (!varNames || !lower || !upper || !values)))
		return fail("rpwa_pwalik_set_on_the_fly_binning", "null pointer");
	std::map<std::string, std::pair<double, double> > bounds;
	std::map<std::string, std::vector<double> > vars;
	for (uint32_t k = 0; k < nmbVars; ++k) {
		if (!varNames[k])
			return fail("rpwa_pwalik_set_on_the_fly_binning", "null variable name");
		const std::string name(varNames[k]);
		if (bounds.count(name))
			return fail("rpwa_pwalik_set_on_the_fly_binning", "binning variable given twice");
		bounds[name] = std::make_pair(lower[k], upper[k]);
		vars[name].assign(values + (size_t)k * nmbEntries, values + (size_t)(k + 1) * nmbEntries);
	}
	if (nmbVars == 0)  // an empty bin definition keeps every entry
		vars[""].assign((size_t)nmbEntries, 0.);
	bool ok = false;
	WITH_LIK(L, ok = lik.setOnTheFlyBinning(bounds, vars));
	if (!ok)
		return fail("rpwa_pwalik_set_on_the_fly_binning", "rejected (amplitudes already added?)");
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_nmb_entries_in_bin(const rpwa_pwalik* L, uint64_t* nmbEntries)
{
	GUARD_BEGIN
	if (!L || !nmbEntries)
		return fail("rpwa_pwalik_nmb_entries_in_bin", "null pointer");
	WITH_LIK(L, *nmbEntries = lik.nmbEntriesInBin());
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_covariance_from_hessian(const rpwa_pwalik* L, const double* hessian, double* covariance)
{
	GUARD_BEGIN
	if (!L || !hessian || !covariance)
		return fail("rpwa_pwalik_covariance_from_hessian", "null pointer");
	WITH_LIK(L, {
		const unsigned int P = lik.nmbPars();
		rpwa::denseMatrix h(P);
		std::memcpy(h.data.data(), hessian, (size_t)P * P * sizeof(double));
		const rpwa::denseMatrix C = lik.CovarianceMatrix(h);
		std::memcpy(covariance, C.data.data(), C.data.size() * sizeof(double));
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_build_prod_amp_arrays(const rpwa_pwalik* L, const double* par, int withFlat,
                                                 double* prodAmps, int32_t* parIndices, uint32_t* n, char* names,
                                                 uint64_t namesLen)
{
	GUARD_BEGIN
	if (!L || !par || !n)
		return fail("rpwa_pwalik_build_prod_amp_arrays", "null pointer");
	std::vector<std::complex<double> > amps;
	std::vector<std::pair<int, int> > idx;
	std::vector<std::string> nm;
	WITH_LIK(L, lik.buildProdAmpArrays(par, amps, idx, nm, withFlat != 0));
	std::string joined;
	for (size_t k = 0; k < nm.size(); ++k)
		joined += (k ? "\n" : "") + nm[k];
	if (names && joined.size() + 1 > namesLen)
		return fail("rpwa_pwalik_build_prod_amp_arrays", "name buffer too small");
	*n = (uint32_t)amps.size();
	for (size_t k = 0; k < amps.size(); ++k) {
		if (prodAmps) {
			prodAmps[2 * k] = amps[k].real();
			prodAmps[2 * k + 1] = amps[k].imag();
		}
		if (parIndices) {
			parIndices[2 * k] = idx[k].first;
			parIndices[2 * k + 1] = idx[k].second;
		}
	}
	if (names)
		std::memcpy(names, joined.c_str(), joined.size() + 1);
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_print(const rpwa_pwalik* L, int funcInfo, char* text, uint64_t textLen, uint64_t* needed)
{
	GUARD_BEGIN
	if (!L)
		return fail("rpwa_pwalik_print", "null pointer");
	std::ostringstream out;
	WITH_LIK(L, {
		if (funcInfo)
			lik.printFuncInfo(out);
		else
			lik.print(out);
	});
	const std::string s = out.str();
	if (needed)
		*needed = s.size() + 1;
	if (text) {
		if (textLen < s.size() + 1)
			return fail("rpwa_pwalik_print", "text buffer too small");
		std::memcpy(text, s.c_str(), s.size() + 1);
	}
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_call_info(const rpwa_pwalik* L, uint32_t calls[5], double totalTime[5])
{
	GUARD_BEGIN
	if (!L)
		return fail("rpwa_pwalik_call_info", "null pointer");
	WITH_LIK(L, {
		typedef typename std::remove_reference<decltype(lik)>::type likT;
		for (int i = 0; i < 5; ++i) {
			const auto& ci = lik.callInfo((typename likT::functionCallEnum)i);
			if (calls)
				calls[i] = ci.nmbCalls;
			if (totalTime)
				totalTime[i] = ci.totalTime;
		}
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_gpu_context(const rpwa_pwalik* L, rpwa_lik_ctx** ctx)
{
	GUARD_BEGIN
	if (!L || !ctx)
		return fail("rpwa_pwalik_gpu_context", "null pointer");
	WITH_LIK(L, *ctx = lik.gpuContext());
	if (!*ctx) {
		set_error("rpwa_pwalik_gpu_context: finishInit has not set up a GPU context");
		return RPWA_ERR_STATE;
	}
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_pwalik_hessian_eigen(const rpwa_pwalik* L, const double* hessian, double* eigenValues,
                                         double* eigenVectors)
{
	GUARD_BEGIN
	if (!L || !hessian || !eigenValues)
		return fail("rpwa_pwalik_hessian_eigen", "null pointer");
	WITH_LIK(L, {
		const unsigned int P = lik.nmbPars();
		rpwa::denseMatrix h(P);
		std::memcpy(h.data.data(), hessian, (size_t)P * P * sizeof(double));
		const auto eigen = lik.HessianEigenVectors(h);
		for (size_t k = 0; k < eigen.size(); ++k) {
			eigenValues[k] = eigen[k].second;
			if (eigenVectors)
				std::memcpy(eigenVectors + k * P, eigen[k].first.data(), P * sizeof(double));
		}
	});
	return RPWA_OK;
	GUARD_END
}

extern "C" int rpwa_fit_start_values(const rpwa_pwalik* L, uint32_t seed, double* out)
{
	GUARD_BEGIN
	if (!L || !out)
		return fail("rpwa_fit_start_values", "null pointer");
	WITH_LIK(L, {
		const std::vector<double> v = rpwa::hli::randomStartValues(lik, seed);
		std::memcpy(out, v.data(), v.size() * sizeof(double));
	});
	return RPWA_OK;
	GUARD_END
}

// shared part of the two driver entry points
static rpwa::minimizerOptions fit_options(const rpwa_fit_options* o, const bool minuit)
{
	rpwa::minimizerOptions m = minuit ? rpwa::hli::minuitDefaults() : rpwa::minimizerOptions();
	if (o) {
		m.memory = o->memory;
		m.maxEvaluations = o->maxEvaluations;
		m.xtolRel = o->xtolRel;
		m.ftolAbs = o->ftolAbs;
		m.edmTolerance = o->edmTolerance;
	}
	return m;
}

static void fit_outputs(const rpwa::hli::fitResultData& r, const unsigned int P, double* outPar,
                        double* outCovariance, double* outEigenValues, rpwa_fit_summary* out)
{
	std::memcpy(outPar, r.fitParameters.data(), (size_t)P * sizeof(double));
	if (outCovariance) {
		if (r.covMatrix.n == P)
			std::memcpy(outCovariance, r.covMatrix.data.data(), (size_t)P * P * sizeof(double));
		else
			std::memset(outCovariance, 0, (size_t)P * P * sizeof(double));
	}
	if (outEigenValues && !r.hessianEigenValues.empty())
		std::memcpy(outEigenValues, r.hessianEigenValues.data(), r.hessianEigenValues.size() * sizeof(double));
	uint32_t nonPositive = 0;
	for (double e : r.hessianEigenValues)
		if (e <= 0.)
			++nonPositive;
	out->logLikelihood = r.logLikelihood;
	out->edm = r.edm;
	out->status = r.status;
	out->converged = r.converged ? 1 : 0;
	out->hasHessian = r.hasHessian ? 1 : 0;
	out->nmbEvaluations = r.nmbEvaluations;
	out->nmbIterations = r.nmbIterations;
	out->nmbNonPositiveEigenvalues = nonPositive;
}

extern "C" int rpwa_fit_nlopt(const rpwa_pwalik* L, uint32_t seed, int checkHessian, const rpwa_fit_options* opts,
                              double* outPar, double* outCovariance, double* outEigenValues, rpwa_fit_summary* out)
{
	GUARD_BEGIN
	if (!L || !outPar || !out)
		return fail("rpwa_fit_nlopt", "null pointer");
	if (opts && (opts->memory == 0 || opts->maxEvaluations == 0))
		return fail("rpwa_fit_nlopt", "memory and maxEvaluations must be positive");
	const rpwa::minimizerOptions m = fit_options(opts, false);
	WITH_LIK(L, {
		const rpwa::hli::fitResultData r =
			rpwa::hli::pwaNloptFit(lik, seed, checkHessian != 0 || outEigenValues, outCovariance == nullptr, false, m);
		fit_outputs(r, lik.nmbPars(), outPar, outCovariance, outEigenValues, out);
	});
	return RPWA_OK;
	}
	catch (const rpwa::hli::signConventionError& ex) {
		set_error(ex.what());
		return RPWA_ERR_INVALID;
	}
	catch (const std::logic_error& ex) {
		set_error(ex.what());
		return RPWA_ERR_STATE;
	}
	catch (const std::exception& ex) {
		set_error(ex.what());
		return RPWA_ERR_HIP;
	}
}

extern "C" int rpwa_fit_minuit(const rpwa_pwalik* L, uint32_t seed, const double* startPar, int checkHessian,
                               const rpwa_fit_options* opts, double* outPar, double* outCovariance,
                               double* outEigenValues, rpwa_fit_summary* out)
{
	GUARD_BEGIN
	if (!L || !outPar || !out)
		return fail("rpwa_fit_minuit", "null pointer");
	if (opts && (opts->memory == 0 || opts->maxEvaluations == 0))
		return fail("rpwa_fit_minuit", "memory and maxEvaluations must be positive");
	const rpwa::minimizerOptions m = fit_options(opts, true);
	WITH_LIK(L, {
		std::vector<double> start;
		if (startPar)
			start.assign(startPar, startPar + lik.nmbPars());
		const rpwa::hli::fitResultData r = rpwa::hli::pwaFit(lik, seed, startPar ? &start : nullptr,
		                                                     checkHessian != 0 || outEigenValues,
		                                                     outCovariance == nullptr, false, m);
		fit_outputs(r, lik.nmbPars(), outPar, outCovariance, outEigenValues, out);
	});
	return RPWA_OK;
	}
	catch (const rpwa::hli::signConventionError& ex) {
		set_error(ex.what());
		return RPWA_ERR_INVALID;
	}
	catch (const std::logic_error& ex) {
		set_error(ex.what());
		return RPWA_ERR_STATE;
	}
	catch (const std::exception& ex) {
		set_error(ex.what());
		return RPWA_ERR_HIP;
	}
}
