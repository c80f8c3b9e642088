"""Python face of the C++ likelihood (include/rootpwa_amd_pwalik.h).

Mirrors what pyRootPwa exposes for this path: the pwaLikelihood class
(pyInterface/bindings/partialWaveFit/pwaLikelihood_py.cc:243-312), the
ampIntegralMatrix class, and the initLikelihood setup sequence of
pyInterface/module/_likelihood.py:4-71, with arrays instead of ROOT files.
"""
import ctypes

import numpy as np

from . import capi

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_ip = ctypes.POINTER(ctypes.c_int)
_cpp = ctypes.POINTER(ctypes.c_char_p)

PROTOTYPES = {
    "rpwa_intmat_create": (_i, [ctypes.POINTER(_vp)]),
    "rpwa_intmat_destroy": (_i, [_vp]),
    "rpwa_intmat_set_wave_names": (_i, [_vp, _u32, _cpp]),
    "rpwa_intmat_nmb_waves": (_i, [_vp, _u32p]),
    "rpwa_intmat_wave_name": (_i, [_vp, _u32, ctypes.c_char_p, _i]),
    "rpwa_intmat_set_sums": (_i, [_vp, _dp, _u64]),
    "rpwa_intmat_get_sums": (_i, [_vp, _dp, _u64p]),
    "rpwa_intmat_set_nmb_events": (_i, [_vp, _u64]),
    "rpwa_intmat_element": (_i, [_vp, _u32, _u32, _dp, _dp]),
    "rpwa_intmat_integrate": (_i, [_vp, _vp, _u64, _vp, _i, _i]),
    "rpwa_intmat_integrate_incoherent": (_i, [_vp, _vp, _u32, _u64, _vp, _i, _i]),
    "rpwa_intmat_add": (_i, [_vp, _vp]),
    "rpwa_intmat_renormalize": (_i, [_vp, _u64]),
    "rpwa_intmat_write_ascii": (_i, [_vp, ctypes.c_char_p]),
    "rpwa_intmat_read_ascii": (_i, [_vp, ctypes.c_char_p]),
    "rpwa_pwalik_create": (_i, [ctypes.POINTER(_vp), _i]),
    "rpwa_pwalik_clone": (_i, [_vp, ctypes.POINTER(_vp)]),
    "rpwa_pwalik_destroy": (_i, [_vp]),
    "rpwa_pwalik_set_device": (_i, [_vp, _i]),
    "rpwa_pwalik_init": (_i, [_vp, _u32, _cpp, _ip, _dp, _u32, ctypes.c_double]),
    "rpwa_pwalik_use_normalized_amps": (_i, [_vp, _i]),
    "rpwa_pwalik_set_prior": (_i, [_vp, _i, ctypes.c_double]),
    "rpwa_pwalik_add_norm_integral": (_i, [_vp, _vp]),
    "rpwa_pwalik_add_acc_integral": (_i, [_vp, _vp, _u32]),
    "rpwa_pwalik_add_amplitude": (_i, [_vp, ctypes.c_char_p, _dp, _u64]),
    "rpwa_pwalik_add_amplitudes_device": (_i, [_vp, _vp, _u64]),
    "rpwa_pwalik_finish_init": (_i, [_vp]),
    "rpwa_pwalik_nmb_pars": (_i, [_vp, _u32p, _u32p]),
    "rpwa_pwalik_nmb_events": (_i, [_vp, _u64p]),
    "rpwa_pwalik_nmb_waves": (_i, [_vp, _i, _u32p]),
    "rpwa_pwalik_parameter": (_i, [_vp, _u32, ctypes.c_char_p, _i, _u32p, _ip, _ip]),
    "rpwa_pwalik_doeval": (_i, [_vp, _dp, _dp]),
    "rpwa_pwalik_gradient": (_i, [_vp, _dp, _dp]),
    "rpwa_pwalik_fdf": (_i, [_vp, _dp, _dp, _dp]),
    "rpwa_pwalik_time_fdf": (_i, [_vp, _dp, ctypes.c_uint32, _dp, _dp, _dp]),
    "rpwa_pwalik_doderivative": (_i, [_vp, _dp, _u32, _dp]),
    "rpwa_pwalik_hessian": (_i, [_vp, _dp, _dp]),
    "rpwa_pwalik_covariance": (_i, [_vp, _dp, _dp]),
    "rpwa_pwalik_correct_param_signs": (_i, [_vp, _dp, _dp]),
    "rpwa_pwalik_copy_from_par": (_i, [_vp, _dp, _dp, _dp]),
    "rpwa_pwalik_copy_to_par": (_i, [_vp, _dp, ctypes.c_double, _dp]),
    "rpwa_pwalik_integral_matrices": (_i, [_vp, _i, _dp, _dp, _dp]),
    "rpwa_pwalik_call_info": (_i, [_vp, _u32p, _dp]),
    "rpwa_pwalik_gpu_context": (_i, [_vp, ctypes.POINTER(_vp)]),
    "rpwa_pwalik_hessian_eigen": (_i, [_vp, _dp, _dp, _dp]),
    "rpwa_pwalik_set_quiet": (_i, [_i]),
    "rpwa_pwalik_enable_gpu": (_i, [_vp, _i]),
    "rpwa_pwalik_settings": (_i, [_vp, _ip, _ip, _dp, _ip]),
    "rpwa_pwalik_set_on_the_fly_binning": (_i, [_vp, _u32, _cpp, _dp, _dp, _u64, _dp]),
    "rpwa_pwalik_nmb_entries_in_bin": (_i, [_vp, _u64p]),
    "rpwa_pwalik_covariance_from_hessian": (_i, [_vp, _dp, _dp]),
    "rpwa_pwalik_build_prod_amp_arrays": (_i, [_vp, _dp, _i, _dp, ctypes.POINTER(ctypes.c_int32), _u32p,
                                               ctypes.c_char_p, _u64]),
    "rpwa_pwalik_print": (_i, [_vp, _i, ctypes.c_char_p, _u64, _u64p]),
    "rpwa_fit_start_values": (_i, [_vp, _u32, _dp]),
    "rpwa_fit_nlopt": (_i, [_vp, _u32, _i, _vp, _dp, _dp, _dp, _vp]),
    "rpwa_fit_minuit": (_i, [_vp, _u32, _dp, _i, _vp, _dp, _dp, _dp, _vp]),
}


class FitOptions(ctypes.Structure):
    """rpwa_fit_options"""
    _fields_ = [("memory", ctypes.c_uint32), ("maxEvaluations", ctypes.c_uint32), ("xtolRel", ctypes.c_double),
                ("ftolAbs", ctypes.c_double), ("edmTolerance", ctypes.c_double)]


class FitSummary(ctypes.Structure):
    """rpwa_fit_summary"""
    _fields_ = [("logLikelihood", ctypes.c_double), ("edm", ctypes.c_double), ("status", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("hasHessian", ctypes.c_int32), ("nmbEvaluations", ctypes.c_uint32),
                ("nmbIterations", ctypes.c_uint32), ("nmbNonPositiveEigenvalues", ctypes.c_uint32)]


# reference settings of the two drivers (pwaNloptFit.cc:105-107, pwaFit.cc:86-93)
NLOPT_DEFAULTS = dict(memory=10, maxEvaluations=50000, xtolRel=1e-4, ftolAbs=1e-6, edmTolerance=1e-10)
MINUIT_DEFAULTS = dict(memory=10, maxEvaluations=40000, xtolRel=0.0, ftolAbs=0.0, edmTolerance=1e-10)

_ready = False


def _lib():
    global _ready
    L = capi.lib()
    if not _ready:
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ready = True
    return L


def _call(name, *args):
    capi.check(name, getattr(_lib(), name)(*args))


def _d(a):
    return a.ctypes.data_as(_dp)


_RAW = {}


def _raw(name):
    """The library function without declared argument types, for the per-
    evaluation calls: the caller passes ctypes objects of the exact C types
    (the handle as c_void_p, pointers as ctypes pointers), which skips the
    per-argument converters (about 1 us of a C0 evaluation's 25 us)."""
    f = _RAW.get(name)
    if f is None:
        _lib()
        f = ctypes.CDLL(capi.LIB_PATH, handle=capi.lib()._handle)[name]
        f.restype = _i
        _RAW[name] = f
    return f


def _names(names):
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    return arr


class ampIntegralMatrix:
    """rpwa::ampIntegralMatrix: I_ij sums over MC events and their count."""

    def __init__(self, waveNames=None):
        self.h = _vp()
        _call("rpwa_intmat_create", ctypes.byref(self.h))
        if waveNames is not None:
            self.setWaveNames(waveNames)

    def __del__(self):
        try:
            if self.h and self.h.value:
                _lib().rpwa_intmat_destroy(self.h)
        except Exception:
            pass

    def setWaveNames(self, names):
        _call("rpwa_intmat_set_wave_names", self.h, len(names), _names(names))

    def nmbWaves(self):
        n = ctypes.c_uint32()
        _call("rpwa_intmat_nmb_waves", self.h, ctypes.byref(n))
        return n.value

    def waveNames(self):
        out = []
        buf = ctypes.create_string_buffer(512)
        for i in range(self.nmbWaves()):
            _call("rpwa_intmat_wave_name", self.h, i, buf, 512)
            out.append(buf.value.decode())
        return out

    def setSums(self, sums, nmbEvents):
        s = np.ascontiguousarray(sums, dtype=np.complex128)
        assert s.shape == (self.nmbWaves(), self.nmbWaves())
        _call("rpwa_intmat_set_sums", self.h, s.view(np.float64).ctypes.data_as(_dp), int(nmbEvents))

    def sums(self):
        W = self.nmbWaves()
        out = np.zeros((W, W), dtype=np.complex128)
        n = ctypes.c_uint64()
        _call("rpwa_intmat_get_sums", self.h, out.view(np.float64).ctypes.data_as(_dp), ctypes.byref(n))
        return out, n.value

    def nmbEvents(self):
        return self.sums()[1]

    def setNmbEvents(self, n):
        _call("rpwa_intmat_set_nmb_events", self.h, int(n))

    def element(self, i, j):
        re, im = ctypes.c_double(), ctypes.c_double()
        if isinstance(i, str) or isinstance(j, str):
            names = self.waveNames()
            i = names.index(i) if isinstance(i, str) else i
            j = names.index(j) if isinstance(j, str) else j
        _call("rpwa_intmat_element", self.h, int(i), int(j), ctypes.byref(re), ctypes.byref(im))
        return complex(re.value, im.value)

    def integrate(self, amps, weights=None, device=0):
        """Host amplitudes [M][W] complex, or [S][M][W] for S incoherent
        sub-amplitudes per wave -> sums on the GPU."""
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        self._keep = (a, w)
        wp = w.ctypes.data_as(_vp) if w is not None else None
        if a.ndim == 3:
            if w is not None and w.shape != (a.shape[1],):
                raise ValueError("one weight per event expected")
            _call("rpwa_intmat_integrate_incoherent", self.h, a.ctypes.data_as(_vp), a.shape[0], a.shape[1], wp,
                  int(device), 0)
        else:
            if a.ndim != 2 or (w is not None and w.shape != (a.shape[0],)):
                raise ValueError("amplitudes [M][W] and one weight per event expected")
            _call("rpwa_intmat_integrate", self.h, a.ctypes.data_as(_vp), a.shape[0], wp, int(device), 0)

    def integrateDevice(self, d_amps, nmbEvents, d_weights=None, device=0, nmbSubAmps=1):
        _call("rpwa_intmat_integrate_incoherent", self.h, _vp(d_amps), int(nmbSubAmps), int(nmbEvents),
              _vp(d_weights) if d_weights else None, int(device), 1)

    def __iadd__(self, other):
        _call("rpwa_intmat_add", self.h, other.h)
        return self

    def renormalize(self, n):
        _call("rpwa_intmat_renormalize", self.h, int(n))

    def writeAscii(self, fileName):
        _call("rpwa_intmat_write_ascii", self.h, fileName.encode())

    def readAscii(self, fileName):
        _call("rpwa_intmat_read_ascii", self.h, fileName.encode())


class fitParameter:
    """pwaLikelihood::fitParameter (partialWaveFit/pwaLikelihood.h:113-148)."""

    def __init__(self, waveName, rank, threshold, fixed, realPart):
        self._waveName, self._rank, self._threshold = waveName, int(rank), float(threshold)
        self._fixed, self._realPart = bool(fixed), bool(realPart)

    def waveName(self):
        return self._waveName

    def rank(self):
        return self._rank

    def threshold(self):
        return self._threshold

    def fixed(self):
        return self._fixed

    def realPart(self):
        return self._realPart

    def parName(self):
        """V<rank>_<wave>_RE / _IM, V_flat_RE for the flat wave (pwaLikelihood.cc:107-122)."""
        rank = "" if self._waveName == "flat" else str(self._rank)
        return "V%s_%s_%s" % (rank, self._waveName, "RE" if self._realPart else "IM")

    def __eq__(self, other):
        return isinstance(other, fitParameter) and (self._waveName, self._rank, self._threshold, self._fixed,
                                                    self._realPart) == (other._waveName, other._rank,
                                                                        other._threshold, other._fixed,
                                                                        other._realPart)

    def __repr__(self):
        return "fitParameter(%s, fixed=%s)" % (self.parName(), self._fixed)


class pwaLikelihood:
    """rpwa::pwaLikelihood evaluated on the GPU (C++ object behind the C ABI)."""

    FLAT = 0
    HALF_CAUCHY = 1
    # functionCallEnum (partialWaveFit/pwaLikelihood.h:77-84)
    FDF, GRADIENT, DOEVAL, DODERIVATIVE, HESSIAN = range(5)

    @staticmethod
    def setQuiet(flag=True):
        _call("rpwa_pwalik_set_quiet", int(bool(flag)))

    def __init__(self, precision=capi.PREC_FP64, _handle=None):
        self.h = _vp()
        if _handle is not None:
            self.h = _handle
        else:
            _call("rpwa_pwalik_create", ctypes.byref(self.h), int(precision))
        self.precision = precision
        self._waves = None
        self._keep = []

    def __del__(self):
        try:
            if self.h and self.h.value:
                _lib().rpwa_pwalik_destroy(self.h)
        except Exception:
            pass

    def Clone(self):
        out = _vp()
        _call("rpwa_pwalik_clone", self.h, ctypes.byref(out))
        c = pwaLikelihood(self.precision, _handle=out)
        c._waves = self._waves
        c._rank = self._rank
        return c

    # setup ---------------------------------------------------------------
    def setDevice(self, device):
        _call("rpwa_pwalik_set_device", self.h, int(device))

    def init(self, waves, rank=1, massBinCenter=0.0):
        """waves: list of (name, reflectivity, threshold)."""
        names = [w[0] for w in waves]
        refl = (ctypes.c_int * len(waves))(*[int(w[1]) for w in waves])
        thr = np.array([float(w[2]) if len(w) > 2 else 0.0 for w in waves])
        _call("rpwa_pwalik_init", self.h, len(waves), _names(names), refl, _d(thr), int(rank), float(massBinCenter))
        self._waves = list(waves)
        self._rank = int(rank)

    def useNormalizedAmps(self, flag=True):
        _call("rpwa_pwalik_use_normalized_amps", self.h, int(bool(flag)))

    def _settings(self):
        un, hc, gpu = _i(), _i(), _i()
        w = ctypes.c_double()
        _call("rpwa_pwalik_settings", self.h, ctypes.byref(un), ctypes.byref(hc), ctypes.byref(w), ctypes.byref(gpu))
        return bool(un.value), (self.HALF_CAUCHY if hc.value else self.FLAT), w.value, bool(gpu.value)

    def normalizedAmpsUsed(self):
        return self._settings()[0]

    def setPriorType(self, prior=0, cauchyWidth=None):
        """setPriorType(priorType) keeps the Cauchy width; giving cauchyWidth also sets it."""
        width = self.cauchyWidth() if cauchyWidth is None else float(cauchyWidth)
        _call("rpwa_pwalik_set_prior", self.h, int(prior == self.HALF_CAUCHY), width)

    def priorType(self):
        return self._settings()[1]

    def setCauchyWidth(self, cauchyWidth):
        _call("rpwa_pwalik_set_prior", self.h, int(self.priorType() == self.HALF_CAUCHY), float(cauchyWidth))

    def cauchyWidth(self):
        return self._settings()[2]

    def enableCuda(self, flag=True):
        """The GPU is the only event-sum backend: with flag False every evaluation raises."""
        _call("rpwa_pwalik_enable_gpu", self.h, int(bool(flag)))

    def cudaEnabled(self):
        return self._settings()[3]

    def setOnTheFlyBinning(self, multibinBoundaries, eventVariables):
        """setOnTheFlyBinning (partialWaveFit/pwaLikelihood.cc:1391-1436).

        multibinBoundaries: {variable: (lower, upper)}, an event is in the bin
        when lower <= value < upper for every variable.  eventVariables:
        {variable: array over all entries}, or a list of such dicts, one per
        event file, in the order in which the amplitude arrays concatenate the
        files.  addAmplitude then takes arrays over all entries."""
        if isinstance(eventVariables, dict):
            eventVariables = [eventVariables]
        names = sorted(multibinBoundaries)
        for v in names:
            if any(v not in ev for ev in eventVariables):
                raise KeyError("binning variable '%s' is not among the event variables" % v)
        cols = [np.concatenate([np.asarray(ev[v], dtype=np.float64).ravel() for ev in eventVariables])
                for v in names]
        nEntries = cols[0].size if cols else sum(
            len(next(iter(ev.values()))) if ev else 0 for ev in eventVariables)
        if any(c.size != nEntries for c in cols):
            raise ValueError("event variables differ in length")
        values = np.ascontiguousarray(np.stack(cols) if cols else np.zeros((0, nEntries)))
        lower = np.array([float(multibinBoundaries[v][0]) for v in names])
        upper = np.array([float(multibinBoundaries[v][1]) for v in names])
        _call("rpwa_pwalik_set_on_the_fly_binning", self.h, len(names), _names(names), _d(lower), _d(upper),
              int(nEntries), _d(values))

    def nmbEntriesInBin(self):
        n = ctypes.c_uint64()
        _call("rpwa_pwalik_nmb_entries_in_bin", self.h, ctypes.byref(n))
        return n.value

    def addNormIntegral(self, m):
        _call("rpwa_pwalik_add_norm_integral", self.h, m.h)

    def addAccIntegral(self, m, accEventsOverride=0):
        _call("rpwa_pwalik_add_acc_integral", self.h, m.h, int(accEventsOverride))

    def addAmplitude(self, waveName, amps):
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        _call("rpwa_pwalik_add_amplitude", self.h, waveName.encode(), a.view(np.float64).ctypes.data_as(_dp),
              a.shape[0])

    def addAmplitudesDevice(self, d_amps, nmbEvents):
        _call("rpwa_pwalik_add_amplitudes_device", self.h, _vp(d_amps), int(nmbEvents))

    def finishInit(self):
        _call("rpwa_pwalik_finish_init", self.h)
        self._nPars = self.nmbPars()  # fixed from here on: evaluations check their input against it

    # accessors -----------------------------------------------------------
    def nmbPars(self):
        n = ctypes.c_uint32()
        _call("rpwa_pwalik_nmb_pars", self.h, ctypes.byref(n), None)
        return n.value

    def NDim(self):
        return self.nmbPars()

    def nmbParsFixed(self):
        n = ctypes.c_uint32()
        _call("rpwa_pwalik_nmb_pars", self.h, None, ctypes.byref(n))
        return n.value

    def nmbEvents(self):
        n = ctypes.c_uint64()
        _call("rpwa_pwalik_nmb_events", self.h, ctypes.byref(n))
        return n.value

    def nmbWaves(self, reflectivity=0):
        n = ctypes.c_uint32()
        _call("rpwa_pwalik_nmb_waves", self.h, int(reflectivity), ctypes.byref(n))
        return n.value

    def rank(self):
        return self._rank

    def parameter(self, i):
        """fitParameter i (partialWaveFit/pwaLikelihood.h:195)."""
        name = ctypes.create_string_buffer(1024)
        r = ctypes.c_uint32()
        fx = ctypes.c_int()
        rp = ctypes.c_int()
        _call("rpwa_pwalik_parameter", self.h, int(i), name, 1024, ctypes.byref(r), ctypes.byref(fx),
              ctypes.byref(rp))
        parName = name.value.decode()
        if parName == "V_flat_RE":
            return fitParameter("flat", 0, 0.0, bool(fx.value), True)
        # V<rank>_<wave>_RE|IM
        wave = parName[len("V%d_" % r.value):-3]
        thr = {w[0]: (float(w[2]) if len(w) > 2 else 0.0) for w in (self._waves or [])}.get(wave, 0.0)
        return fitParameter(wave, r.value, thr, bool(fx.value), bool(rp.value))

    def parameters(self):
        return [self.parameter(i) for i in range(self.nmbPars())]

    # evaluation ------------------------------------------------------------
    def _par(self, par):
        p = np.ascontiguousarray(par, dtype=np.float64)
        n = getattr(self, "_nPars", None)
        assert p.shape == ((n if n is not None else self.nmbPars()),)
        return p

    def _eval_buffers(self, par):
        """The parameters copied into a buffer of this object whose C pointer is
        kept (converting a numpy array to a ctypes pointer costs microseconds,
        as much as an evaluation at small sizes), the gradient buffer and the
        value's c_double with its reference."""
        b = getattr(self, "_evalBufs", None)
        n = getattr(self, "_nPars", None)
        if b is None or n is None or b[0].size != n:
            n = n if n is not None else self.nmbPars()
            pb, gb, v = np.zeros(n), np.zeros(n), ctypes.c_double()
            b = self._evalBufs = (pb, gb, _d(pb), _d(gb), v, ctypes.byref(v))
        if type(par) is np.ndarray and par.shape == b[0].shape:
            np.copyto(b[0], par)
        else:
            np.copyto(b[0], self._par(par))
        return b

    def DoEval(self, par):
        _, _, pp, _, v, vr = self._eval_buffers(par)
        rc = _raw("rpwa_pwalik_doeval")(self.h, pp, vr)
        if rc:
            capi.check("rpwa_pwalik_doeval", rc)
        return v.value

    def Gradient(self, par):
        _, gb, pp, gp, _, _ = self._eval_buffers(par)
        rc = _raw("rpwa_pwalik_gradient")(self.h, pp, gp)
        if rc:
            capi.check("rpwa_pwalik_gradient", rc)
        return gb.copy()

    def FdF(self, par):
        _, gb, pp, gp, v, vr = self._eval_buffers(par)
        rc = _raw("rpwa_pwalik_fdf")(self.h, pp, vr, gp)
        if rc:
            capi.check("rpwa_pwalik_fdf", rc)
        return v.value, gb.copy()

    def timeFdF(self, par, n):
        """n FdF calls at par from C++ (rpwa_pwalik_time_fdf): the native call rate
        without this binding's per-call overhead; returns (seconds, value, gradient)."""
        p = self._par(par)
        g = np.zeros_like(p)
        v, sec = ctypes.c_double(), ctypes.c_double()
        _call("rpwa_pwalik_time_fdf", self.h, _d(p), int(n), ctypes.byref(v), _d(g), ctypes.byref(sec))
        return sec.value, v.value, g

    def DoDerivative(self, par, k):
        p = self._par(par)
        v = ctypes.c_double()
        _call("rpwa_pwalik_doderivative", self.h, _d(p), int(k), ctypes.byref(v))
        return v.value

    def Hessian(self, par):
        p = self._par(par)
        H = np.zeros((p.size, p.size))
        _call("rpwa_pwalik_hessian", self.h, _d(p), _d(H))
        return H

    def CovarianceMatrix(self, parOrHessian):
        """CovarianceMatrix(par) or CovarianceMatrix(hessian) (pwaLikelihood_py.cc:270-271)."""
        a = np.asarray(parOrHessian, dtype=np.float64)
        P = self.nmbPars()
        C = np.zeros((P, P))
        if a.ndim == 2:
            if a.shape != (P, P):
                raise ValueError("Hessian must be %d x %d" % (P, P))
            _call("rpwa_pwalik_covariance_from_hessian", self.h, _d(np.ascontiguousarray(a)), _d(C))
        else:
            _call("rpwa_pwalik_covariance", self.h, _d(self._par(a)), _d(C))
        return C

    def HessianEigenVectors(self, hessian):
        """[(eigenvector over all parameters, eigenvalue)], descending eigenvalues
        (partialWaveFit/pwaLikelihood.cc:864-911)."""
        H = np.ascontiguousarray(hessian, dtype=np.float64)
        P = self.nmbPars()
        nFree = P - self.nmbParsFixed()
        vals = np.zeros(max(nFree, 1))
        vecs = np.zeros((max(nFree, 1), P))
        _call("rpwa_pwalik_hessian_eigen", self.h, _d(H), _d(vals), _d(vecs))
        return [(vecs[k].copy(), float(vals[k])) for k in range(nFree)]

    def CorrectParamSigns(self, par):
        p = self._par(par)
        out = np.zeros_like(p)
        _call("rpwa_pwalik_correct_param_signs", self.h, _d(p), _d(out))
        return out

    def copyFromParArray(self, par):
        p = self._par(par)
        nW = (self.nmbWaves(-1), self.nmbWaves(+1))
        prod = np.zeros((self._rank, 2, max(nW)), dtype=np.complex128)
        flat = ctypes.c_double()
        _call("rpwa_pwalik_copy_from_par", self.h, _d(p), prod.view(np.float64).ctypes.data_as(_dp),
              ctypes.byref(flat))
        return prod, flat.value

    def copyToParArray(self, prod, flat):
        prod = np.ascontiguousarray(prod, dtype=np.complex128)
        par = np.zeros(self.nmbPars())
        _call("rpwa_pwalik_copy_to_par", self.h, prod.view(np.float64).ctypes.data_as(_dp), float(flat), _d(par))
        return par

    def integralMatrices(self, withFlat=False):
        n = self.nmbWaves() + (1 if withFlat else 0)
        N = np.zeros((n, n), dtype=np.complex128)
        A = np.zeros((n, n), dtype=np.complex128)
        ps = np.zeros(n)
        _call("rpwa_pwalik_integral_matrices", self.h, int(bool(withFlat)), N.view(np.float64).ctypes.data_as(_dp),
              A.view(np.float64).ctypes.data_as(_dp), _d(ps))
        return N, A, ps

    def getIntegralMatrices(self, withFlat=False):
        return self.integralMatrices(withFlat)

    def buildProdAmpArrays(self, par, withFlat=False):
        """(prodAmps, parIndices, prodAmpNames) (partialWaveFit/pwaLikelihood.cc:1624-1661);
        parIndices holds (re, im) parameter indices, -1 where there is none."""
        p = self._par(par)
        cap = self.nmbPars()
        amps = np.zeros(2 * cap)
        idx = (ctypes.c_int32 * (2 * cap))()
        n = ctypes.c_uint32()
        namesLen = sum(len(w[0]) + 24 for w in (self._waves or [])) * max(self._rank, 1) + 64
        names = ctypes.create_string_buffer(namesLen)
        _call("rpwa_pwalik_build_prod_amp_arrays", self.h, _d(p), int(bool(withFlat)), _d(amps), idx,
              ctypes.byref(n), names, namesLen)
        k = n.value
        prod = amps[:2 * k].view(np.complex128).copy()
        pairs = [(idx[2 * j], idx[2 * j + 1]) for j in range(k)]
        return prod, pairs, names.value.decode().split("\n") if k else []

    def _text(self, funcInfo):
        need = ctypes.c_uint64()
        _call("rpwa_pwalik_print", self.h, int(funcInfo), None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value)
        _call("rpwa_pwalik_print", self.h, int(funcInfo), buf, need.value, None)
        return buf.value.decode()

    def printFuncInfo(self):
        return self._text(1)

    def __str__(self):
        return self._text(0)

    def ncalls(self, func=0):
        return self.callInfo()[("FdF", "Gradient", "DoEval", "DoDerivative", "Hessian")[func]][0]

    def gpuContext(self):
        """Borrowed capi.LikelihoodContext view of the event-sum context."""
        h = _vp()
        _call("rpwa_pwalik_gpu_context", self.h, ctypes.byref(h))
        view = capi.LikelihoodContext.__new__(capi.LikelihoodContext)
        view.h = h
        view.precision = self.precision
        view.nW = (self.nmbWaves(-1), self.nmbWaves(+1))
        view.borrowed = True
        return view

    def callInfo(self):
        calls = (ctypes.c_uint32 * 5)()
        times = np.zeros(5)
        _call("rpwa_pwalik_call_info", self.h, calls, _d(times))
        return dict(zip(("FdF", "Gradient", "DoEval", "DoDerivative", "Hessian"), zip(list(calls), times.tolist())))


def initLikelihood(waves, amplitudes, normIntegral, accIntegral=None, rank=1, massBinCenter=0.0,
                   accEventsOverride=0, precision=capi.PREC_FP64, cauchyWidth=None, device=0,
                   useNormalizedAmps=True, multibinBoundaries=None, eventVariables=None):
    """The setup sequence of pyRootPwa.initLikelihood (pyInterface/module/_likelihood.py:4-71).

    waves: [(name, reflectivity, threshold)]; amplitudes: dict name -> complex
    array of data events; normIntegral / accIntegral: ampIntegralMatrix.
    With eventVariables (see pwaLikelihood.setOnTheFlyBinning) the amplitude
    arrays cover all entries of the event files and the events inside
    multibinBoundaries (empty or None: all) are kept."""
    L = pwaLikelihood(precision)
    L.setDevice(device)
    L.useNormalizedAmps(useNormalizedAmps)
    if cauchyWidth is not None:
        L.setPriorType(pwaLikelihood.HALF_CAUCHY)
        L.setCauchyWidth(cauchyWidth)
    L.init(waves, rank, massBinCenter)
    L.addNormIntegral(normIntegral)
    L.addAccIntegral(accIntegral if accIntegral is not None else normIntegral, accEventsOverride)
    if eventVariables is not None:
        L.setOnTheFlyBinning(multibinBoundaries or {}, eventVariables)
    for w in waves:
        L.addAmplitude(w[0], amplitudes[w[0]])
    L.finishInit()
    return L


class fitResult:
    """What fitResult::fill receives in the drivers (pwaNloptFit.cc:385-398), as arrays."""

    def __init__(self, L, par, cov, eig, summary):
        self.fitParameters = par
        self.covMatrix = cov
        self.hessianEigenValues = eig
        self._logLikelihood = summary.logLikelihood
        self._converged = bool(summary.converged)
        self._hasHessian = bool(summary.hasHessian)
        self.status = summary.status
        self.edm = summary.edm
        self.nmbEvaluations = summary.nmbEvaluations
        self.nmbIterations = summary.nmbIterations
        self.nmbNonPositiveEigenvalues = summary.nmbNonPositiveEigenvalues
        self.parNames = [L.parameter(i).parName() for i in range(L.nmbPars())]

    def logLikelihood(self):
        return self._logLikelihood

    def converged(self):
        return self._converged

    def covMatrixValid(self):
        return self._hasHessian

    def fitParameter(self, name):
        return float(self.fitParameters[self.parNames.index(name)])


def _fit(entry, L, seed, checkHessian, saveSpace, opts, defaults, startValues=None):
    o = dict(defaults)
    o.update(opts)
    options = FitOptions(**o)
    P = L.nmbPars()
    par = np.zeros(P)
    cov = None if saveSpace else np.zeros((P, P))
    wantEig = checkHessian or not saveSpace
    eig = np.zeros(max(P - L.nmbParsFixed(), 1)) if wantEig else None
    summary = FitSummary()
    args = [L.h, int(seed) & 0xFFFFFFFF]
    if entry == "rpwa_fit_minuit":
        sv = None if startValues is None else np.ascontiguousarray(startValues, dtype=np.float64)
        args.append(_d(sv) if sv is not None else None)
    args += [int(bool(checkHessian)), ctypes.byref(options), _d(par), _d(cov) if cov is not None else None,
             _d(eig) if eig is not None else None, ctypes.byref(summary)]
    _call(entry, *args)
    if eig is not None:
        eig = eig[:P - L.nmbParsFixed()]
    return fitResult(L, par, cov, eig, summary)


def startValues(likelihood, seed=0):
    """The drivers' random start values for all parameters (fixed ones 0)."""
    out = np.zeros(likelihood.nmbPars())
    _call("rpwa_fit_start_values", likelihood.h, int(seed) & 0xFFFFFFFF, _d(out))
    return out


def pwaNloptFit(likelihood, seed=0, checkHessian=False, saveSpace=False, **opts):
    """pwaNloptFit (highLevelInterface/pwaNloptFit.cc:84-402) on the GPU likelihood.
    opts override NLOPT_DEFAULTS."""
    return _fit("rpwa_fit_nlopt", likelihood, seed, checkHessian, saveSpace, opts, NLOPT_DEFAULTS)


def pwaFit(likelihood, seed=0, startValues=None, checkHessian=False, saveSpace=False, **opts):
    """pwaFit (highLevelInterface/pwaFit.cc:64-433): Minuit's call pattern
    (DoEval, then Gradient); startValues replaces the start-value file."""
    return _fit("rpwa_fit_minuit", likelihood, seed, checkHessian, saveSpace, opts, MINUIT_DEFAULTS, startValues)


def fitAttempts(likelihood, fitFunction=pwaNloptFit, seed=0, attempts=1, **kwargs):
    """The attempts loop of pyRootPwa's _pwaFit (pyInterface/module/_fit.py:132-165):
    seeds drawn with random.randint(1000, 2**32-1) from random.seed(seed), all
    seeds 0 if seed is 0; returns (results, index of best, index of best converged)."""
    import random
    if attempts == 1:
        seeds = [seed]
    elif seed == 0:
        seeds = [0] * attempts
    else:
        random.seed(seed)
        seeds = []
        for _ in range(attempts):
            while True:
                v = random.randint(1000, 2 ** 32 - 1)
                if v not in seeds:
                    break
            seeds.append(v)
    results, iBest, iBestValid = [], -1, -1
    best, bestValid = np.inf, np.inf
    for s in seeds:
        r = fitFunction(likelihood, seed=s, **kwargs)
        results.append(r)
        if r.logLikelihood() < best:
            best, iBest = r.logLikelihood(), len(results) - 1
        if r.converged() and r.logLikelihood() < bestValid:
            bestValid, iBestValid = r.logLikelihood(), len(results) - 1
    return results, iBest, iBestValid
