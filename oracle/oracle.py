"""ctypes loader of the CPU oracle (liborc.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
module, and only as the checker.  It wraps oracle/pwa_oracle.c, the plain-C
restatement of ROOTPWA's pwaLikelihood (partialWaveFit/pwaLikelihood.cc) and
ampIntegralMatrix::integrate (decayAmplitude/ampIntegralMatrix.cc), and the CPU
restatement of the synthetic-amplitude generator (oracle/gen_oracle.c).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_lik_new.restype = ctypes.c_void_p
        L.orc_lik_new.argtypes = [ctypes.c_uint, _up, _dp, ctypes.c_double]
        L.orc_lik_free.argtypes = [ctypes.c_void_p]
        for name in ("orc_lik_nmb_pars", "orc_lik_nmb_pars_fixed", "orc_lik_wmax"):
            getattr(L, name).restype = ctypes.c_uint
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_lik_nmb_events.restype = ctypes.c_uint64
        L.orc_lik_nmb_events.argtypes = [ctypes.c_void_p]
        L.orc_lik_tot_acc.restype = ctypes.c_double
        L.orc_lik_tot_acc.argtypes = [ctypes.c_void_p]
        L.orc_lik_par_info.argtypes = [ctypes.c_void_p, ctypes.c_uint] + [_ip] * 5
        L.orc_lik_par_map.argtypes = [ctypes.c_void_p, _ip]
        L.orc_lik_use_normalized_amps.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_lik_set_prior.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_lik_add_integrals.restype = ctypes.c_int
        L.orc_lik_add_integrals.argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint64, _dp, ctypes.c_uint64,
                                            ctypes.c_uint64]
        L.orc_lik_add_amplitude.restype = ctypes.c_int
        L.orc_lik_add_amplitude.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, _dp, ctypes.c_uint64]
        L.orc_lik_finish_init.restype = ctypes.c_int
        L.orc_lik_finish_init.argtypes = [ctypes.c_void_p]
        L.orc_lik_get_acc_matrix.argtypes = [ctypes.c_void_p, _dp]
        L.orc_lik_get_norm_matrix.argtypes = [ctypes.c_void_p, _dp]
        L.orc_lik_get_decay_amps.argtypes = [ctypes.c_void_p, ctypes.c_uint, _dp]
        L.orc_lik_copy_from_par.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_lik_copy_to_par.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.orc_lik_doeval.restype = ctypes.c_double
        L.orc_lik_doeval.argtypes = [ctypes.c_void_p, _dp]
        L.orc_lik_gradient.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_lik_fdf.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_lik_fdf_mt.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int]
        L.orc_lik_doderivative.restype = ctypes.c_double
        L.orc_lik_doderivative.argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint]
        L.orc_lik_hessian.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_lik_correct_param_signs.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_event_loglik.restype = ctypes.c_double
        L.orc_event_loglik.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.orc_event_deriv.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, _dp]
        L.orc_integral.argtypes = [_dp, ctypes.c_uint64, ctypes.c_uint, _dp, _dp]
        L.orc_integral_incoherent.argtypes = [_dp, ctypes.c_uint, ctypes.c_uint64, ctypes.c_uint, _dp, _dp]
        L.orc_intensity.argtypes = [_dp, ctypes.c_uint64, _up, _dp, ctypes.c_uint, _dp, _dp]
        L.orc_kahan_float_demo.argtypes = [ctypes.c_uint64, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]
        L.orc_gen_hash.restype = ctypes.c_uint64
        L.orc_gen_hash.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_gen_wave_scales.argtypes = [ctypes.c_uint64, ctypes.c_uint32, _dp]
        L.orc_gen_prod_amps.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _u32p, _dp]
        L.orc_gen_amps.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_uint32, _dp, _dp]
        L.orc_gen_accept_reject.restype = ctypes.c_uint64
        L.orc_gen_accept_reject.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _u32p,
                                            ctypes.c_uint32, _dp, _dp, _dp, _u64p, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def cplx_array(a):
    """complex ndarray -> contiguous complex128 (interleaved doubles)."""
    return np.ascontiguousarray(a, dtype=np.complex128)


class OracleLikelihood:
    """pwaLikelihood<complex<double>> restated on the CPU (the checker)."""

    def __init__(self, nWavesRefl, rank=1, thresholds=None, massBinCenter=0.0):
        self.L = lib()
        self.nW = (int(nWavesRefl[0]), int(nWavesRefl[1]))
        self.rank = int(rank)
        self.wmax = max(self.nW)
        nw = (ctypes.c_uint * 2)(*self.nW)
        thr = None
        if thresholds is not None:
            thr = np.ascontiguousarray(thresholds, dtype=np.float64).reshape(2, self.wmax)
            self._thr = thr
        self.h = self.L.orc_lik_new(self.rank, nw, _d(thr) if thr is not None else None, float(massBinCenter))
        if not self.h:
            raise ValueError("orc_lik_new failed")
        self.nPars = self.L.orc_lik_nmb_pars(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_lik_free(self.h)
            self.h = None

    # setup -------------------------------------------------------------
    def use_normalized_amps(self, flag=True):
        self.L.orc_lik_use_normalized_amps(self.h, int(bool(flag)))

    def set_prior(self, half_cauchy, width=0.5):
        self.L.orc_lik_set_prior(self.h, 1 if half_cauchy else 0, float(width))

    def add_integrals(self, normSums, normEvents, accSums=None, accEvents=None, accOverride=0):
        n = cplx_array(normSums)
        a = cplx_array(accSums if accSums is not None else normSums)
        accEvents = normEvents if accEvents is None else accEvents
        rc = self.L.orc_lik_add_integrals(self.h, n.view(np.float64).ctypes.data_as(_dp), int(normEvents),
                                          a.view(np.float64).ctypes.data_as(_dp), int(accEvents),
                                          int(accOverride))
        if rc != 0:
            raise RuntimeError("orc_lik_add_integrals: %d" % rc)

    def add_amplitude(self, refl, wave, amps):
        a = cplx_array(amps)
        rc = self.L.orc_lik_add_amplitude(self.h, int(refl), int(wave), a.view(np.float64).ctypes.data_as(_dp),
                                          a.shape[0])
        if rc != 0:
            raise RuntimeError("orc_lik_add_amplitude(%d,%d): %d" % (refl, wave, rc))

    def add_amplitudes(self, amps):
        """amps: [N][W0+W1] complex, refl -1 waves first."""
        amps = np.asarray(amps)
        for r in range(2):
            for w in range(self.nW[r]):
                col = w if r == 0 else self.nW[0] + w
                self.add_amplitude(r, w, amps[:, col])

    def finish_init(self):
        rc = self.L.orc_lik_finish_init(self.h)
        if rc != 0:
            raise RuntimeError("orc_lik_finish_init: %d" % rc)

    # accessors ---------------------------------------------------------
    def nmb_pars_fixed(self):
        return self.L.orc_lik_nmb_pars_fixed(self.h)

    def nmb_events(self):
        return self.L.orc_lik_nmb_events(self.h)

    def tot_acc(self):
        return self.L.orc_lik_tot_acc(self.h)

    def par_info(self, i):
        v = [ctypes.c_int() for _ in range(5)]
        self.L.orc_lik_par_info(self.h, i, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def par_map(self):
        out = np.zeros((self.rank, 2, self.wmax, 2), dtype=np.int32)
        self.L.orc_lik_par_map(self.h, out.ctypes.data_as(_ip))
        return out

    def acc_matrix(self):
        out = np.zeros((2, self.wmax, 2, self.wmax), dtype=np.complex128)
        self.L.orc_lik_get_acc_matrix(self.h, out.view(np.float64).ctypes.data_as(_dp))
        return out

    def norm_matrix(self):
        out = np.zeros((2, self.wmax, 2, self.wmax), dtype=np.complex128)
        self.L.orc_lik_get_norm_matrix(self.h, out.view(np.float64).ctypes.data_as(_dp))
        return out

    def decay_amps(self, refl):
        out = np.zeros((self.nmb_events(), self.nW[refl]), dtype=np.complex128)
        if out.size:
            self.L.orc_lik_get_decay_amps(self.h, refl, out.view(np.float64).ctypes.data_as(_dp))
        return out

    # mapping -----------------------------------------------------------
    def copy_from_par(self, par):
        par = np.ascontiguousarray(par, dtype=np.float64)
        prod = np.zeros((self.rank, 2, self.wmax), dtype=np.complex128)
        flat = ctypes.c_double()
        self.L.orc_lik_copy_from_par(self.h, _d(par), prod.view(np.float64).ctypes.data_as(_dp), ctypes.byref(flat))
        return prod, flat.value

    def copy_to_par(self, prod, flat):
        prod = cplx_array(prod)
        par = np.zeros(self.nPars)
        self.L.orc_lik_copy_to_par(self.h, prod.view(np.float64).ctypes.data_as(_dp), float(flat), _d(par))
        return par

    # evaluation --------------------------------------------------------
    def doeval(self, par):
        par = np.ascontiguousarray(par, dtype=np.float64)
        return self.L.orc_lik_doeval(self.h, _d(par))

    def gradient(self, par):
        par = np.ascontiguousarray(par, dtype=np.float64)
        g = np.zeros(self.nPars)
        self.L.orc_lik_gradient(self.h, _d(par), _d(g))
        return g

    def fdf(self, par, threads=1):
        par = np.ascontiguousarray(par, dtype=np.float64)
        g = np.zeros(self.nPars)
        f = ctypes.c_double()
        if threads > 1:
            self.L.orc_lik_fdf_mt(self.h, _d(par), ctypes.byref(f), _d(g), int(threads))
        else:
            self.L.orc_lik_fdf(self.h, _d(par), ctypes.byref(f), _d(g))
        return f.value, g

    def doderivative(self, par, k):
        par = np.ascontiguousarray(par, dtype=np.float64)
        return self.L.orc_lik_doderivative(self.h, _d(par), int(k))

    def hessian(self, par):
        par = np.ascontiguousarray(par, dtype=np.float64)
        H = np.zeros((self.nPars, self.nPars))
        self.L.orc_lik_hessian(self.h, _d(par), _d(H))
        return H

    def correct_param_signs(self, par):
        par = np.ascontiguousarray(par, dtype=np.float64)
        out = np.zeros(self.nPars)
        self.L.orc_lik_correct_param_signs(self.h, _d(par), _d(out))
        return out

    def event_loglik(self, prod, flat):
        prod = cplx_array(prod)
        return self.L.orc_event_loglik(self.h, prod.view(np.float64).ctypes.data_as(_dp), float(flat))

    def event_deriv(self, prod, flat):
        prod = cplx_array(prod)
        d = np.zeros((self.rank, 2, self.wmax), dtype=np.complex128)
        df = ctypes.c_double()
        self.L.orc_event_deriv(self.h, prod.view(np.float64).ctypes.data_as(_dp), float(flat),
                               d.view(np.float64).ctypes.data_as(_dp), ctypes.byref(df))
        return d, df.value


def integral(amps, weights=None):
    """ampIntegralMatrix::integrate sums, [W][W] complex, of amps [M][W] or,
    with incoherent sub-amplitudes, [S][M][W]."""
    a = np.asarray(amps)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    wp = _d(w) if w is not None else None
    if a.ndim == 3:
        a = np.ascontiguousarray(a, dtype=np.complex128)
        S, M, W = a.shape
        out = np.zeros((W, W), dtype=np.complex128)
        lib().orc_integral_incoherent(a.view(np.float64).ctypes.data_as(_dp), S, M, W, wp,
                                      out.view(np.float64).ctypes.data_as(_dp))
        return out
    a = cplx_array(amps)
    M, W = a.shape
    out = np.zeros((W, W), dtype=np.complex128)
    lib().orc_integral(a.view(np.float64).ctypes.data_as(_dp), M, W, wp, out.view(np.float64).ctypes.data_as(_dp))
    return out


def intensity(amps, nW, prod, waveNorm=None):
    """modelIntensity::getIntensity per event of amps [N][W0+W1], prod [rank][2][Wmax]."""
    a = cplx_array(amps)
    p = cplx_array(prod)
    out = np.zeros(a.shape[0])
    nw = (ctypes.c_uint * 2)(*nW)
    n = None if waveNorm is None else np.ascontiguousarray(waveNorm, dtype=np.float64)
    lib().orc_intensity(a.view(np.float64).ctypes.data_as(_dp), a.shape[0], nw, p.view(np.float64).ctypes.data_as(_dp),
                        p.shape[0], _d(n) if n is not None else None, _d(out))
    return out


def kahan_float_demo(n, x):
    a, b = ctypes.c_float(), ctypes.c_float()
    lib().orc_kahan_float_demo(int(n), float(x), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


# generator restatement ------------------------------------------------------
def gen_hash(seed, stream, ctr):
    return lib().orc_gen_hash(seed, stream, ctr)


def gen_wave_scales(seed, W):
    out = np.zeros(W)
    lib().orc_gen_wave_scales(seed, W, _d(out))
    return out


def gen_prod_amps(seed, stream, rank, nW):
    wmax = max(nW)
    out = np.zeros((rank, 2, wmax), dtype=np.complex128)
    nw = (ctypes.c_uint32 * 2)(*nW)
    lib().orc_gen_prod_amps(seed, stream, rank, nw, out.view(np.float64).ctypes.data_as(_dp))
    return out


def gen_amps(seed, stream, e0, n, scales):
    scales = np.ascontiguousarray(scales, dtype=np.float64)
    W = scales.shape[0]
    out = np.zeros((n, W), dtype=np.complex128)
    lib().orc_gen_amps(seed, stream, e0, n, W, _d(scales), out.view(np.float64).ctypes.data_as(_dp))
    return out


def gen_accept_reject(seed, stream, nWanted, nW, rank, prod, scales):
    scales = np.ascontiguousarray(scales, dtype=np.float64)
    prod = cplx_array(prod)
    W = nW[0] + nW[1]
    out = np.zeros((nWanted, W), dtype=np.complex128)
    idx = np.zeros(nWanted, dtype=np.uint64)
    smax = ctypes.c_double()
    nw = (ctypes.c_uint32 * 2)(*nW)
    pool = lib().orc_gen_accept_reject(seed, stream, nWanted, nw, rank, prod.view(np.float64).ctypes.data_as(_dp),
                                       _d(scales), out.view(np.float64).ctypes.data_as(_dp),
                                       idx.ctypes.data_as(_u64p), ctypes.byref(smax))
    if pool == 0:
        raise RuntimeError("accept-reject failed")
    return out, idx, pool, smax.value
