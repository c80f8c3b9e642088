"""Python face of the batched fits (include/rootpwa_amd_batch.h): many bins,
each with many starts, minimised together on one GPU.

Each bin's likelihood is pwaLikelihood's for rank 1 with all waves in one
sector (partialWaveFit/pwaLikelihood.cc); each start follows pwaNloptFit's flow
(highLevelInterface/pwaNloptFit.cc:152-244); the bins and starts are the
attempts loop of pyInterface/module/_fit.py:132-165 run at once."""
import ctypes

import numpy as np

from . import capi
from .likelihood import FitOptions, NLOPT_DEFAULTS

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)

PROTOTYPES = {
    "rpwa_batch_create": (_i, [ctypes.POINTER(_vp), _i, _u32, _u32]),
    "rpwa_batch_destroy": (_i, [_vp]),
    "rpwa_batch_add_bin": (_i, [_vp, _vp, _u64, _i, _dp, _u64, _dp, _u64, _u32p]),
    "rpwa_batch_finish": (_i, [_vp]),
    "rpwa_batch_nmb_pars": (_i, [_vp, _u32p]),
    "rpwa_batch_nmb_bins": (_i, [_vp, _u32p]),
    "rpwa_batch_eval": (_i, [_vp, _dp, _dp, _dp]),
    "rpwa_batch_start_values": (_i, [_vp, _u32, _dp]),
    "rpwa_batch_fit": (_i, [_vp, _dp, _u32, _vp, _dp, _dp, _i32p, _u32p]),
    "rpwa_batch_set_tile": (_i, [_vp, _u32]),
    "rpwa_batch_info": (_i, [_vp, _u64p]),
    "rpwa_batch_time_event_kernel": (_i, [_vp, _u32, _dp]),
}

MAX_WAVES = 96
MAX_STARTS = 64

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


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class BatchFit:
    """All bins share the number of waves W and of starts S; P = 2 W parameters per fit."""

    def __init__(self, nmbWaves, nmbStarts, device=0):
        self.h = _vp()
        _call("rpwa_batch_create", ctypes.byref(self.h), int(device), int(nmbWaves), int(nmbStarts))
        self.W = int(nmbWaves)
        self.S = int(nmbStarts)
        self.P = 2 * self.W
        self.nEvents = []

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            _lib().rpwa_batch_destroy(h)
            self.h = None

    def addBin(self, amps, normSums, nmbNormEvents, accSums=None, nmbAccEvents=0):
        """amps: [N][W] complex data amplitudes; normSums / accSums: [W][W] complex
        sums over nmbNormEvents / nmbAccEvents MC events (accSums None: acceptance 1)."""
        a = _c128(amps)
        assert a.ndim == 2 and a.shape[1] == self.W
        return self._add(a.ctypes.data, a.shape[0], 0, normSums, nmbNormEvents, accSums, nmbAccEvents)

    def addBinDevice(self, d_amps, nEvents, normSums, nmbNormEvents, accSums=None, nmbAccEvents=0):
        """d_amps: device address (or capi.DeviceBuffer) of [nEvents][W] complex128."""
        ptr = d_amps.address if isinstance(d_amps, capi.DeviceBuffer) else int(d_amps)
        return self._add(ptr, int(nEvents), 1, normSums, nmbNormEvents, accSums, nmbAccEvents)

    def _add(self, ptr, n, onDevice, normSums, nmbNormEvents, accSums, nmbAccEvents):
        norm = _c128(normSums)
        acc = _c128(accSums) if accSums is not None else None
        idx = ctypes.c_uint32()
        _call("rpwa_batch_add_bin", self.h, _vp(ptr), int(n), int(onDevice),
              norm.view(np.float64).ctypes.data_as(_dp), int(nmbNormEvents),
              acc.view(np.float64).ctypes.data_as(_dp) if acc is not None else None, int(nmbAccEvents),
              ctypes.byref(idx))
        self.nEvents.append(int(n))
        return idx.value

    def finish(self):
        _call("rpwa_batch_finish", self.h)

    def nmbBins(self):
        n = ctypes.c_uint32()
        _call("rpwa_batch_nmb_bins", self.h, ctypes.byref(n))
        return n.value

    def eval(self, par):
        """par [nBins][S][P] -> (value [nBins][S], gradient [nBins][S][P])"""
        nb = self.nmbBins()
        p = np.ascontiguousarray(par, dtype=np.float64).reshape(nb, self.S, self.P)
        v = np.zeros((nb, self.S))
        g = np.zeros((nb, self.S, self.P))
        _call("rpwa_batch_eval", self.h, _d(p), _d(v), _d(g))
        return v, g

    def startValues(self, seed=0):
        out = np.zeros((self.nmbBins(), self.S, self.P))
        _call("rpwa_batch_start_values", self.h, int(seed) & 0xFFFFFFFF, _d(out))
        return out

    def fit(self, startPar=None, seed=0, **opts):
        """Minimises every fit; returns (par [nBins][S][P], value [nBins][S], status, nmbEval)."""
        nb = self.nmbBins()
        o = dict(NLOPT_DEFAULTS)
        o.update(opts)
        options = FitOptions(**o)
        par = np.zeros((nb, self.S, self.P))
        val = np.zeros((nb, self.S))
        st = np.zeros((nb, self.S), dtype=np.int32)
        ne = np.zeros((nb, self.S), dtype=np.uint32)
        sp = None
        if startPar is not None:
            sp = np.ascontiguousarray(startPar, dtype=np.float64).reshape(nb, self.S, self.P)
        _call("rpwa_batch_fit", self.h, _d(sp) if sp is not None else None, int(seed) & 0xFFFFFFFF,
              ctypes.byref(options), _d(par), _d(val), st.ctypes.data_as(_i32p), ne.ctypes.data_as(_u32p))
        return par, val, st, ne

    def setTile(self, wgPerCu):
        _call("rpwa_batch_set_tile", self.h, int(wgPerCu))

    def info(self):
        a = (ctypes.c_uint64 * 8)()
        _call("rpwa_batch_info", self.h, a)
        keys = ("bins", "tiles", "slots", "workgroups", "wgPerCu", "wavesPadded", "events", "lastEvaluations")
        return dict(zip(keys, list(a)))

    def timeEventKernel(self, reps=10):
        ms = ctypes.c_double()
        _call("rpwa_batch_time_event_kernel", self.h, int(reps), ctypes.byref(ms))
        return ms.value
