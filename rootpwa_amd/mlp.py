"""Python face of the learned initializer's network (include/rootpwa_amd_mlp.h)."""
import ctypes

import numpy as np

from . import capi

_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64

PROTOTYPES = {
    "rpwa_mlp_create": (_i, [ctypes.POINTER(_vp), _i, _u32, _u32, _u32, _u32, _u64]),
    "rpwa_mlp_destroy": (_i, [_vp]),
    "rpwa_mlp_nmb_params": (_i, [_vp, ctypes.POINTER(_u64)]),
    "rpwa_mlp_get_params": (_i, [_vp, _fp]),
    "rpwa_mlp_set_params": (_i, [_vp, _fp]),
    "rpwa_mlp_forward": (_i, [_vp, _fp, _u32, _fp]),
    "rpwa_mlp_loss_grad": (_i, [_vp, _fp, _fp, _u32, _fp, _fp]),
    "rpwa_mlp_train": (_i, [_vp, _fp, _fp, _u32, _u32, _u32, ctypes.c_float, _u64, _fp]),
}

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


def _f(a):
    return a.ctypes.data_as(_fp)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    def __init__(self, nIn, nHidden, nHiddenLayers, nOut, seed=0, device=0):
        self.h = _vp()
        _call("rpwa_mlp_create", ctypes.byref(self.h), int(device), int(nIn), int(nHidden), int(nHiddenLayers),
              int(nOut), int(seed) & 0xFFFFFFFFFFFFFFFF)
        self.shape = (int(nIn), int(nHidden), int(nHiddenLayers), int(nOut))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            _lib().rpwa_mlp_destroy(h)
            self.h = None

    def nParams(self):
        n = _u64()
        _call("rpwa_mlp_nmb_params", self.h, ctypes.byref(n))
        return n.value

    def params(self):
        p = np.zeros(self.nParams(), dtype=np.float32)
        _call("rpwa_mlp_get_params", self.h, _f(p))
        return p

    def setParams(self, p):
        p = _f32(p)
        assert p.size == self.nParams()
        _call("rpwa_mlp_set_params", self.h, _f(p))

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.shape[3]), dtype=np.float32)
        _call("rpwa_mlp_forward", self.h, _f(x), x.shape[0], _f(y))
        return y

    def lossGrad(self, x, t):
        x, t = _f32(x), _f32(t)
        loss = ctypes.c_float()
        g = np.zeros(self.nParams(), dtype=np.float32)
        _call("rpwa_mlp_loss_grad", self.h, _f(x), _f(t), x.shape[0], ctypes.byref(loss), _f(g))
        return loss.value, g

    def train(self, x, t, batchSize=64, epochs=100, lr=1e-3, seed=0):
        x, t = _f32(x), _f32(t)
        hist = np.zeros(epochs, dtype=np.float32)
        _call("rpwa_mlp_train", self.h, _f(x), _f(t), x.shape[0], int(batchSize), int(epochs), float(lr),
              int(seed) & 0xFFFFFFFFFFFFFFFF, _f(hist))
        return hist
