"""ctypes binding of the C ABI declared in include/rootpwa_amd.h.

This is what a Python caller of the reference binds for this path: the
low-level event-sum backend (the role cuda::likelihoodInterface plays in
ROOTPWA) plus the integral and generator entry points.  The shared library is
built in-tree (rootpwa_amd/librootpwa_amd.so); importing this module without it
raises, there is no fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librootpwa_amd.so")

RPWA_OK = 0
RPWA_ERR_INVALID = 1
RPWA_ERR_STATE = 2
RPWA_ERR_HIP = 3
RPWA_ERR_NOMEM = 4
RPWA_ERR_UNSUPPORTED = 5

PREC_FP64 = 0
PREC_MIXED = 1

LAYOUT_REFL_WAVE_EVT = 0
LAYOUT_EVT_WAVE = 1

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_d = ctypes.c_double

# name -> (restype, argtypes); mirrors include/rootpwa_amd.h
PROTOTYPES = {
    "rpwa_last_error": (ctypes.c_char_p, []),
    "rpwa_device_count": (_i, []),
    "rpwa_lik_create": (_i, [ctypes.POINTER(_vp), _i, _i]),
    "rpwa_lik_retain": (_i, [_vp]),
    "rpwa_lik_release": (_i, [_vp]),
    "rpwa_lik_load_decay_amps": (_i, [_vp, _dp, _u64, _u64, _u32p, _i]),
    "rpwa_lik_load_decay_amps_device": (_i, [_vp, _vp, _u64, _u32p, _dp]),
    "rpwa_lik_log_likelihood": (_i, [_vp, _dp, _u32, _d, _u32, _dp]),
    "rpwa_lik_log_likelihood_deriv": (_i, [_vp, _dp, _u32, _d, _u32, _dp, _dp]),
    "rpwa_lik_fdf": (_i, [_vp, _dp, _u32, _d, _u32, _dp, _dp, _dp]),
    "rpwa_lik_hessian": (_i, [_vp, _dp, _u32, _d, _u32, _dp, _dp, _dp]),
    "rpwa_lik_submit": (_i, [_vp, _dp, _u32, _d, _u32]),
    "rpwa_lik_collect": (_i, [_vp, _dp, _dp, _dp]),
    "rpwa_lik_kernel_time": (_i, [_vp, _dp]),
    "rpwa_lik_reset_kernel_time": (_i, [_vp]),
    "rpwa_lik_set_timing": (_i, [_vp, _i]),
    "rpwa_lik_kernel_times": (_i, [_vp, _dp, _dp, _u64p]),
    "rpwa_lik_info": (_i, [_vp, _u64p]),
    "rpwa_lik_set_tile": (_i, [_vp, _u32]),
    "rpwa_lik_set_stage_tile": (_i, [_vp, _u32]),
    "rpwa_lik_set_stage_variant": (_i, [_vp, _u32, _i]),
    "rpwa_lik_kernel_variant": (_i, [_vp, _u32, _u32p]),
    "rpwa_lik_set_autotune": (_i, [_vp, _i]),
    "rpwa_lik_stream": (_i, [_vp, ctypes.POINTER(_vp)]),
    "rpwa_dev_alloc": (_i, [_i, _u64, ctypes.POINTER(_vp)]),
    "rpwa_dev_free": (_i, [_vp]),
    "rpwa_dev_copy": (_i, [_vp, _vp, _u64, _i]),
    "rpwa_dev_synchronize": (_i, [_i]),
    "rpwa_dev_info": (_i, [_i, ctypes.c_char_p, _i, ctypes.POINTER(_i), _u64p, ctypes.POINTER(_i)]),
    "rpwa_wave_nonzero_device": (_i, [_i, _vp, _u64, _u32, _u32, ctypes.POINTER(_i)]),
    "rpwa_bandwidth_probe": (_i, [_i, _u64, _i, _dp, _dp]),
    "rpwa_integral": (_i, [_i, _dp, _u64, _u32, _dp, _dp]),
    "rpwa_integral_device": (_i, [_i, _vp, _u64, _u32, _vp, _dp]),
    "rpwa_integral_device_noalloc": (_i, [_i, _vp, _u64, _u32, _vp, _dp]),
    "rpwa_data_moments_device": (_i, [_i, _vp, _u64, _u32, _dp, _i, _dp]),
    "rpwa_intensity_device": (_i, [_i, _vp, _u64, _u32p, _dp, _u32, _u32, _dp, _vp]),
    "rpwa_intensity": (_i, [_i, _dp, _u64, _u32p, _dp, _u32, _u32, _dp, _dp]),
    "rpwa_gen_wave_scales": (_i, [_u64, _u32, _dp]),
    "rpwa_gen_prod_amps": (_i, [_u64, _u32, _u32, _u32p, _dp]),
    "rpwa_gen_amps_device": (_i, [_i, _u64, _u32, _u64, _u64, _u32, _dp, _vp]),
    "rpwa_gen_accept_reject_device": (_i, [_i, _u64, _u32, _u64, _u32p, _u32, _dp, _dp, _vp, _u64p, _dp]),
}

_lib = None


class RpwaError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__("%s failed (status %d): %s" % (fn, code, msg))
        self.code = code


def lib():
    """Load librootpwa_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("rootpwa_amd: %s is missing; run __graft_entry__.build() (or "
                              "python rootpwa_amd/build.py) first" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(fn, code):
    if code != RPWA_OK:
        raise RpwaError(fn, code, lib().rpwa_last_error().decode())


def call(name, *args):
    check(name, getattr(lib(), name)(*args))


def device_count():
    return lib().rpwa_device_count()


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _cd(a):
    return a.view(np.float64).ctypes.data_as(_dp)


class LikelihoodContext:
    """Owning wrapper of an rpwa_lik_ctx (the event-sum backend)."""

    def __init__(self, device=0, precision=PREC_FP64):
        self.h = _vp()
        call("rpwa_lik_create", ctypes.byref(self.h), int(device), int(precision))
        self.precision = precision
        self.nW = None

    def close(self):
        if getattr(self, "borrowed", False):
            return
        if self.h and self.h.value:
            call("rpwa_lik_release", self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, amps, nW, layout=LAYOUT_EVT_WAVE, nEvents=None):
        """amps: complex array; EVT_WAVE: [N][W0+W1]; REFL_WAVE_EVT: [2][Wmax][N]."""
        a = _cplx(amps)
        if layout == LAYOUT_EVT_WAVE:
            nEvents = a.shape[0]
        nw = (ctypes.c_uint32 * 2)(*nW)
        call("rpwa_lik_load_decay_amps", self.h, _cd(a), a.size, int(nEvents), nw, int(layout))
        self.nW = tuple(nW)

    def load_device(self, d_ptr, nEvents, nW, waveNorm=None):
        nw = (ctypes.c_uint32 * 2)(*nW)
        sc = None
        if waveNorm is not None:
            self._scale = np.ascontiguousarray(waveNorm, dtype=np.float64)
            sc = _dptr(self._scale)
        call("rpwa_lik_load_decay_amps_device", self.h, _vp(d_ptr), int(nEvents), nw, sc)
        self.nW = tuple(nW)

    @property
    def wmax(self):
        return max(self.nW)

    def _prod(self, prod):
        p = _cplx(prod)
        rank = p.shape[0]
        assert p.shape == (rank, 2, self.wmax), p.shape
        return p, rank

    def log_likelihood(self, prod, flat):
        p, rank = self._prod(prod)
        out = ctypes.c_double()
        call("rpwa_lik_log_likelihood", self.h, _cd(p), p.size, float(flat), rank, ctypes.byref(out))
        return out.value

    def log_likelihood_deriv(self, prod, flat):
        p, rank = self._prod(prod)
        d = np.zeros_like(p)
        df = ctypes.c_double()
        call("rpwa_lik_log_likelihood_deriv", self.h, _cd(p), p.size, float(flat), rank, _cd(d), ctypes.byref(df))
        return d, df.value

    def fdf(self, prod, flat):
        p, rank = self._prod(prod)
        d = np.zeros_like(p)
        df = ctypes.c_double()
        f = ctypes.c_double()
        call("rpwa_lik_fdf", self.h, _cd(p), p.size, float(flat), rank, ctypes.byref(f), _cd(d), ctypes.byref(df))
        return f.value, d, df.value

    def submit(self, prod, flat):
        p, rank = self._prod(prod)
        self._pending = p
        call("rpwa_lik_submit", self.h, _cd(p), p.size, float(flat), rank)

    def collect(self):
        p = self._pending
        d = np.zeros_like(p)
        df = ctypes.c_double()
        f = ctypes.c_double()
        call("rpwa_lik_collect", self.h, ctypes.byref(f), _cd(d), ctypes.byref(df))
        return f.value, d, df.value

    def hessian(self, prod, flat):
        p, rank = self._prod(prod)
        nd = p.size
        H = np.zeros((nd, nd, 3))
        ft = np.zeros_like(p)
        hf = ctypes.c_double()
        call("rpwa_lik_hessian", self.h, _cd(p), p.size, float(flat), rank, _dptr(H), _cd(ft), ctypes.byref(hf))
        return H, ft, hf.value

    def kernel_time(self):
        t = ctypes.c_double()
        call("rpwa_lik_kernel_time", self.h, ctypes.byref(t))
        return t.value

    def kernel_times(self):
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        call("rpwa_lik_kernel_times", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n))
        return a.value, b.value, n.value

    def set_timing(self, enable=True):
        """Kernel timing with HIP events: True or 1 times every evaluation, N > 1
        every N-th, False or 0 turns it off."""
        call("rpwa_lik_set_timing", self.h, max(0, int(enable)))

    def reset_kernel_time(self):
        call("rpwa_lik_reset_kernel_time", self.h)

    def info(self):
        out = (ctypes.c_uint64 * 9)()
        call("rpwa_lik_info", self.h, out)
        keys = ("events", "w_neg", "w_pos", "precision", "lanes", "slots", "workgroups", "threads", "bytes")
        return dict(zip(keys, list(out)))

    def set_tile(self, lanes):
        call("rpwa_lik_set_tile", self.h, int(lanes))

    def set_stage_tile(self, events):
        call("rpwa_lik_set_stage_tile", self.h, int(events))

    def set_stage_variant(self, events, variant):
        """variant 0: VALU S and gradient; 1 (or True): gradient on FP32 MFMA;
        2: wave-tile variant (S on FP64 MFMA, gradient on FP32 MFMA)."""
        call("rpwa_lik_set_stage_variant", self.h, int(events), int(variant))

    def kernel_variant(self, rank):
        out = (ctypes.c_uint32 * 4)()
        call("rpwa_lik_kernel_variant", self.h, int(rank), out)
        return dict(zip(("lanes", "slots", "gradient_mfma", "workgroups"), list(out)))

    def set_autotune(self, enable):
        call("rpwa_lik_set_autotune", self.h, int(bool(enable)))

    def stream(self):
        s = _vp()
        call("rpwa_lik_stream", self.h, ctypes.byref(s))
        return s.value


class DeviceBuffer:
    """Device allocation through the library's own HIP runtime."""

    def __init__(self, nbytes, device=0):
        self.ptr = _vp()
        self.nbytes = int(nbytes)
        self.device = device
        call("rpwa_dev_alloc", int(device), self.nbytes, ctypes.byref(self.ptr))

    @property
    def address(self):
        return self.ptr.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        call("rpwa_dev_copy", self.ptr, a.ctypes.data_as(_vp), a.nbytes, 0)

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        call("rpwa_dev_copy", out.ctypes.data_as(_vp), self.ptr, out.nbytes, 1)
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            call("rpwa_dev_free", self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(device=0):
    call("rpwa_dev_synchronize", int(device))


def device_info(device=0):
    name = ctypes.create_string_buffer(64)
    cu = ctypes.c_int()
    mem = ctypes.c_uint64()
    clk = ctypes.c_int()
    call("rpwa_dev_info", int(device), name, 64, ctypes.byref(cu), ctypes.byref(mem), ctypes.byref(clk))
    return dict(arch=name.value.decode(), cus=cu.value, memory=mem.value, clock_khz=clk.value)


def wave_nonzero_device(d_amps, N, W, wave, device=0):
    """True if column `wave` of the device array [N][W] complex128 has a non-zero element."""
    out = _i()
    call("rpwa_wave_nonzero_device", device, _vp(d_amps), N, W, wave, ctypes.byref(out))
    return bool(out.value)


def bandwidth_probe(nbytes=4 << 30, repetitions=10, device=0):
    """(read GB/s, copy GB/s) of streaming kernels over nbytes of HBM."""
    r, c = ctypes.c_double(), ctypes.c_double()
    call("rpwa_bandwidth_probe", int(device), int(nbytes), int(repetitions), ctypes.byref(r), ctypes.byref(c))
    return r.value, c.value


def integral(amps, weights=None, device=0):
    a = _cplx(amps)
    M, W = a.shape
    out = np.zeros((W, W), dtype=np.complex128)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    call("rpwa_integral", int(device), _cd(a), M, W, _dptr(w) if w is not None else None, _cd(out))
    return out


def integral_device(d_amps, M, W, d_weights=None, device=0):
    out = np.zeros((W, W), dtype=np.complex128)
    call("rpwa_integral_device", int(device), _vp(d_amps), int(M), int(W),
         _vp(d_weights) if d_weights else None, _cd(out))
    return out


def integral_device_noalloc(d_amps, M, W, d_weights=None, device=0):
    """rpwa_integral_device_noalloc: the same sums from a static device work area."""
    out = np.zeros((W, W), dtype=np.complex128)
    call("rpwa_integral_device_noalloc", int(device), _vp(d_amps), int(M), int(W),
         _vp(d_weights) if d_weights else None, _cd(out))
    return out


def data_moments_device(d_amps, N, W, normDiag, perEventNorm=True, device=0):
    """[W][W] complex mean moments of amplitudes normalised by sqrt(normDiag) (per-event
    normalised with perEventNorm), FP64 MFMA on the device."""
    nd = np.ascontiguousarray(normDiag, dtype=np.float64)
    out = np.zeros((W, W), dtype=np.complex128)
    call("rpwa_data_moments_device", int(device), _vp(d_amps), int(N), int(W), _dptr(nd), int(bool(perEventNorm)),
         _cd(out))
    return out


def intensity(amps, nW, prod, waveNorm=None, device=0):
    """Per-event model intensity sum_r sum_refl |sum_i T a_i|^2 of host amplitudes
    [N][W0+W1] and production amplitudes [rank][2][Wmax] (rpwa_intensity)."""
    a = _cplx(amps)
    p = _cplx(prod)
    rank = p.shape[0]
    nw = (ctypes.c_uint32 * 2)(*nW)
    out = np.zeros(a.shape[0])
    nrm = None if waveNorm is None else np.ascontiguousarray(waveNorm, dtype=np.float64)
    call("rpwa_intensity", int(device), _cd(a), a.shape[0], nw, _cd(p), p.size, rank,
         _dptr(nrm) if nrm is not None else None, _dptr(out))
    return out


def intensity_device(d_amps, N, nW, prod, d_out, waveNorm=None, device=0):
    """rpwa_intensity_device: amplitudes [N][W0+W1] and the result [N] in device memory."""
    p = _cplx(prod)
    nw = (ctypes.c_uint32 * 2)(*nW)
    nrm = None if waveNorm is None else np.ascontiguousarray(waveNorm, dtype=np.float64)
    call("rpwa_intensity_device", int(device), _vp(d_amps), int(N), nw, _cd(p), p.size, p.shape[0],
         _dptr(nrm) if nrm is not None else None, _vp(d_out))


def gen_wave_scales(seed, W):
    out = np.zeros(W)
    call("rpwa_gen_wave_scales", int(seed), int(W), _dptr(out))
    return out


def gen_prod_amps(seed, stream, rank, nW):
    out = np.zeros((rank, 2, max(nW)), dtype=np.complex128)
    nw = (ctypes.c_uint32 * 2)(*nW)
    call("rpwa_gen_prod_amps", int(seed), int(stream), int(rank), nw, _cd(out))
    return out


def gen_amps_device(d_out, seed, stream, e0, n, scales, device=0):
    sc = np.ascontiguousarray(scales, dtype=np.float64)
    call("rpwa_gen_amps_device", int(device), int(seed), int(stream), int(e0), int(n), sc.shape[0], _dptr(sc),
         _vp(d_out))


def gen_accept_reject_device(d_out, seed, stream, nWanted, nW, rank, prod, scales, device=0):
    sc = np.ascontiguousarray(scales, dtype=np.float64)
    p = _cplx(prod)
    nw = (ctypes.c_uint32 * 2)(*nW)
    pool = ctypes.c_uint64()
    smax = ctypes.c_double()
    call("rpwa_gen_accept_reject_device", int(device), int(seed), int(stream), int(nWanted), nw, int(rank),
         _cd(p), _dptr(sc), _vp(d_out), ctypes.byref(pool), ctypes.byref(smax))
    return pool.value, smax.value
