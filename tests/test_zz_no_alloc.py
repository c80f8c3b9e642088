"""rpwa_integral_device_noalloc (a static device work area instead of
hipMalloc'd buffers) against rpwa_integral_device: bitwise equal sums, two
calls in a row with the second smaller, with and without importance weights;
and the size limit of the work area."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _upload(gpu, a):
    a = np.ascontiguousarray(a)
    d = gpu.DeviceBuffer(a.nbytes)
    d.upload(a)
    return d


def _amps(rng, M, W):
    return (rng.normal(size=(M, W)) + 1j * rng.normal(size=(M, W))) * rng.uniform(0.1, 1.0, size=W)


@pytest.mark.parametrize("weighted", [False, True], ids=["unweighted", "weighted"])
def test_noalloc_integral_equals_allocating_one(gpu, weighted):
    rng = np.random.default_rng(31 + weighted)
    # the second call is smaller in events and in waves; the third returns to a
    # wider row after it, so stale contents of the work area would show
    for M, W in ((300_001, 88), (20_000, 17), (70_000, 128)):
        a = _amps(rng, M, W)
        d = _upload(gpu, a)
        dw = _upload(gpu, rng.uniform(0.5, 2.0, size=M)) if weighted else None
        print("noalloc integral M", M, "W", W, "weighted", weighted, flush=True)
        ref = gpu.integral_device(d.address, M, W, dw.address if dw else None)
        got = gpu.integral_device_noalloc(d.address, M, W, dw.address if dw else None)
        assert np.array_equal(got.view(np.float64), ref.view(np.float64))
        again = gpu.integral_device_noalloc(d.address, M, W, dw.address if dw else None)
        assert np.array_equal(again.view(np.float64), ref.view(np.float64))
        # and the sums are the integral (a Hermitian matrix with the plain sums' diagonal)
        if not weighted:
            diag = np.sum(np.abs(a) ** 2, axis=0)
            assert np.allclose(np.real(np.diag(got)), diag, rtol=1e-12)
        assert np.allclose(got, np.conj(got.T), rtol=0, atol=1e-9 * np.max(np.abs(got)))
        d.free()
        if dw:
            dw.free()


def test_noalloc_integral_limits(gpu):
    rng = np.random.default_rng(5)
    a = _amps(rng, 100, 129)
    d = _upload(gpu, a)
    out = np.zeros((129, 129), dtype=np.complex128)
    L = gpu.lib()
    rc = L.rpwa_integral_device_noalloc(0, ctypes.c_void_p(d.address), 100, 129, None,
                                        out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == gpu.RPWA_ERR_UNSUPPORTED
    assert b"128" in L.rpwa_last_error()
    rc = L.rpwa_integral_device_noalloc(0, None, 100, 12, None,
                                        out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == gpu.RPWA_ERR_INVALID
    d.free()
