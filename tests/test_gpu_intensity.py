"""Per-event model intensity (rpwa_intensity, modelIntensity::getIntensity)
on the GPU against the oracle (oracle/pwa_oracle.c orc_intensity).

Tolerance per event: 1e-13 of sum_r sum_refl (sum_i |T_i| |a_i(e)|)^2, the
size of the terms whose rounding the sum of products can lose (the intensity
itself can be much smaller where the terms cancel)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _bound(a, nW, prod, norm):
    an = np.abs(a) / (norm if norm is not None else 1.0)
    b = np.zeros(a.shape[0])
    for s, cols in ((0, slice(0, nW[0])), (1, slice(nW[0], nW[0] + nW[1]))):
        if nW[s]:
            b += np.sum((an[:, cols] @ np.abs(prod[:, s, :nW[s]]).T) ** 2, axis=1)
    return b


@pytest.mark.parametrize("nW,rank", [((0, 16), 1), ((8, 80), 2), ((3, 11), 1), ((3, 11), 4), ((5, 2), 6),
                                     ((1, 1), 1), ((12, 0), 3), ((40, 200), 2)])
@pytest.mark.parametrize("N", [1, 17, 4097])
def test_intensity_matches_oracle(gpu, nW, rank, N):
    rng = np.random.default_rng(N + 7 * rank + sum(nW))
    W = sum(nW)
    a = (rng.normal(size=(N, W)) + 1j * rng.normal(size=(N, W))) * rng.uniform(0.1, 1.0, size=W)
    prod = H.random_prod(rng, rank, nW, 1.3)
    norm = rng.uniform(0.3, 3.0, size=W)
    for nrm in (None, norm):
        print("intensity", nW, "rank", rank, "events", N, "norm", nrm is not None, flush=True)
        got = gpu.intensity(a, nW, prod, nrm)
        ref = O.intensity(a, nW, prod, nrm)
        assert np.all(np.abs(got - ref) <= 1e-13 * _bound(a, nW, prod, nrm))


def test_intensity_full_size_device(gpu):
    """2M events x (8 + 80) waves, rank 2, in device memory: against the oracle
    on every event, and with the likelihood's normalisation its
    sum_e -ln(I_e + f^2) is the FP64 likelihood's event part."""
    from rootpwa_amd import synth
    nW, rank, N, M = (8, 80), 2, 2_000_000, 400_000
    d_data, sums, truth, _ = synth.generate_sample(20261207, nW, rank, N, M)
    norm = np.sqrt(np.real(np.diag(sums)) / M)
    prod = H.random_prod(np.random.default_rng(3), rank, nW, 0.9)
    d_out = gpu.DeviceBuffer(N * 8)
    gpu.intensity_device(d_data.address, N, nW, prod, d_out.address, norm)
    got = d_out.download((N,), np.float64)
    data = d_data.download((N, 88), np.complex128)
    ref = O.intensity(data, nW, prod, norm)
    assert np.all(np.abs(got - ref) <= 1e-13 * _bound(data, nW, prod, norm))
    ctx = gpu.LikelihoodContext(0, gpu.PREC_FP64)
    ctx.load(data / norm, nW)
    f = ctx.log_likelihood(prod, 0.2)
    assert abs(-np.sum(np.log(got + 0.04)) - f) <= 1e-10 * abs(f)
    ctx.close()
    d_out.free()
    d_data.free()
