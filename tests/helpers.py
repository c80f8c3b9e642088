"""Shared helpers of the test suite: synthetic workloads built with the CPU
restatement of the generator and checked likelihoods built with the oracle."""
import numpy as np

from oracle import oracle as O

STREAM_TRUE_T = 2
STREAM_MC = 3
STREAM_DATA = 4


def dataset(nW, rank, nData, nMC, seed):
    """Amplitudes of an MC pool and of accept-rejected data (the benchmark recipe)."""
    W = nW[0] + nW[1]
    scales = O.gen_wave_scales(seed, W)
    mc = O.gen_amps(seed, STREAM_MC, 0, nMC, scales)
    truth = O.gen_prod_amps(seed, STREAM_TRUE_T, rank, nW)
    data, idx, pool, smax = O.gen_accept_reject(seed, STREAM_DATA, nData, nW, rank, truth, scales)
    return dict(scales=scales, mc=mc, truth=truth, data=data, idx=idx, pool=pool, smax=smax,
                integral=O.integral(mc), nMC=nMC, nW=tuple(nW), rank=rank)


def oracle_likelihood(ds, prior=None):
    L = O.OracleLikelihood(ds["nW"], ds["rank"])
    if prior is not None:
        L.set_prior(True, prior)
    L.add_integrals(ds["integral"], ds["nMC"])
    L.add_amplitudes(ds["data"])
    L.finish_init()
    return L


def normalised_amps(L):
    """The oracle's normalised decay amplitudes as one [N][W0+W1] array."""
    return np.concatenate([L.decay_amps(0), L.decay_amps(1)], axis=1)


def random_prod(rng, rank, nW, scale=1.0):
    """Production amplitudes obeying the rank restrictions of pwaLikelihood."""
    wmax = max(nW)
    p = (rng.normal(size=(rank, 2, wmax)) + 1j * rng.normal(size=(rank, 2, wmax))) * scale
    for r in range(rank):
        for s in range(2):
            p[r, s, nW[s]:] = 0
            p[r, s, :r] = 0
            if r < nW[s]:
                p[r, s, r] = p[r, s, r].real
    return p


def random_par(rng, L, scale=1.0):
    return rng.normal(size=L.nPars) * scale


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
