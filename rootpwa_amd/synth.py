"""Synthetic bins for batched fits, generated on the GPU with the library's
counter-based generator (rpwa_gen_*), the recipe of the benchmark workloads:
A_i(e) = (N(0,1) + i N(0,1)) s_i with s_i ~ U[0.1, 1], each event's wave vector
scaled by one lognormal(0, 0.5) factor; per bin a true T ~ complex N(0,1);
data accept-rejected from a fresh pool by |sum_i T_i A_i|^2; the normalisation
integral from an MC pool of mcFactor times the data count (acceptance 1).
Every bin is rank 1 with all waves in the refl +1 sector."""
import numpy as np

from . import capi

STREAM_TRUE_T, STREAM_MC, STREAM_DATA = 2, 3, 4


def bin_sizes(seed, nBins, lo=20_000, hi=400_000):
    """Data counts log-uniform in [lo, hi] (numpy PCG64 seeded with seed)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=nBins)).astype(np.int64)


def bin_seed(seed, b):
    return (int(seed) * 1_000_003 + 7919 * int(b)) & 0xFFFFFFFF


def generate_bin(seed, W, nData, mcFactor=5, device=0):
    """One bin on the device: returns (device buffer [nData][W] complex128,
    integral sums [W][W], nMC, truth T [W])."""
    nW = (0, W)
    scales = capi.gen_wave_scales(seed, W)
    truth = capi.gen_prod_amps(seed, STREAM_TRUE_T, 1, nW)
    M = int(mcFactor * nData)
    d_mc = capi.DeviceBuffer(M * W * 16, device)
    capi.gen_amps_device(d_mc.address, seed, STREAM_MC, 0, M, scales, device)
    sums = capi.integral_device(d_mc.address, M, W, device=device)
    d_mc.free()
    d_data = capi.DeviceBuffer(int(nData) * W * 16, device)
    capi.gen_accept_reject_device(d_data.address, seed, STREAM_DATA, int(nData), nW, 1, truth, scales, device)
    return d_data, sums, M, truth[0, 1, :W].copy()


def generate_sample(seed, nW, rank, nData, nMC, device=0):
    """A single-bin workload on the device with both reflectivity sectors:
    returns (device buffer [nData][W0+W1] complex128 of accept-rejected data,
    integral sums [W][W] of an nMC-event MC pool, the true production
    amplitudes [rank][2][Wmax], the size of the data pool)."""
    W = nW[0] + nW[1]
    scales = capi.gen_wave_scales(seed, W)
    truth = capi.gen_prod_amps(seed, STREAM_TRUE_T, rank, nW)
    d_mc = capi.DeviceBuffer(int(nMC) * W * 16, device)
    capi.gen_amps_device(d_mc.address, seed, STREAM_MC, 0, int(nMC), scales, device)
    sums = capi.integral_device(d_mc.address, int(nMC), W, device=device)
    d_mc.free()
    d_data = capi.DeviceBuffer(int(nData) * W * 16, device)
    pool, _ = capi.gen_accept_reject_device(d_data.address, seed, STREAM_DATA, int(nData), nW, rank, truth, scales,
                                            device)
    return d_data, sums, truth, pool


def fill_batch(batch, seed, sizes, mcFactor=5, device=0, log=None):
    """Adds one generated bin per entry of sizes to a rootpwa_amd.batch.BatchFit;
    returns the true production amplitudes [nBins][W]."""
    truths = []
    for b, n in enumerate(sizes):
        d_data, sums, M, truth = generate_bin(bin_seed(seed, b), batch.W, int(n), mcFactor, device)
        batch.addBinDevice(d_data, int(n), sums, M)
        d_data.free()
        truths.append(truth)
        if log is not None and (b % 8 == 7 or b == len(sizes) - 1):
            log("generated bin %d / %d" % (b + 1, len(sizes)))
    return np.array(truths)
