"""The five benchmark configurations at their full workload sizes on the GPU,
against the oracle.

C0  100k events x 16 waves (one sector), rank 1, 200k MC; FP64; against the
    single-thread oracle FdF to 1e-10.
C1  2M events x (8 + 80) waves, rank 2, 5M MC, generated on the GPU and loaded
    through pwaLikelihood::addAmplitudesDevice; FP64 FdF and a full
    Minuit-pattern fit (pwaFit) from one random start; lnL and gradient against
    the oracle (16 threads) to 1e-10 relative.
C2  the same workload with complex64 storage: |dlnL| < 1e-3 absolute against
    the FP64 oracle, gradient to 1e-5 of its scale.
C3  48 bins, data counts log-uniform in [20k, 400k], 88 waves, rank 1, 64
    starts: the batch's value and gradient against the oracle in the smallest,
    the median and the largest bin, at a start and at the fitted minimum, with
    |dlnL| < 1e-3 absolute.
C4  the learned initializer's network at its full width (775 -> 3 x 512 GELU
    -> 176) against the FP64 oracle, and the protocol on 88-wave bins.

The gradient tolerances at a minimum are relative to the gradient's size at
the start, since the gradient itself vanishes there."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _waves(nW):
    return [("neg%02d" % i, -1, 0.0) for i in range(nW[0])] + [("pos%02d" % i, +1, 0.0) for i in range(nW[1])]


def _device_likelihood(PL, precision, nW, rank, d_data, N, sums, M):
    waves = _waves(nW)
    names = [w[0] for w in waves]
    I = PL.ampIntegralMatrix(names)
    I.setSums(sums, M)
    L = PL.pwaLikelihood(precision)
    L.init(waves, rank)
    L.addNormIntegral(I)
    L.addAccIntegral(I)
    L.addAmplitudesDevice(d_data.address, N)
    L.finishInit()
    return L


def _oracle(nW, rank, data, sums, M):
    Lo = O.OracleLikelihood(nW, rank)
    Lo.add_integrals(sums, M)
    Lo.add_amplitudes(data)
    Lo.finish_init()
    return Lo


def test_c0_full_size(gpu):
    from rootpwa_amd import likelihood as PL
    from rootpwa_amd import synth
    nW, rank, N, M = (0, 16), 1, 100_000, 200_000
    d_data, sums, truth, _ = synth.generate_sample(20261202, nW, rank, N, M)
    L = _device_likelihood(PL, gpu.PREC_FP64, nW, rank, d_data, N, sums, M)
    data = d_data.download((N, 16), np.complex128)
    Lo = _oracle(nW, rank, data, sums, M)
    assert L.nmbPars() == 32
    rng = np.random.default_rng(20261202 + 5)
    for k in range(3):
        par = rng.normal(size=32) * np.sqrt(N / 16)
        f, g = L.FdF(par)
        f_ref, g_ref = Lo.fdf(par)
        print("C0 point", k, f, f_ref, flush=True)
        assert abs(f - f_ref) <= 1e-10 * abs(f_ref)
        assert np.max(np.abs(g - g_ref)) <= 1e-10 * np.max(np.abs(g_ref))
        assert L.DoEval(par) == f
        # the native timing loop returns the same bits (rpwa_pwalik_time_fdf)
        sec, f2, g2 = L.timeFdF(par, 3)
        assert f2 == f and np.array_equal(g2, g) and sec > 0
    d_data.free()


@pytest.fixture(scope="module")
def c1():
    """C1/C2 workload: generated on the GPU, loaded by device ingest into an
    FP64 and a complex64 likelihood, and the oracle on a host copy."""
    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL
    from rootpwa_amd import synth
    nW, rank, N, M = (8, 80), 2, 2_000_000, 5_000_000
    d_data, sums, truth, pool = synth.generate_sample(20261203, nW, rank, N, M)
    print("C1 data pool", pool, flush=True)
    L64 = _device_likelihood(PL, capi.PREC_FP64, nW, rank, d_data, N, sums, M)
    Lmix = _device_likelihood(PL, capi.PREC_MIXED, nW, rank, d_data, N, sums, M)
    data = d_data.download((N, 88), np.complex128)
    Lo = _oracle(nW, rank, data, sums, M)
    del data
    yield dict(L64=L64, Lmix=Lmix, Lo=Lo, N=N, d_data=d_data)
    d_data.free()


def _start(L, N, seed):
    return np.random.default_rng(seed).normal(size=L.nmbPars()) * np.sqrt(N / 88)


def test_c1_fp64_full_size_and_fit(gpu, c1):
    from rootpwa_amd import likelihood as PL
    L, Lo, N = c1["L64"], c1["Lo"], c1["N"]
    assert L.nmbPars() == 345
    par = _start(L, N, 7)
    f, g = L.FdF(par)
    f_ref, g_ref = Lo.fdf(par, threads=ORACLE_THREADS)
    gscale = np.max(np.abs(g_ref))
    print("C1 start", f, f_ref, abs(f - f_ref) / abs(f_ref), np.max(np.abs(g - g_ref)) / gscale, flush=True)
    assert abs(f - f_ref) <= 1e-10 * abs(f_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-10 * gscale
    # DoEval then Gradient (the Minuit pattern) equals FdF bitwise
    assert L.DoEval(par) == f and np.array_equal(L.Gradient(par), g)
    # a full Minuit-pattern fit from one random start (the drivers draw it from the seed)
    r = PL.pwaFit(L, seed=11)
    x = r.fitParameters
    fm, gm = L.FdF(x)
    fm_ref, gm_ref = Lo.fdf(x, threads=ORACLE_THREADS)
    print("C1 fit: lnL", r.logLikelihood(), "evaluations", r.nmbEvaluations, "converged", r.converged(),
          "oracle", fm_ref, "grad max", np.max(np.abs(gm_ref)), flush=True)
    assert r.logLikelihood() == fm
    assert abs(fm - fm_ref) <= 1e-10 * abs(fm_ref)
    assert np.max(np.abs(gm - gm_ref)) <= 1e-10 * gscale
    assert fm < f
    c1["fp64_minimum"] = x


def test_c2_mixed_full_size(gpu, c1):
    L, Lo, N = c1["Lmix"], c1["Lo"], c1["N"]
    par = _start(L, N, 7)
    f, g = L.FdF(par)
    f_ref, g_ref = Lo.fdf(par, threads=ORACLE_THREADS)
    gscale = np.max(np.abs(g_ref))
    print("C2 start |dlnL|", abs(f - f_ref), "grad", np.max(np.abs(g - g_ref)) / gscale, flush=True)
    assert abs(f - f_ref) < 1e-3
    assert np.max(np.abs(g - g_ref)) <= 1e-5 * gscale
    assert L.DoEval(par) == f
    # at the FP64 fit's minimum (or, if that test did not run, after a fit of its own)
    x = c1.get("fp64_minimum")
    if x is None:
        from rootpwa_amd import likelihood as PL
        x = PL.pwaNloptFit(L, seed=11).fitParameters
    fm, gm = L.FdF(x)
    fm_ref, gm_ref = Lo.fdf(x, threads=ORACLE_THREADS)
    print("C2 minimum |dlnL|", abs(fm - fm_ref), "grad", np.max(np.abs(gm - gm_ref)) / gscale, flush=True)
    assert abs(fm - fm_ref) < 1e-3
    assert np.max(np.abs(gm - gm_ref)) <= 1e-5 * gscale


def _fp32(x):
    x = np.array(x, dtype=np.float64)
    x[..., :-1] = x[..., :-1].astype(np.float32).astype(np.float64)
    return x


def test_c3_batched_full_size(gpu):
    from rootpwa_amd import batch as B
    from rootpwa_amd import synth
    W, S, seed = 88, 64, 20261202 + 3
    sizes = synth.bin_sizes(seed, 48)
    bf = B.BatchFit(W, S)
    synth.fill_batch(bf, seed, sizes, log=lambda m: print(m, flush=True))
    bf.finish()
    starts = _fp32(bf.startValues(seed=2))
    v0, g0 = bf.eval(starts)
    par, val, st, ne = bf.fit(startPar=starts)
    print("C3 fits: status", {int(k): int(c) for k, c in zip(*np.unique(st, return_counts=True))},
          "evaluations per fit", float(ne.mean()), flush=True)
    pts = _fp32(par)
    v, g = bf.eval(pts)
    order = np.argsort(sizes)
    for b in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        n = int(sizes[b])
        d_data, sums, M, _ = synth.generate_bin(synth.bin_seed(seed, b), W, n)
        data = d_data.download((n, W), np.complex128)
        d_data.free()
        Lo = _oracle((0, W), 1, data, sums, M)
        del data
        s = int(np.argmin(val[b]))
        f0_ref, g0_ref = Lo.fdf(starts[b, s], threads=ORACLE_THREADS)
        f_ref, g_ref = Lo.fdf(pts[b, s], threads=ORACLE_THREADS)
        gscale = np.max(np.abs(g0_ref))
        # the batch contracts FP32-rounded production amplitudes: the oracle at the
        # FP64 minimiser point itself differs from the value at its rounding by
        f_x = Lo.doeval(par[b, s])
        print("C3 bin", b, "events", n, "start |dlnL|", abs(v0[b, s] - f0_ref), "minimum |dlnL|",
              abs(v[b, s] - f_ref), "grad at start", np.max(np.abs(g0[b, s] - g0_ref)) / gscale,
              "grad at minimum", np.max(np.abs(g[b, s] - g_ref)) / gscale,
              "|L(x) - L(fl32(x))|", abs(f_x - f_ref), flush=True)
        assert abs(v0[b, s] - f0_ref) < 1e-3
        assert abs(v[b, s] - f_ref) < 1e-3
        assert np.max(np.abs(g0[b, s] - g0_ref)) <= 1e-5 * gscale
        assert np.max(np.abs(g[b, s] - g_ref)) <= 1e-5 * gscale
        assert abs(f_x - f_ref) < 1e-3
        del Lo


def test_c4_network_full_width(gpu):
    """The initializer's network at the configuration's width: forward, loss
    and gradient of the FP32 kernels against the FP64 oracle."""
    from rootpwa_amd.initializer import Initializer
    from rootpwa_amd.mlp import MLP
    shape = Initializer(88).shape
    assert shape == (88 + 2 * 256 + 2 * 88 - 1, 512, 3, 176)
    m = MLP(*shape, seed=7)
    p = m.params()
    layers = MO.unpack(p.astype(np.float64), *shape)
    rng = np.random.default_rng(8)
    for rows in (1, 64, 200):
        x = rng.normal(size=(rows, shape[0])).astype(np.float32)
        t = rng.normal(size=(rows, shape[3])).astype(np.float32)
        y = m.forward(x)
        ref, _ = MO.forward(layers, x.astype(np.float64))
        err_y = float(np.max(np.abs(y - ref)) / np.max(np.abs(ref)))
        loss, gr = m.lossGrad(x, t)
        lref, gref = MO.loss_grad(layers, x.astype(np.float64), t.astype(np.float64))
        err_g = float(np.max(np.abs(gr - gref)) / np.max(np.abs(gref)))
        print("C4 network rows", rows, "forward", err_y, "loss", abs(loss - lref) / lref, "gradient", err_g,
              flush=True)
        assert err_y < 1e-5
        assert abs(loss - lref) / lref < 1e-5
        assert err_g < 1e-4


def test_c4_protocol_88_waves(gpu):
    """The learned-initializer protocol on 88-wave bins (fewer and smaller bins
    than the benchmark's 2000 + 200): moments and features of the
    configuration's size, the spectral estimate aligned with the truth, a
    network trained on the batch's best-of-16 fits, and its starts refined by
    the plain fit against random starts on held-out bins."""
    from rootpwa_amd import initializer as I
    from rootpwa_amd import synth
    W = 88
    train_sizes = synth.bin_sizes(501, 192, lo=20_000, hi=40_000)
    test_sizes = synth.bin_sizes(502, 24, lo=20_000, hi=40_000)
    data = I.fitted_bins(5001, train_sizes, W, nStarts=16, fitSeed=5)
    init = I.Initializer(W, seed=3)
    init.train(data["moments"], data["best"], data["nEvents"], epochs=200, batchSize=32, lr=1e-3)
    X = init.features(data["moments"][:4])
    assert X.shape == (4, init.shape[0])
    # the spectral estimate against the best fit (the global phase is fixed by the anchor)
    spec = init.spectral(data["moments"])
    best = np.asarray(data["best"])[:, :-1]
    cos = np.abs(np.sum(spec[:, :-1] * best, axis=1)) / (np.linalg.norm(spec[:, :-1], axis=1) *
                                                           np.linalg.norm(best, axis=1))
    print("C4 spectral cosine median", float(np.median(cos)), "validation loss", init.validationLoss, flush=True)
    assert np.median(cos) > 0.99
    assert init.validationLoss < 1.0
    rep = I.evaluate(init, 6002, test_sizes, W, nRandom=16, nMixed=4, fitSeed=9)
    print({k: v for k, v in rep.items() if k != "per_bin"}, flush=True)
    assert rep["start_gap_median"]["mlp"] < 0.2 * rep["start_gap_median"]["random"]
    assert rep["mlp_start_reaches_best"] >= 0.8
    assert rep["evaluations_per_fit"]["mlp_start"] < rep["evaluations_per_fit"]["random_start"]
