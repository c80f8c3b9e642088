"""The learned initializer: the network's GPU kernels against the FP64 oracle
(oracle/mlp_oracle.py), training determinism, and what the trained network
gains over random starts on held-out bins (small shapes).

Tolerances: FP32 kernels vs FP64 oracle, relative to the largest magnitude:
forward 1e-5, loss 1e-5, gradient 1e-4."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

SHAPE = (13, 24, 2, 5)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def net():
    from rootpwa_amd.mlp import MLP
    m = MLP(*SHAPE, seed=5)
    rng = np.random.default_rng(3)
    p = (rng.normal(size=m.nParams()) * 0.5).astype(np.float32)
    m.setParams(p)
    return m, p


def test_parameter_layout_and_init():
    from rootpwa_amd.mlp import MLP
    m = MLP(*SHAPE, seed=1)
    n = sum(i * o + o for i, o in M.shapes(*SHAPE))
    assert m.nParams() == n
    p = m.params()
    layers = M.unpack(p.astype(np.float64), *SHAPE)
    for (W, b), (i, o) in zip(layers, M.shapes(*SHAPE)):
        assert np.all(b == 0) and np.max(np.abs(W)) <= np.sqrt(6.0 / (i + o))  # Glorot uniform
    assert np.array_equal(MLP(*SHAPE, seed=1).params(), p) and not np.array_equal(MLP(*SHAPE, seed=2).params(), p)


def test_forward_matches_oracle(gpu, net):
    m, p = net
    rng = np.random.default_rng(4)
    for rows in (1, 7, 64, 131):
        x = rng.normal(size=(rows, SHAPE[0])).astype(np.float32)
        y = m.forward(x)
        ref, _ = M.forward(M.unpack(p.astype(np.float64), *SHAPE), x.astype(np.float64))
        assert rel(y, ref) < 1e-5


def test_loss_and_gradient_match_oracle(gpu, net):
    m, p = net
    rng = np.random.default_rng(5)
    for rows in (3, 64, 100):
        x = rng.normal(size=(rows, SHAPE[0])).astype(np.float32)
        t = rng.normal(size=(rows, SHAPE[3])).astype(np.float32)
        loss, g = m.lossGrad(x, t)
        lref, gref = M.loss_grad(M.unpack(p.astype(np.float64), *SHAPE), x.astype(np.float64), t.astype(np.float64))
        print("rows", rows, "loss", loss, lref, "grad rel", rel(g, gref))
        assert abs(loss - lref) / lref < 1e-5
        assert rel(g, gref) < 1e-4


def test_training_reduces_loss_deterministically(gpu):
    from rootpwa_amd.mlp import MLP
    rng = np.random.default_rng(6)
    x = rng.normal(size=(512, 10))
    A = rng.normal(size=(10, 6)) / 3
    t = np.tanh(x @ A) @ rng.normal(size=(6, 4))
    runs = []
    for _ in range(2):
        m = MLP(10, 32, 2, 4, seed=9)
        hist = m.train(x, t, batchSize=64, epochs=60, lr=3e-3, seed=4)
        runs.append((hist, m.params()))
    print("loss", runs[0][0][0], "->", runs[0][0][-1])
    assert runs[0][0][-1] < 0.1 * runs[0][0][0]
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


def test_initializer_gains_over_random_starts(gpu):
    """Small version of the learned-initializer protocol: 8 waves, bins of
    2k-8k events, best-of-16 training fits on 400 bins, 32 held-out bins.
    The network's start must lie at least 5x closer to the best minimum (in
    -lnL, median over bins) than random starts do, its fit must reach the best
    minimum in at least 90% of the bins, and it must need fewer evaluations
    than the random starts' fits do on average."""
    from rootpwa_amd import initializer as I
    from rootpwa_amd import synth
    W = 8
    train_sizes = synth.bin_sizes(101, 400, lo=2000, hi=8000)
    test_sizes = synth.bin_sizes(202, 32, lo=2000, hi=8000)
    data = I.fitted_bins(1001, train_sizes, W, nStarts=16, fitSeed=5)
    init = I.Initializer(W, nPairs=16, nHidden=64, nHiddenLayers=3, seed=3)
    hist = init.train(data["moments"], data["best"], data["nEvents"], epochs=300, batchSize=32, lr=1e-3)
    print("training loss", hist[0], "->", hist[-1], "validation", init.validationLoss)
    assert init.validationLoss < 1.0  # the correction beats its mean
    rep = I.evaluate(init, 2002, test_sizes, W, nRandom=16, nMixed=4, fitSeed=9)
    print({k: v for k, v in rep.items() if k != "per_bin"})
    assert rep["start_gap_median"]["mlp"] < 0.2 * rep["start_gap_median"]["random"]
    assert rep["mlp_start_reaches_best"] >= 0.9
    assert rep["evaluations_per_fit"]["mlp_start"] < rep["evaluations_per_fit"]["random_start"]
