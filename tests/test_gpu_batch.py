"""Batched fits (include/rootpwa_amd_batch.h) against the oracle.

Every bin of a batch is pwaLikelihood with rank 1 and all waves in the refl +1
sector, so the oracle for bin b is OracleLikelihood((0, W), 1) on the bin's
data and integral.  The batch stores amplitudes as complex64 and contracts on
FP32 MFMA with the production amplitudes rounded to FP32; the test points are
therefore FP32-representable, so the oracle sees exactly the production
amplitudes the kernel uses and the differences measure only the complex64
amplitude storage and the FP32 contraction.  Tolerances: |d lnL| <= min(1e-6 (sqrt(N W / 8) + 1), 1e-3)
(absolute, per fit, N data events, W waves: the FP32 rounding of each event's
sigma grows as the root of the 2 W summed terms and the events' errors add as
a random walk) and gradient max-abs error <= 1e-5 of the gradient's
max-abs value."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def make_bins(W, sizes, seed, nMC_factor=5):
    """Per bin: dataset (refl +1 sector only), oracle likelihood."""
    bins = []
    for k, n in enumerate(sizes):
        ds = H.dataset((0, W), 1, n, max(nMC_factor * n, 2000), seed + 17 * k)
        Lo = H.oracle_likelihood(ds)
        bins.append((ds, Lo))
    return bins


def make_batch(W, S, bins):
    from rootpwa_amd import batch as B
    bf = B.BatchFit(W, S)
    for ds, _ in bins:
        bf.addBin(ds["data"], ds["integral"], ds["nMC"])
    bf.finish()
    return bf


def fp32_params(rng, nb, S, P, scale):
    par = rng.normal(size=(nb, S, P)) * scale
    par[..., :-1] = par[..., :-1].astype(np.float32).astype(np.float64)
    return par


def f_tol(n, W):
    # never looser than the 1e-3 the single-fit complex64 path is held to
    return min(1e-6 * (np.sqrt(n * W / 8.0) + 1.0), 1e-3)


def check_against_oracle(bf, bins, par, starts=None):
    v, g = bf.eval(par)
    W = par.shape[2] // 2
    worst_f, worst_g = 0.0, 0.0
    for bi, (ds, Lo) in enumerate(bins):
        n = Lo.nmb_events()
        for s in (range(par.shape[1]) if starts is None else starts):
            f_ref, g_ref = Lo.fdf(par[bi, s])
            df = abs(v[bi, s] - f_ref)
            dg = np.max(np.abs(g[bi, s] - g_ref)) / np.max(np.abs(g_ref))
            print("bin", bi, "N", n, "start", s, "f", v[bi, s], "df", df, "df/tol", df / f_tol(n, W), "grad rel", dg)
            worst_f = max(worst_f, df / f_tol(n, W))
            worst_g = max(worst_g, dg)
    print("worst df/tol", worst_f, "worst grad rel", worst_g)
    assert worst_f <= 1.0
    assert worst_g <= 1e-5
    return v, g


@pytest.fixture(scope="module")
def small():
    bins = make_bins(8, [37, 1000, 2500, 64], seed=300)
    return bins, make_batch(8, 5, bins)


def test_eval_matches_oracle_small(gpu, small):
    bins, bf = small
    rng = np.random.default_rng(1)
    for scale in (0.5, 3.0, 20.0):
        check_against_oracle(bf, bins, fp32_params(rng, len(bins), 5, 16, scale))


def test_eval_matches_oracle_88_waves(gpu):
    bins = make_bins(88, [20000, 3001], seed=310)
    bf = make_batch(88, 64, bins)
    print(bf.info())
    rng = np.random.default_rng(2)
    par = fp32_params(rng, len(bins), 64, 176, 2.0)
    # a subset of the starts against the oracle, all of them for determinism
    v, g = check_against_oracle(bf, bins, par, starts=(0, 17, 38, 63))
    v2, g2 = bf.eval(par)
    assert np.array_equal(v, v2) and np.array_equal(g, g2)


def test_sign_convention_and_determinism(gpu, small):
    bins, bf = small
    rng = np.random.default_rng(3)
    par = fp32_params(rng, len(bins), 5, 16, 2.0)
    v, g = bf.eval(par)
    flipped = par.copy()
    flipped[..., :-1] *= -1
    flipped[..., -1] *= -1
    v2, _ = bf.eval(flipped)
    assert np.array_equal(v, v2)  # the contraction is exactly odd in T
    v3, g3 = bf.eval(par)
    assert np.array_equal(v, v3) and np.array_equal(g, g3)


def test_start_values(gpu, small):
    """start (bin i, start s) = the drivers' draws from MT19937 seeded with
    splitmix64(seed * 2^32 + i * 2^16 + s) (32 low bits)."""
    bins, bf = small
    sv = bf.startValues(seed=9)

    def splitmix64(x):
        m = (1 << 64) - 1
        x = (x + 0x9E3779B97F4A7C15) & m
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
        return x ^ (x >> 31)

    for bi, (ds, Lo) in enumerate(bins):
        for s in range(5):
            seed = splitmix64((9 << 32) + (bi << 16) + s) & 0xFFFFFFFF
            raw = np.random.RandomState(seed).randint(0, 2 ** 32, size=80, dtype=np.uint64)
            u = raw[raw != 0].astype(np.float64) * 2.0 ** -32
            sq = np.sqrt(float(Lo.nmb_events()))
            expect = np.array([(0.01 + (sq - 0.01) * u[2 * k]) * (-1 if u[2 * k + 1] > 0.5 else 1)
                               for k in range(16)])
            assert np.array_equal(sv[bi, s], expect)


def oracle_newton(Lo, x, steps=8):
    x = np.array(x, dtype=np.float64)
    for _ in range(steps):
        x -= np.linalg.solve(Lo.hessian(x), Lo.gradient(x))
    return x


def test_batched_fits_reach_the_oracle_minimum(gpu):
    bins = make_bins(8, [3000, 6000, 1500], seed=320)
    bf = make_batch(8, 8, bins)
    par, val, st, ne = bf.fit(seed=4)
    print("status", st.tolist(), "evals", ne.tolist())
    print(bf.info())
    assert np.all(st > 0) or np.mean(st > 0) >= 0.75
    for bi, (ds, Lo) in enumerate(bins):
        best = int(np.argmin(val[bi]))
        x = par[bi, best]
        assert x[0] >= 0 and x[-1] >= 0  # sign convention applied
        xs = oracle_newton(Lo, x)
        cov = np.linalg.inv(Lo.hessian(xs))
        sigma = np.sqrt(np.diag(cov))
        shift = np.max(np.abs(x - xs) / sigma)
        f_star = Lo.doeval(xs)
        print("bin", bi, "best", val[bi, best], "f*", f_star, "shift/sigma", shift)
        assert shift < 0.05
        assert abs(val[bi, best] - f_star) < f_tol(Lo.nmb_events(), 8) + 1e-4
    # the returned values are the batch's values at the returned (sign-corrected) points
    v, _ = bf.eval(par)
    assert np.array_equal(v, val)
