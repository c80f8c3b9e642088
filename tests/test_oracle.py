"""Pins of the CPU oracle (runs without a GPU).

Each check fixes the oracle by something other than its own formula: the
golden vectors computed in 60-digit decimal arithmetic (tests/golden), a
closed form, mathematical identities (Euler homogeneity, phase and sign
invariance, the extended-ML yield identity at a minimum), finite differences,
an independent vectorised numpy formulation, and the reference's own
compensated-sum scenario.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "likelihood_golden.json")


def _cplx(a):
    a = np.asarray(a)
    return a[..., 0] + 1j * a[..., 1]


def golden_cases():
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"]


def oracle_from_golden(c):
    L = O.OracleLikelihood(tuple(c["nW"]), c["rank"])
    if c["cauchy_width"] is not None:
        L.set_prior(True, c["cauchy_width"])
    L.use_normalized_amps(c.get("normalised", True))
    acc = _cplx(c["acc_sums"]) if "acc_sums" in c else None
    L.add_integrals(_cplx(c["integral_sums"]), c["mc_events"], acc, c.get("acc_events"), c.get("acc_override", 0))
    L.add_amplitudes(_cplx(c["data"]))
    L.finish_init()
    return L


@pytest.mark.parametrize("name", list(golden_cases().keys()))
def test_golden_vectors(name):
    c = golden_cases()[name]
    L = oracle_from_golden(c)
    par = np.array(c["par"])
    f, g = L.fdf(par)
    assert abs(f - c["value"]) <= 1e-13 * max(1.0, abs(c["value"]))
    gref = np.array(c["gradient"])
    assert np.max(np.abs(g - gref)) <= 1e-13 * np.max(np.abs(gref))
    assert L.doeval(par) == f
    assert np.array_equal(L.gradient(par), g)
    assert H.rel_err(O.integral(_cplx(c["mc"])), _cplx(c["integral_sums"])) <= 1e-14
    if "acc_sums" in c:
        # the acceptance integral is the integral over the accepted MC subset
        acc = _cplx(c["mc"])[np.array(c["acc_index"], dtype=np.int64)]
        assert H.rel_err(O.integral(acc), _cplx(c["acc_sums"])) <= 1e-14
        assert L.tot_acc() == c["acc_events"] / (c["acc_override"] or c["mc_events"])


@pytest.mark.parametrize("name", list(golden_cases().keys()))
@pytest.mark.parametrize("threads", [2, 3, 7, 16, 64])
def test_threaded_fdf_on_golden_vectors(name, threads):
    """orc_lik_fdf_mt (the checker of the full-size GPU runs) reproduces the
    golden vectors, also with more threads than events (empty ranges)."""
    c = golden_cases()[name]
    L = oracle_from_golden(c)
    par = np.array(c["par"])
    f, g = L.fdf(par, threads=threads)
    assert abs(f - c["value"]) <= 1e-13 * max(1.0, abs(c["value"]))
    gref = np.array(c["gradient"])
    assert np.max(np.abs(g - gref)) <= 1e-13 * np.max(np.abs(gref))


@pytest.mark.parametrize("threads", list(range(1, 17)))
def test_threaded_fdf_equals_single_thread(threads):
    """Range partials merged with their compensations: orc_lik_fdf_mt equals
    orc_lik_fdf to the rounding of the Kahan merge (a few ulp of the sums)."""
    ds = H.dataset((3, 9), 2, 3001, 2000, seed=77)
    L = H.oracle_likelihood(ds)
    rng = np.random.default_rng(threads)
    par = H.random_par(rng, L, 0.7)
    f1, g1 = L.fdf(par)
    fm, gm = L.fdf(par, threads=threads)
    assert abs(fm - f1) <= 4e-16 * 3001 ** 0.5 * abs(f1) + 1e-12
    assert np.max(np.abs(gm - g1)) <= 1e-14 * np.max(np.abs(g1))
    if threads == 1:
        assert fm == f1 and np.array_equal(gm, g1)


def test_threaded_fdf_fewer_events_than_threads():
    ds = H.dataset((2, 3), 1, 5, 200, seed=78)
    L = H.oracle_likelihood(ds)
    par = H.random_par(np.random.default_rng(5), L, 0.7)
    f1, g1 = L.fdf(par)
    for threads in (6, 16):
        fm, gm = L.fdf(par, threads=threads)
        assert abs(fm - f1) <= 1e-15 * abs(f1) + 1e-14
        assert np.max(np.abs(gm - g1)) <= 1e-15 * np.max(np.abs(g1))


def test_kahan_scenario_of_the_reference():
    # test/utilities/testSumAccumulators.cc:46-60: 1e6 additions of 1e-6 in float
    naive, comp = O.kahan_float_demo(1000000, 1e-6)
    assert comp == pytest.approx(1.0, abs=1e-7)
    assert abs(naive - 1.0) > 1e-3


@pytest.mark.parametrize("nW,rank,P", [((0, 16), 1, 32), ((8, 80), 2, 345), ((0, 88), 1, 176), ((7, 34), 2, 157),
                                       ((3, 3), 4, 19)])
def test_parameter_counts(nW, rank, P):
    L = O.OracleLikelihood(nW, rank)
    assert L.nPars == P
    # rank restrictions: zero below the diagonal, real anchor, flat last
    m = L.par_map()
    for r in range(rank):
        for s in range(2):
            for i in range(nW[s]):
                if i < r:
                    assert tuple(m[r, s, i]) == (-1, -1)
                elif i == r:
                    assert m[r, s, i, 0] >= 0 and m[r, s, i, 1] == -1
                else:
                    assert m[r, s, i, 1] == m[r, s, i, 0] + 1
    assert L.par_info(P - 1)[1] == -1  # flat wave
    rng = np.random.default_rng(0)
    par = rng.normal(size=P)
    prod, flat = L.copy_from_par(par)
    assert np.array_equal(L.copy_to_par(prod, flat), par)


def test_thresholds_fix_parameters():
    nW = (2, 3)
    thr = np.zeros((2, 3))
    thr[1, 2] = 2.0   # above the bin centre -> fixed
    thr[0, 1] = 1.0   # below -> free
    L = O.OracleLikelihood(nW, 1, thresholds=thr, massBinCenter=1.5)
    fixed = [L.par_info(i)[3] for i in range(L.nPars)]
    assert sum(fixed) == 2 and L.nmb_pars_fixed() == 2


def test_closed_form_one_wave():
    # one wave, rank 1: -lnL = -sum ln(T^2 |A|^2 + f^2) + (T^2 I + f^2 eta)
    rng = np.random.default_rng(1)
    a = rng.normal(size=50) + 1j * rng.normal(size=50)
    mc = rng.normal(size=400) + 1j * rng.normal(size=400)
    Isum = np.sum(np.abs(mc) ** 2)
    L = O.OracleLikelihood((0, 1), 1)
    L.add_integrals(np.array([[Isum]]), 400)
    L.add_amplitudes(a[:, None])
    L.finish_init()
    T, f = 0.7, 0.25
    An = a / np.sqrt(Isum / 400)
    expect = -np.sum(np.log(T * T * np.abs(An) ** 2 + f * f)) + (T * T + f * f)
    assert L.doeval(np.array([T, f])) == pytest.approx(expect, rel=1e-14)
    # non-normalised amplitudes: nu = N and no rescaling
    L2 = O.OracleLikelihood((0, 1), 1)
    L2.use_normalized_amps(False)
    L2.add_integrals(np.array([[Isum]]), 400)
    L2.add_amplitudes(a[:, None])
    L2.finish_init()
    expect2 = -np.sum(np.log(T * T * np.abs(a) ** 2 + f * f)) + 50 * (T * T * Isum / 400 + f * f)
    assert L2.doeval(np.array([T, f])) == pytest.approx(expect2, rel=1e-14)


@pytest.mark.parametrize("override", [0, 1000])
def test_closed_form_one_wave_acceptance(override):
    """One wave, rank 1, acceptance integral over an accepted subset
    (addAccIntegral, pwaLikelihood.cc:1088-1107): with n = accEventsOverride
    or the MC count, eta = N_acc / n and
    -lnL = -sum ln(T^2 |A|^2 / I_n + f^2) + T^2 (S_acc / n) / I_n + f^2 eta,
    I_n = S_mc / M; computed here from the raw sums."""
    rng = np.random.default_rng(11)
    a = rng.normal(size=40) + 1j * rng.normal(size=40)
    mc = rng.normal(size=600) + 1j * rng.normal(size=600)
    accepted = rng.random(600) < 0.3
    S_mc = np.sum(np.abs(mc) ** 2)
    S_acc = np.sum(np.abs(mc[accepted]) ** 2)
    n_acc = int(accepted.sum())
    L = O.OracleLikelihood((0, 1), 1)
    L.add_integrals(np.array([[S_mc]]), 600, np.array([[S_acc]]), n_acc, override)
    L.add_amplitudes(a[:, None])
    L.finish_init()
    T, f = 1.3, 0.4
    n = override or 600
    I_n = S_mc / 600
    eta = n_acc / n
    expect = -np.sum(np.log(T * T * np.abs(a) ** 2 / I_n + f * f)) + T * T * (S_acc / n) / I_n + f * f * eta
    assert L.doeval(np.array([T, f])) == pytest.approx(expect, rel=1e-14)
    # gradient: d/dT = sum -2 T |A|^2/I_n / sigma + 2 T (S_acc/n)/I_n, d/df = sum -2 f / sigma + 2 f eta
    sigma = T * T * np.abs(a) ** 2 / I_n + f * f
    g = L.gradient(np.array([T, f]))
    assert g[0] == pytest.approx(np.sum(-2 * T * np.abs(a) ** 2 / I_n / sigma) + 2 * T * S_acc / n / I_n, rel=1e-13)
    assert g[1] == pytest.approx(np.sum(-2 * f / sigma) + 2 * f * eta, rel=1e-13)


def test_intensity_closed_form_and_brute_force():
    """orc_intensity (modelIntensity::getIntensity): one wave, rank 1:
    |T|^2 |A|^2 / n^2; several waves and ranks: an independent einsum."""
    rng = np.random.default_rng(12)
    a = rng.normal(size=(30, 1)) + 1j * rng.normal(size=(30, 1))
    T = np.array([[[0.0], [0.8 - 0.3j]]])
    got = O.intensity(a, (0, 1), T, waveNorm=np.array([1.7]))
    assert np.allclose(got, abs(0.8 - 0.3j) ** 2 * np.abs(a[:, 0]) ** 2 / 1.7 ** 2, rtol=1e-14, atol=0)
    nW, rank = (3, 5), 3
    a = rng.normal(size=(50, 8)) + 1j * rng.normal(size=(50, 8))
    prod = H.random_prod(rng, rank, nW)
    n = rng.uniform(0.5, 2.0, size=8)
    an = a / n
    expect = np.zeros(50)
    for s, cols in ((0, slice(0, 3)), (1, slice(3, 8))):
        S = np.einsum("ri,ei->er", prod[:, s, :nW[s]], an[:, cols])
        expect += np.sum(np.abs(S) ** 2, axis=1)
    assert np.allclose(O.intensity(a, nW, prod, n), expect, rtol=1e-13, atol=0)
    # ranks add up; a global phase per (rank, reflectivity) block changes nothing
    per_rank = sum(O.intensity(a, nW, prod[r:r + 1], n) for r in range(rank))
    assert np.allclose(O.intensity(a, nW, prod, n), per_rank, rtol=1e-14, atol=0)
    rot = prod.copy()
    rot[1, 0] *= np.exp(0.7j)
    rot[2, 1] *= np.exp(-2.1j)
    assert np.allclose(O.intensity(a, nW, rot, n), expect, rtol=1e-13, atol=0)


def test_intensity_is_the_likelihood_event_term(small):
    """With the likelihood's normalisation (raw amplitudes divided by
    sqrt(Re I_ii / M)), sum_e -ln(intensity_e + f^2) is the golden-pinned
    likelihood oracle's event part."""
    ds, L = small
    prod = H.random_prod(np.random.default_rng(4), 2, ds["nW"], 0.7)
    norm = np.sqrt(np.real(np.diag(ds["integral"])) / ds["nMC"])
    inten = O.intensity(ds["data"], ds["nW"], prod, norm)
    assert -np.sum(np.log(inten + 0.3 ** 2)) == pytest.approx(L.event_loglik(prod, 0.3), rel=1e-13)
    # the same from the likelihood's own normalised amplitudes without a divisor
    inten2 = O.intensity(H.normalised_amps(L), ds["nW"], prod)
    assert np.allclose(inten2, inten, rtol=1e-13, atol=0)


def numpy_neg_lnl(L, prod, flat):
    """Independent vectorised formulation (einsum over the sector blocks)."""
    A = [L.decay_amps(0), L.decay_amps(1)]
    Iacc = L.acc_matrix()
    sigma = np.full(L.nmb_events(), flat * flat)
    norm = flat * flat * L.tot_acc()
    for s in range(2):
        n = L.nW[s]
        if n == 0:
            continue
        S = np.einsum("ri,ei->er", prod[:, s, :n], A[s])
        sigma += np.sum(np.abs(S) ** 2, axis=1)
        norm += np.real(np.einsum("ri,rj,ij->", prod[:, s, :n], np.conj(prod[:, s, :n]), Iacc[s, :n, s, :n]))
    return -np.sum(np.log(sigma)) + norm


@pytest.fixture(scope="module")
def small():
    ds = H.dataset((3, 5), 2, 400, 2000, seed=21)
    return ds, H.oracle_likelihood(ds)


def test_matches_vectorised_formulation(small):
    ds, L = small
    rng = np.random.default_rng(2)
    for _ in range(3):
        par = rng.normal(size=L.nPars)
        prod, flat = L.copy_from_par(par)
        assert L.doeval(par) == pytest.approx(numpy_neg_lnl(L, prod, flat), rel=1e-12)


def test_gradient_finite_differences(small):
    ds, L = small
    rng = np.random.default_rng(3)
    par = rng.normal(size=L.nPars)
    g = L.gradient(par)
    h = 1e-6
    num = np.array([(L.doeval(par + h * e) - L.doeval(par - h * e)) / (2 * h) for e in np.eye(L.nPars)])
    assert np.max(np.abs(num - g)) <= 1e-6 * np.max(np.abs(g))


def test_euler_identity(small):
    # sigma_e and the normalisation term are homogeneous of degree 2:
    # sum_k par_k dL/dpar_k = -2N + 2 nu NormTerm
    ds, L = small
    rng = np.random.default_rng(4)
    par = rng.normal(size=L.nPars)
    f, g = L.fdf(par)
    prod, flat = L.copy_from_par(par)
    event = L.event_loglik(prod, flat)
    norm = f - event
    assert np.dot(par, g) == pytest.approx(-2 * L.nmb_events() + 2 * norm, rel=1e-11)


def test_phase_and_sign_invariance(small):
    ds, L = small
    rng = np.random.default_rng(5)
    par = rng.normal(size=L.nPars)
    prod, flat = L.copy_from_par(par)
    f0 = L.event_loglik(prod, flat)
    for r in range(2):
        for s in range(2):
            p = prod.copy()
            p[r, s] *= np.exp(1j * 0.7)
            assert L.event_loglik(p, flat) == pytest.approx(f0, rel=1e-13)
    # CorrectParamSigns flips (rank, sector) blocks whose real anchor is negative: bitwise equal
    q = L.correct_param_signs(par)
    pq, fq = L.copy_from_par(q)
    assert fq >= 0
    for r in range(2):
        for s in range(2):
            assert not (pq[r, s, r].real < 0 and pq[r, s, r].imag == 0)
    assert L.doeval(q) == L.doeval(par)


def test_doderivative_cache(small):
    ds, L = small
    par = np.random.default_rng(6).normal(size=L.nPars)
    f, g = L.fdf(par)
    assert all(L.doderivative(par, k) == g[k] for k in range(0, L.nPars, 7))
    par2 = par + 0.01
    g2 = L.gradient(par2)
    assert L.doderivative(par2, 3) == g2[3]


def test_half_cauchy_prior_closed_form(small):
    ds, _ = small
    L = H.oracle_likelihood(ds, prior=0.6)
    Lf = H.oracle_likelihood(ds)
    par = np.random.default_rng(7).normal(size=L.nPars)
    prod, flat = L.copy_from_par(par)
    r2 = np.abs(prod[prod != 0]) ** 2
    prior = np.sum(np.log(1 + r2 / 0.36))
    assert L.doeval(par) - Lf.doeval(par) == pytest.approx(prior, rel=1e-12)
    g = L.gradient(par)
    h = 1e-6
    num = np.array([(L.doeval(par + h * e) - L.doeval(par - h * e)) / (2 * h) for e in np.eye(L.nPars)])
    assert np.max(np.abs(num - g)) <= 1e-6 * np.max(np.abs(g))


def test_hessian_finite_differences(small):
    ds, L = small
    for prior in (None, 0.5):
        Lp = H.oracle_likelihood(ds, prior=prior) if prior else L
        par = np.random.default_rng(8).normal(size=Lp.nPars)
        Hm = Lp.hessian(par)
        assert np.allclose(Hm, Hm.T, rtol=1e-10, atol=1e-9 * np.max(np.abs(Hm)))
        h = 1e-5
        num = np.array([(Lp.gradient(par + h * e) - Lp.gradient(par - h * e)) / (2 * h) for e in np.eye(Lp.nPars)])
        assert np.max(np.abs(num - Hm)) <= 1e-5 * np.max(np.abs(Hm))


def test_yield_identity_at_minimum():
    # at an extended-ML minimum with normalised amplitudes, NormTerm = N
    from scipy.optimize import minimize
    ds = H.dataset((0, 4), 1, 600, 3000, seed=31)
    L = H.oracle_likelihood(ds)
    x0 = np.full(L.nPars, 3.0)
    res = minimize(lambda p: L.fdf(p), x0, jac=True, method="L-BFGS-B", options=dict(gtol=1e-9, maxiter=2000))
    x = res.x
    for _ in range(6):  # Newton polish with the analytic Hessian
        x = x - np.linalg.solve(L.hessian(x), L.gradient(x))
    f, g = L.fdf(x)
    assert np.max(np.abs(g)) < 1e-8
    prod, flat = L.copy_from_par(x)
    norm = f - L.event_loglik(prod, flat)
    assert norm == pytest.approx(L.nmb_events(), rel=1e-10)


def test_integral_properties():
    rng = np.random.default_rng(9)
    a = rng.normal(size=(500, 6)) + 1j * rng.normal(size=(500, 6))
    I = O.integral(a)
    assert np.allclose(I, I.conj().T, rtol=0, atol=1e-12 * np.max(np.abs(I)))
    assert np.all(np.diag(I).real > 0) and np.allclose(np.diag(I).imag, 0)
    assert np.min(np.linalg.eigvalsh(I)) > 0
    # I_ij = sum_e A_i conj(A_j)
    assert H.rel_err(I, a.T @ np.conj(a)) <= 1e-13
    # weights 1 reduce to the plain sum; constant weights cancel in the renormalisation
    assert H.rel_err(O.integral(a, np.ones(500)), I) <= 1e-14
    assert H.rel_err(O.integral(a, np.full(500, 3.0)), I) <= 1e-14
    # additive over event subsets
    assert H.rel_err(O.integral(a[:200]) + O.integral(a[200:]), I) <= 1e-13
    # importance weights: sum (1/w) A A^H renormalised by mean(1/w)
    w = rng.uniform(0.5, 2, size=500)
    ref = (a.T * (1 / w)) @ np.conj(a) / np.mean(1 / w)
    assert H.rel_err(O.integral(a, w), ref) <= 1e-13


def test_integral_incoherent_sub_amplitudes():
    """ampIntegralMatrix.cc:432-466: sub-amplitudes of one event add incoherently."""
    rng = np.random.default_rng(19)
    a = rng.normal(size=(2, 400, 5)) + 1j * rng.normal(size=(2, 400, 5))
    w = rng.uniform(0.5, 2, size=400)
    # one sub-amplitude is the plain integral
    assert np.array_equal(O.integral(a[:1]), O.integral(a[0]))
    assert np.array_equal(O.integral(a[:1], w), O.integral(a[0], w))
    # a zero second sub-amplitude changes nothing; a duplicated one doubles the sums exactly
    z = np.stack([a[0], np.zeros_like(a[0])])
    assert np.array_equal(O.integral(z, w), O.integral(a[0], w))
    assert np.array_equal(O.integral(np.stack([a[0], a[0]]), w), 2 * O.integral(a[0], w))
    # sum over events and sub-amplitudes of (1/w) A_i conj(A_j), renormalised by mean(1/w) over events
    ref = np.einsum("sei,sej->ij", a * (1 / w)[None, :, None], np.conj(a)) / np.mean(1 / w)
    I = O.integral(a, w)
    assert H.rel_err(I, ref) <= 1e-13
    assert H.rel_err(I, O.integral(a[0], w) + O.integral(a[1], w)) <= 1e-13


def test_zero_wave_is_fixed():
    # all-zero amplitudes of a wave whose acceptance integral is zero fix its parameters
    # (pwaLikelihood.cc:1248-1282), without changing nmbParsFixed
    rng = np.random.default_rng(10)
    a = rng.normal(size=(50, 3)) + 1j * rng.normal(size=(50, 3))
    a[:, 2] = 0
    mc = rng.normal(size=(300, 3)) + 1j * rng.normal(size=(300, 3))
    mc[:, 2] = 0
    I = O.integral(mc)
    I[2, 2] = 1.0  # normalisation integral non-zero, acceptance zero
    acc = O.integral(mc)
    L = O.OracleLikelihood((0, 3), 1)
    L.add_integrals(I, 300, acc, 300)
    L.add_amplitudes(a)
    fixed = [L.par_info(i)[3] for i in range(L.nPars)]
    assert fixed == [0, 0, 0, 1, 1, 0] and L.nmb_pars_fixed() == 0


def test_generator_restatement():
    seed = 5
    s = O.gen_wave_scales(seed, 1000)
    assert np.all((s >= 0.1) & (s <= 1.0)) and abs(np.mean(s) - 0.55) < 0.02
    a = O.gen_amps(seed, 3, 0, 20000, np.ones(4))
    # lognormal(0, 0.5) per event times complex normals: E|A|^2 = 2 E[exp(N(0,1))] = 2 e^0.5
    assert np.mean(np.abs(a) ** 2) == pytest.approx(2 * np.exp(0.5), rel=0.03)
    # counter-based: any sub-range regenerates identically
    assert np.array_equal(O.gen_amps(seed, 3, 500, 10, np.ones(4)), a[500:510])
    # accept-reject keeps pool order and draws events proportional to the intensity
    prod = O.gen_prod_amps(seed, 2, 1, (0, 4))
    data, idx, pool, smax = O.gen_accept_reject(seed, 4, 300, (0, 4), 1, prod, np.ones(4))
    assert np.all(np.diff(idx.astype(np.int64)) > 0)
    sig = np.abs(data @ prod[0, 1]) ** 2
    pool_a = O.gen_amps(seed, 4, 0, 20000, np.ones(4))
    assert np.mean(sig) > 1.5 * np.mean(np.abs(pool_a @ prod[0, 1]) ** 2)


# ---- the learned initializer's network (oracle/mlp_oracle.py) ---------------------------------

def test_mlp_oracle_gelu_closed_forms():
    from oracle import mlp_oracle as M
    import math
    z = np.array([-30.0, -1.0, 0.0, 0.5, 1.0, 30.0])
    g = M.gelu(z)
    # GELU(z) = z Phi(z): Phi(0) = 1/2, Phi(1) = 0.841344746068543, tails 0 and 1
    assert g[2] == 0.0 and abs(g[0]) < 1e-150 and g[5] == 30.0
    assert g[4] == pytest.approx(0.841344746068542948585, rel=1e-15)
    assert g[1] == pytest.approx(-1.0 * (1.0 - 0.841344746068542948585), rel=1e-14)
    # derivative: Phi(z) + z phi(z); at 0 it is 1/2, and it matches central differences
    assert M.gelu_grad(np.array([0.0]))[0] == 0.5
    h = 1e-6
    zz = np.linspace(-4, 4, 33)
    fd = (M.gelu(zz + h) - M.gelu(zz - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(zz))) < 1e-9
    assert math.isclose(float(M.gelu(np.array([0.5]))[0]), 0.5 * 0.691462461274013103637, rel_tol=1e-15)


def test_mlp_oracle_gradient_against_finite_differences():
    """The hand-written backward pass against central differences of the loss, for every parameter
    of a small network (a dropped term, a transposed product or a wrong activation derivative fails)."""
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(11)
    shape = (5, 7, 2, 3)
    n = sum(i * o + o for i, o in M.shapes(*shape))
    p = rng.normal(size=n) * 0.7
    x = rng.normal(size=(6, 5))
    t = rng.normal(size=(6, 3))
    layers = M.unpack(p, *shape)
    loss, g = M.loss_grad(layers, x, t)
    assert loss == pytest.approx(M.loss_only(layers, x, t), rel=1e-15)
    h = 1e-6
    fd = np.zeros(n)
    for k in range(n):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd[k] = (M.loss_only(M.unpack(pp, *shape), x, t) - M.loss_only(M.unpack(pm, *shape), x, t)) / (2 * h)
    assert np.max(np.abs(fd - g)) / np.max(np.abs(g)) < 1e-8


def test_mlp_oracle_linear_limit():
    """With no hidden layer the network is y = x W + b and the MSE gradient is the normal-equation residual."""
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(12)
    x = rng.normal(size=(9, 4))
    t = rng.normal(size=(9, 2))
    p = rng.normal(size=4 * 2 + 2)
    layers = M.unpack(p, 4, 1, 0, 2)
    W, b = layers[0]
    loss, g = M.loss_grad(layers, x, t)
    r = x @ W + b - t
    assert loss == pytest.approx(np.sum(r ** 2) / r.size, rel=1e-14)
    assert np.allclose(g[:8].reshape(4, 2), 2 * x.T @ r / r.size, rtol=1e-13)
    assert np.allclose(g[8:], 2 * r.sum(axis=0) / r.size, rtol=1e-13)
