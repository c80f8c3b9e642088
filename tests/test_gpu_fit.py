"""The fit drivers (pwaNloptFit / pwaFit flows) on the GPU likelihood.

Minimiser trajectories cannot match NLopt's or Minuit2's (neither library is
available), so parity is judged on the minimum: the oracle's Newton iteration
started from the GPU result must not move it by more than a small fraction of
the statistical error, and the oracle's gradient there must vanish at that
scale.  The drivers' own contracts are checked exactly: the returned minimum is
bitwise DoEval of the sign-corrected parameters, fixed parameters are zero, the
covariance is the inverse Hessian, the eigenvalues are the Hessian's."""
import numpy as np
import pytest

from tests import helpers as H
from tests.test_gpu_pwalik import build_pair

pytestmark = pytest.mark.gpu


def oracle_newton(Lo, x, free, steps=8):
    """Newton iteration on the free parameters with the oracle's analytic Hessian."""
    x = np.array(x, dtype=np.float64)
    for _ in range(steps):
        g = Lo.gradient(x)[free]
        Hm = Lo.hessian(x)[np.ix_(free, free)]
        x[free] -= np.linalg.solve(Hm, g)
    return x


def check_minimum(Lo, Lg, res, tol_sigma):
    P = Lg.nmbPars()
    fixed = np.array([Lg.parameter(i).fixed() for i in range(P)], dtype=bool)
    free = np.flatnonzero(~fixed)
    x = res.fitParameters
    assert np.all(x[fixed] == 0.0)
    # the minimum the driver reports is bitwise the value at its parameters
    assert Lg.DoEval(x) == res.logLikelihood()
    xs = oracle_newton(Lo, x, free)
    Hm = Lo.hessian(xs)
    cov = np.linalg.inv(Hm[np.ix_(free, free)])
    sigma = np.sqrt(np.diag(cov))
    assert np.all(np.isfinite(sigma))
    shift = np.max(np.abs(x[free] - xs[free]) / sigma)
    f_star = Lo.doeval(xs)
    print("max shift / sigma", shift, "f - f*", res.logLikelihood() - f_star, "evals", res.nmbEvaluations,
          "status", res.status)
    assert shift < tol_sigma
    assert res.logLikelihood() - f_star < 0.5 * tol_sigma ** 2 + 1e-9 * abs(f_star)
    assert res.logLikelihood() >= f_star - 1e-9 * abs(f_star)
    return xs, free


@pytest.fixture(scope="module")
def rank1():
    return build_pair((2, 5), 1, seed=41, nData=6000, nMC=30000)


@pytest.fixture(scope="module")
def rank2_thresholds():
    from rootpwa_amd import likelihood as PL
    ds = H.dataset((3, 5), 2, 5000, 30000, seed=43)
    waves = [("neg00", -1, 0.0), ("neg01", -1, 0.0), ("neg02", -1, 4.0),
             ("pos00", +1, 0.0), ("pos01", +1, 0.0), ("pos02", +1, 0.0), ("pos03", +1, 9.0), ("pos04", +1, 0.0)]
    from oracle import oracle as O
    thr = np.array([[0.0, 0.0, 4.0, 0.0, 0.0], [0.0, 0.0, 0.0, 9.0, 0.0]])  # [refl][wave], padded
    Lo = O.OracleLikelihood((3, 5), 2, thresholds=thr, massBinCenter=2.0)
    Lo.add_integrals(ds["integral"], ds["nMC"])
    Lo.add_amplitudes(ds["data"])
    Lo.finish_init()
    names = [w[0] for w in waves]
    I = PL.ampIntegralMatrix(names)
    I.setSums(ds["integral"], ds["nMC"])
    amps = {names[k]: ds["data"][:, k] for k in range(len(names))}
    Lg = PL.initLikelihood(waves, amps, I, rank=2, massBinCenter=2.0)
    return ds, Lo, Lg


def test_nlopt_flow_reaches_the_oracle_minimum(gpu, rank1):
    from rootpwa_amd import likelihood as PL
    ds, Lo, Lg = rank1
    res = PL.pwaNloptFit(Lg, seed=12345, checkHessian=True)
    assert res.status > 0 and res.converged() and res.covMatrixValid()
    check_minimum(Lo, Lg, res, tol_sigma=0.05)
    # covariance = inverse Hessian at the result; eigenvalues of that Hessian
    Hm = Lg.Hessian(res.fitParameters)
    assert H.rel_err(res.covMatrix, Lg.CovarianceMatrix(res.fitParameters)) <= 1e-12
    ev = np.linalg.eigvalsh(Hm)[::-1]
    assert np.all(res.hessianEigenValues > 0)
    assert np.max(np.abs(res.hessianEigenValues - ev)) <= 1e-9 * np.max(np.abs(ev))


def test_minuit_flow_converges_tighter(gpu, rank1):
    from rootpwa_amd import likelihood as PL
    ds, Lo, Lg = rank1
    calls0 = Lg.callInfo()
    res = PL.pwaFit(Lg, seed=777)
    calls1 = Lg.callInfo()
    assert res.status > 0 and res.converged() and res.covMatrixValid()
    check_minimum(Lo, Lg, res, tol_sigma=2e-3)
    # Minuit's call pattern: DoEval calls, a Gradient call after each accepted step
    assert calls1["DoEval"][0] - calls0["DoEval"][0] >= res.nmbEvaluations
    assert calls1["Gradient"][0] > calls0["Gradient"][0]
    assert calls1["FdF"][0] == calls0["FdF"][0]


def test_minuit_flow_from_start_values(gpu, rank1):
    from rootpwa_amd import likelihood as PL
    ds, Lo, Lg = rank1
    first = PL.pwaNloptFit(Lg, seed=5)
    again = PL.pwaFit(Lg, startValues=first.fitParameters)
    assert again.converged()
    assert again.logLikelihood() <= first.logLikelihood()
    assert again.nmbEvaluations < 200


def test_fixed_parameters_and_rank2(gpu, rank2_thresholds):
    from rootpwa_amd import likelihood as PL
    ds, Lo, Lg = rank2_thresholds
    assert Lg.nmbParsFixed() > 0
    res = PL.pwaNloptFit(Lg, seed=99, checkHessian=True)
    assert res.status > 0
    assert len(res.hessianEigenValues) == Lg.nmbPars() - Lg.nmbParsFixed()
    xs, free = check_minimum(Lo, Lg, res, tol_sigma=0.05)
    P = Lg.nmbPars()
    fixed = [i for i in range(P) if Lg.parameter(i).fixed()]
    assert np.all(res.covMatrix[fixed, :] == 0) and np.all(res.covMatrix[:, fixed] == 0)


def test_mixed_precision_fit(gpu):
    from rootpwa_amd import likelihood as PL
    ds, Lo, L64 = build_pair((2, 5), 1, seed=41, nData=6000, nMC=30000)
    _, _, L32 = build_pair((2, 5), 1, seed=41, nData=6000, nMC=30000, precision=1)
    r64 = PL.pwaFit(L64, seed=3)
    r32 = PL.pwaFit(L32, seed=3)
    assert r32.converged()
    # complex64 amplitudes move the minimum by far less than its statistical error
    cov = r64.covMatrix
    sig = np.sqrt(np.maximum(np.diag(cov), 1e-300))
    free = np.diag(cov) > 0
    assert np.max(np.abs(r32.fitParameters - r64.fitParameters)[free] / sig[free]) < 1e-3
    assert abs(r32.logLikelihood() - r64.logLikelihood()) < 1e-3


def test_attempts_loop(gpu, rank1):
    from rootpwa_amd import likelihood as PL
    ds, Lo, Lg = rank1
    results, iBest, iBestValid = PL.fitAttempts(Lg, PL.pwaNloptFit, seed=2024, attempts=3, saveSpace=True)
    assert len(results) == 3 and 0 <= iBest < 3
    assert results[iBest].logLikelihood() == min(r.logLikelihood() for r in results)
    assert iBestValid == -1 or results[iBestValid].converged()
    assert all(r.covMatrix is None for r in results)
