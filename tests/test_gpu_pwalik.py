"""GPU parity of the drop-in likelihood class (C++ pwaLikelihood behind the
C ABI) against the oracle: DoEval, Gradient, FdF, DoDerivative,
CorrectParamSigns, priors, and the bitwise contracts of the fit drivers."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def wave_list(nW):
    waves = [("neg%02d" % i, -1, 0.0) for i in range(nW[0])] + [("pos%02d" % i, +1, 0.0) for i in range(nW[1])]
    return waves


def build_pair(nW, rank, seed, prior=None, precision=0, nData=900, nMC=4000):
    from rootpwa_amd import likelihood as PL
    ds = H.dataset(nW, rank, nData, nMC, seed)
    Lo = H.oracle_likelihood(ds, prior=prior)
    waves = wave_list(nW)
    names = [w[0] for w in waves]
    I = PL.ampIntegralMatrix(names)
    I.setSums(ds["integral"], ds["nMC"])
    amps = {names[k]: ds["data"][:, k] for k in range(len(names))}
    Lg = PL.initLikelihood(waves, amps, I, rank=rank, precision=precision, cauchyWidth=prior)
    return ds, Lo, Lg


# ranks above 4 run the staged kernel in blocks of 4 ranks (5 and 6: two blocks, 9: three)
CASES = [((0, 16), 1, None), ((8, 80), 2, None), ((3, 5), 2, 0.5), ((7, 34), 2, None), ((4, 6), 3, 0.7),
         ((3, 7), 5, None), ((6, 9), 6, 0.6), ((2, 12), 9, None)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: "W%d+%d_r%d_%s" % (c[0][0], c[0][1], c[1],
                                                                              "cauchy" if c[2] else "flat"))
def pair(request):
    nW, rank, prior = request.param
    return build_pair(nW, rank, seed=100 + sum(nW) + rank, prior=prior)


def test_parameter_layout(gpu, pair):
    ds, Lo, Lg = pair
    assert Lg.nmbPars() == Lo.nPars
    assert Lg.nmbParsFixed() == Lo.nmb_pars_fixed()
    assert Lg.nmbEvents() == Lo.nmb_events()
    rng = np.random.default_rng(1)
    par = rng.normal(size=Lo.nPars)
    pg, fg = Lg.copyFromParArray(par)
    po, fo = Lo.copy_from_par(par)
    assert np.array_equal(pg, po) and fg == fo
    assert np.array_equal(Lg.copyToParArray(pg, fg), Lo.copy_to_par(po, fo))


def test_values_and_gradients(gpu, pair):
    ds, Lo, Lg = pair
    rng = np.random.default_rng(2)
    for scale in (0.3, 1.0, 3.0):
        par = rng.normal(size=Lo.nPars) * scale
        f_ref, g_ref = Lo.fdf(par)
        f, g = Lg.FdF(par)
        print("f", f, f_ref, "rel", abs(f - f_ref) / abs(f_ref))
        assert abs(f - f_ref) <= 1e-10 * max(abs(f_ref), 1.0)
        assert H.rel_err(g, g_ref) <= 1e-10
        assert abs(Lg.DoEval(par) - Lo.doeval(par)) <= 1e-10 * max(abs(f_ref), 1.0)
        assert H.rel_err(Lg.Gradient(par), Lo.gradient(par)) <= 1e-10
        for k in (0, Lo.nPars // 2, Lo.nPars - 1):
            assert abs(Lg.DoDerivative(par, k) - g_ref[k]) <= 1e-10 * np.max(np.abs(g_ref))


def test_bitwise_contracts(gpu, pair):
    """pwaNloptFit.cc:238-244 and pwaFit.cc:295-300 compare DoEval bitwise."""
    ds, Lo, Lg = pair
    rng = np.random.default_rng(3)
    par = rng.normal(size=Lo.nPars)
    f, g = Lg.FdF(par)
    assert Lg.DoEval(par) == f
    assert np.array_equal(Lg.Gradient(par), g)
    fixed = Lg.CorrectParamSigns(par)
    assert np.array_equal(fixed, Lo.correct_param_signs(par))
    assert Lg.DoEval(fixed) == f
    # negative anchors flip whole blocks, bitwise invariant
    par2 = par.copy()
    pg, fl = Lg.copyFromParArray(par2)
    pg[:, :, :] *= -1
    par2 = Lg.copyToParArray(pg, -fl)
    assert Lg.DoEval(par2) == f


def test_clone_shares_context(gpu, pair):
    ds, Lo, Lg = pair
    c = Lg.Clone()
    par = np.random.default_rng(4).normal(size=Lo.nPars)
    assert c.DoEval(par) == Lg.DoEval(par)
    del c
    assert np.isfinite(Lg.DoEval(par))


def test_mixed_precision_error(gpu):
    ds, Lo, Lg = build_pair((8, 80), 2, seed=7, precision=1, nData=3000)
    par = np.random.default_rng(5).normal(size=Lo.nPars)
    f_ref, g_ref = Lo.fdf(par)
    f, g = Lg.FdF(par)
    print("mixed |dlnL|", abs(f - f_ref), "grad rel", H.rel_err(g, g_ref))
    assert abs(f - f_ref) < 1e-3
    assert H.rel_err(g, g_ref) < 1e-5


@pytest.mark.parametrize("rank", [5, 6])
def test_high_rank_mixed_precision(gpu, rank):
    """Ranks above 4 with complex64 storage: value and gradient against the
    oracle, DoEval bitwise FdF, and bitwise invariance under the sign
    convention."""
    ds, Lo, Lg = build_pair((4, 9), rank, seed=60 + rank, precision=1, nData=2500)
    rng = np.random.default_rng(rank)
    par = rng.normal(size=Lo.nPars)
    f_ref, g_ref = Lo.fdf(par)
    f, g = Lg.FdF(par)
    print("rank", rank, "mixed |dlnL|", abs(f - f_ref), "grad rel", H.rel_err(g, g_ref))
    assert abs(f - f_ref) < 1e-6 * max(1.0, abs(f_ref))
    assert H.rel_err(g, g_ref) < 1e-5
    assert Lg.DoEval(par) == f and np.array_equal(Lg.Gradient(par), g)
    assert Lg.DoEval(Lg.CorrectParamSigns(par)) == f


def test_integral_matrix_class(gpu, tmp_path):
    from rootpwa_amd import likelihood as PL
    rng = np.random.default_rng(6)
    a = rng.normal(size=(2000, 5)) + 1j * rng.normal(size=(2000, 5))
    m = PL.ampIntegralMatrix(["a", "b", "c", "d", "e"])
    m.integrate(a)
    sums, n = m.sums()
    assert n == 2000
    assert H.rel_err(sums, O.integral(a)) <= 1e-12
    assert abs(m.element(1, 2) - O.integral(a)[1, 2] / 2000) <= 1e-12 * abs(O.integral(a)[1, 2] / 2000)
    m2 = PL.ampIntegralMatrix(["a", "b", "c", "d", "e"])
    m2.integrate(a[:700])
    m3 = PL.ampIntegralMatrix(["a", "b", "c", "d", "e"])
    m3.integrate(a[700:])
    m2 += m3
    s2, n2 = m2.sums()
    assert n2 == 2000 and H.rel_err(s2, sums) <= 1e-12
    f = str(tmp_path / "int.txt")
    m.writeAscii(f)
    m4 = PL.ampIntegralMatrix()
    m4.readAscii(f)
    s4, n4 = m4.sums()
    assert n4 == 2000 and np.array_equal(s4, sums) and m4.waveNames() == ["a", "b", "c", "d", "e"]


def test_hessian_matches_oracle(gpu, pair):
    """pwaLikelihood::Hessian (pwaLikelihood.cc:642-859): GPU event sums + host terms."""
    ds, Lo, Lg = pair
    par = np.random.default_rng(11).normal(size=Lo.nPars)
    Hg = Lg.Hessian(par)
    Ho = Lo.hessian(par)
    scale = np.max(np.abs(Ho))
    print("hessian max abs diff / max", np.max(np.abs(Hg - Ho)) / scale)
    assert np.max(np.abs(Hg - Ho)) <= 1e-10 * scale
    # covariance = inverse of the Hessian on the free parameters
    C = Lg.CovarianceMatrix(par)
    free = [i for i in range(Lo.nPars) if not Lg.parameter(i).fixed()]
    Hf = Ho[np.ix_(free, free)]
    assert np.allclose(C[np.ix_(free, free)] @ Hf, np.eye(len(free)), atol=1e-6)


def test_hessian_mixed_precision(gpu):
    ds, Lo, Lg = build_pair((3, 9), 2, seed=13, precision=1, nData=1500)
    par = np.random.default_rng(12).normal(size=Lo.nPars)
    Hg = Lg.Hessian(par)
    Ho = Lo.hessian(par)
    assert np.max(np.abs(Hg - Ho)) <= 1e-5 * np.max(np.abs(Ho))


def test_on_the_fly_binning_matches_prefiltered(gpu):
    """A likelihood that selects its events with setOnTheFlyBinning
    (pwaLikelihood.cc:1143-1215,1391-1436) is the likelihood of the selected
    events: identical values and gradients, and the oracle's on them."""
    from rootpwa_amd import likelihood as PL
    nW, rank = (3, 7), 2
    ds = H.dataset(nW, rank, 1200, 3000, 77)
    rng = np.random.default_rng(78)
    mass = rng.uniform(0.5, 2.5, size=ds["data"].shape[0])
    tPrime = rng.uniform(0.0, 1.0, size=mass.size)
    bounds = {"mass": (1.0, 1.8), "tPrime": (0.1, 0.6)}
    keep = (mass >= 1.0) & (mass < 1.8) & (tPrime >= 0.1) & (tPrime < 0.6)
    waves = wave_list(nW)
    names = [w[0] for w in waves]
    I = PL.ampIntegralMatrix(names)
    I.setSums(ds["integral"], ds["nMC"])
    amps = {names[k]: ds["data"][:, k] for k in range(len(names))}
    half = mass.size // 2
    files = [{"mass": mass[:half], "tPrime": tPrime[:half]}, {"mass": mass[half:], "tPrime": tPrime[half:]}]
    Lb = PL.initLikelihood(waves, amps, I, rank=rank, multibinBoundaries=bounds, eventVariables=files)
    Lf = PL.initLikelihood(waves, {k: v[keep] for k, v in amps.items()}, I, rank=rank)
    assert Lb.nmbEvents() == Lf.nmbEvents() == int(keep.sum()) > 0
    sub = dict(ds, data=ds["data"][keep])
    Lo = H.oracle_likelihood(sub)
    par = rng.normal(size=Lb.nmbPars())
    # the same kernel variant in both contexts (the autotuner may pick different ones,
    # whose sums differ in the last bits)
    Lb.gpuContext().set_tile(16)
    Lf.gpuContext().set_tile(16)
    fb, gb = Lb.FdF(par)
    ff, gf = Lf.FdF(par)
    assert fb == ff and np.array_equal(gb, gf)
    fo, go = Lo.fdf(par)
    assert abs(fb - fo) <= 1e-10 * abs(fo) and H.rel_err(gb, go) <= 1e-10


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_cpp_consumer_matches_oracle(gpu, tmp_path, precision):
    """The C++ consumer (tests/cpp/pwalik_dropin.cc: rpwa::pwaLikelihood through
    IGradientFunctionMultiDim and Clone, as pwaFit / pwaNloptFit call it) against
    the oracle, with its bitwise DoEval / Gradient / DoDerivative contracts."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "pwalik_dropin")
    assert os.path.exists(exe), "build() links tests/cpp/pwalik_dropin"
    nW, rank = (3, 7), 2
    ds = H.dataset(nW, rank, 600, 2000, 91)
    Lo = H.oracle_likelihood(ds)
    rng = np.random.default_rng(92)
    pts = [rng.normal(size=Lo.nPars) * s for s in (0.5, 1.0, 2.0)]
    W = sum(nW)
    with open(tmp_path / "in.txt", "w") as fh:
        fh.write("%d %d %d %d %d %d\n" % (nW[0], nW[1], rank, ds["data"].shape[0], ds["nMC"], precision))
        fh.write(" ".join("%.17g %.17g" % (z.real, z.imag) for z in ds["integral"].ravel()) + "\n")
        fh.write(" ".join("%.17g %.17g" % (z.real, z.imag) for z in ds["data"].ravel()) + "\n")
        fh.write("%d\n" % len(pts))
        for p in pts:
            fh.write(" ".join("%.17g" % x for x in p) + "\n")
    r = subprocess.run([exe, str(tmp_path / "in.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == 2 * len(pts)
    tol_f, tol_g = (1e-12, 1e-10) if precision == 0 else (1e-6, 1e-5)
    for k, p in enumerate(pts):
        vals = np.array([float(x) for x in lines[2 * k].split()])
        assert lines[2 * k + 1] == "bitwise 1"
        f_ref, g_ref = Lo.fdf(p)
        assert abs(vals[0] - f_ref) <= tol_f * abs(f_ref)
        assert H.rel_err(vals[1:], g_ref) <= tol_g
