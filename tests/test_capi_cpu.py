"""The C ABI without a GPU: the library loads, exports every declared entry
point, and its host-side logic (parameter mapping, sign convention, integral
matrix bookkeeping, error paths) matches the oracle.  No kernel runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in ("rootpwa_amd.h", "rootpwa_amd_pwalik.h", "rootpwa_amd_batch.h",
                                                           "rootpwa_amd_mlp.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(rpwa_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from rootpwa_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = [n for h in HEADERS for n in declared_functions(h)]
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_python_bindings_cover_the_headers():
    from rootpwa_amd import batch, capi, likelihood, mlp
    names = [n for h in HEADERS for n in declared_functions(h)]
    bound = set(capi.PROTOTYPES) | set(likelihood.PROTOTYPES) | set(batch.PROTOTYPES) | set(mlp.PROTOTYPES)
    assert [n for n in names if n not in bound] == []


def test_errors_without_gpu():
    from rootpwa_amd import capi
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.RpwaError) as e:
        capi.LikelihoodContext()
    assert e.value.code == capi.RPWA_ERR_INVALID and "no GPU" in str(e.value)
    # null context is an error, never a crash
    assert capi.lib().rpwa_lik_release(None) == capi.RPWA_ERR_INVALID
    assert b"null context" in capi.lib().rpwa_last_error()


def wave_list(nW):
    return [("neg%02d" % i, -1, 0.0) for i in range(nW[0])] + [("pos%02d" % i, +1, 0.0) for i in range(nW[1])]


@pytest.mark.parametrize("nW,rank", [((0, 16), 1), ((8, 80), 2), ((3, 5), 3), ((4, 0), 1)])
def test_parameter_mapping_matches_oracle(nW, rank):
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    L.init(wave_list(nW), rank)
    Lo = O.OracleLikelihood(nW, rank)
    assert L.nmbPars() == Lo.nPars
    assert L.nmbWaves() == sum(nW) and L.nmbWaves(-1) == nW[0] and L.nmbWaves(+1) == nW[1]
    rng = np.random.default_rng(1)
    par = rng.normal(size=Lo.nPars)
    pg, fg = L.copyFromParArray(par)
    po, fo = Lo.copy_from_par(par)
    assert np.array_equal(pg, po) and fg == fo
    assert np.array_equal(L.copyToParArray(pg, fg), par)
    # CorrectParamSigns on parameters with negative anchors
    par[::3] *= -1
    assert np.array_equal(L.CorrectParamSigns(par), Lo.correct_param_signs(par))


def test_parameter_names():
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    L.init([("1-(1++)0+rho770_01_pi-", +1, 0.0), ("1-(0-+)0+f0980_00_pi-", +1, 0.0),
            ("1-(1++)1-rho770_01_pi-", -1, 0.0)], 2)
    names = [L.parameter(i).parName() for i in range(L.nmbPars())]
    # pwaLikelihood.cc:107-122: V<rank>_<wave>_RE/_IM, V_flat_RE for the flat wave
    assert names == ["V0_1-(1++)1-rho770_01_pi-_RE",
                     "V0_1-(1++)0+rho770_01_pi-_RE", "V0_1-(0-+)0+f0980_00_pi-_RE", "V0_1-(0-+)0+f0980_00_pi-_IM",
                     "V1_1-(0-+)0+f0980_00_pi-_RE",
                     "V_flat_RE"]


def test_thresholds_fix_parameters():
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    L.init([("a", -1, 0.0), ("b", -1, 1.0), ("c", +1, 0.0), ("d", +1, 0.0), ("e", +1, 2.0)], 1, massBinCenter=1.5)
    fixed = [L.parameter(i).fixed() for i in range(L.nmbPars())]
    assert sum(fixed) == 2 and L.nmbParsFixed() == 2


def test_call_order_errors():
    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    L.init(wave_list((1, 2)), 1)
    m = PL.ampIntegralMatrix(["neg00", "pos00", "pos01"])
    with pytest.raises(capi.RpwaError):
        L.addAccIntegral(m)            # before addNormIntegral
    with pytest.raises(capi.RpwaError):
        L.addAmplitude("pos00", np.ones(3))  # before the integrals
    with pytest.raises(capi.RpwaError) as e:
        L.DoEval(np.zeros(L.nmbPars()))  # before finishInit
    assert e.value.code == capi.RPWA_ERR_STATE
    bad = PL.ampIntegralMatrix(["neg00", "pos00"])
    with pytest.raises(capi.RpwaError):
        L.addNormIntegral(bad)         # a wave of the list is missing


def test_integral_matrix_bookkeeping(tmp_path):
    from rootpwa_amd import likelihood as PL
    rng = np.random.default_rng(2)
    a = rng.normal(size=(300, 4)) + 1j * rng.normal(size=(300, 4))
    sums = O.integral(a)
    m = PL.ampIntegralMatrix(["w0", "w1", "w2", "w3"])
    m.setSums(sums, 300)
    # element() divides component-wise (ampIntegralMatrix.cc:251, complex / double)
    cdiv = lambda z, d: complex(z.real / d, z.imag / d)
    assert m.element(1, 3) == cdiv(sums[1, 3], 300)
    assert m.element("w2", "w0") == cdiv(sums[2, 0], 300)
    f = str(tmp_path / "integral.txt")
    m.writeAscii(f)
    m2 = PL.ampIntegralMatrix()
    m2.readAscii(f)
    s2, n2 = m2.sums()
    assert n2 == 300 and m2.waveNames() == ["w0", "w1", "w2", "w3"] and np.array_equal(s2, sums)
    m2 += m
    s3, n3 = m2.sums()
    assert n3 == 600 and np.allclose(s3, 2 * sums, rtol=1e-15)
    m2.renormalize(300)
    s4, n4 = m2.sums()
    assert n4 == 300 and np.allclose(s4, sums, rtol=1e-15)


def test_norm_integral_reordering_against_oracle():
    """addNormIntegral/addAccIntegral take elements by wave name: a permuted
    integral matrix gives the same likelihood set-up."""
    from rootpwa_amd import likelihood as PL
    ds = H.dataset((2, 3), 1, 50, 500, seed=3)
    names = [w[0] for w in wave_list((2, 3))]
    perm = [4, 2, 0, 3, 1]
    m = PL.ampIntegralMatrix([names[p] for p in perm])
    m.setSums(ds["integral"][np.ix_(perm, perm)], ds["nMC"])
    for i, a in enumerate(perm):
        for j, b in enumerate(perm):
            z = ds["integral"][a, b]
            assert m.element(names[a], names[b]) == complex(z.real / ds["nMC"], z.imag / ds["nMC"])


def test_start_values_follow_mt19937():
    """The drivers' start values (pwaNloptFit.cc:192-203): Uniform(0.01, sqrt(N))
    and a sign draw per parameter from MT19937 seeded by init_genrand, pinned
    against numpy's independent MT19937 (RandomState with an integer seed)."""
    from rootpwa_amd import likelihood as PL
    nW = (2, 3)
    waves = [("a", -1, 0.0), ("b", -1, 9.0), ("c", +1, 0.0), ("d", +1, 0.0), ("e", +1, 0.0)]
    L = PL.pwaLikelihood()
    L.init(waves, 1, massBinCenter=1.0)
    m = PL.ampIntegralMatrix([w[0] for w in waves])
    m.setSums(np.eye(5, dtype=np.complex128) * 10, 10)
    L.addNormIntegral(m)
    L.addAccIntegral(m)
    nEv = 1234
    for w in waves:
        L.addAmplitude(w[0], np.ones(nEv, dtype=np.complex128))
    for seed in (0, 1, 123456, 2 ** 32 - 1):
        raw = np.random.RandomState(seed).randint(0, 2 ** 32, size=4 * L.nmbPars() + 8, dtype=np.uint64)
        raw = raw[raw != 0]  # TRandom3::Rndm skips zero draws
        u = raw.astype(np.float64) * 2.0 ** -32
        expect = np.zeros(L.nmbPars())
        for i in range(L.nmbPars()):
            v = 0.01 + (np.sqrt(float(nEv)) - 0.01) * u[2 * i]
            if u[2 * i + 1] > 0.5:
                v = -v
            expect[i] = 0.0 if L.parameter(i).fixed() else v
        got = PL.startValues(L, seed)
        assert np.array_equal(got, expect), seed
    assert L.nmbParsFixed() == 2 and PL.startValues(L, 5)[0] != 0.0


def test_hessian_eigenvectors_against_lapack():
    """HessianEigenVectors (pwaLikelihood.cc:864-911): eigen-decomposition of the
    Hessian reduced to the free parameters, eigenvectors zero at fixed ones,
    eigenvalues descending; checked against numpy.linalg.eigh."""
    from rootpwa_amd import likelihood as PL
    waves = [("a", -1, 0.0), ("b", -1, 5.0), ("c", +1, 0.0), ("d", +1, 0.0), ("e", +1, 3.0), ("f", +1, 0.0)]
    L = PL.pwaLikelihood()
    L.init(waves, 2, massBinCenter=2.0)
    P = L.nmbPars()
    fixed = np.array([L.parameter(i).fixed() for i in range(P)], dtype=bool)
    assert fixed.sum() == L.nmbParsFixed() > 0
    rng = np.random.default_rng(5)
    B = rng.normal(size=(P, P))
    Hm = B @ B.T - 3.0 * np.eye(P)  # symmetric, some negative eigenvalues
    Hm[2, 5] += 1e-9               # asymmetry is symmetrised away
    eig = L.HessianEigenVectors(Hm)
    free = np.flatnonzero(~fixed)
    Hr = 0.5 * (Hm + Hm.T)[np.ix_(free, free)]
    ref_vals, ref_vecs = np.linalg.eigh(Hr)
    ref_vals = ref_vals[::-1]
    ref_vecs = ref_vecs[:, ::-1]
    assert len(eig) == free.size
    vals = np.array([e[1] for e in eig])
    assert np.all(np.diff(vals) <= 0)
    assert np.max(np.abs(vals - ref_vals)) <= 1e-11 * np.max(np.abs(ref_vals))
    for k, (v, lam) in enumerate(eig):
        assert np.all(v[fixed] == 0.0)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-12
        assert np.linalg.norm(Hr @ v[free] - lam * v[free]) <= 1e-10 * np.max(np.abs(ref_vals))
        assert abs(abs(v[free] @ ref_vecs[:, k]) - 1.0) < 1e-8


@pytest.mark.parametrize("kind", ["degenerate", "indefinite", "large"])
def test_hessian_eigenvectors_special_spectra(kind):
    """Repeated eigenvalues (any orthonormal basis of the eigenspace is right),
    negative eigenvalues (a Hessian away from the minimum), and the size of the
    2M x 88 rank-2 fit (345 parameters): A v = lambda v, V orthonormal,
    eigenvalues equal to LAPACK's."""
    from rootpwa_amd import likelihood as PL
    nW = (8, 80) if kind == "large" else (2, 5)
    L = PL.pwaLikelihood()
    L.init(wave_list(nW), 2)
    P = L.nmbPars()
    rng = np.random.default_rng(8)
    Q, _ = np.linalg.qr(rng.normal(size=(P, P)))
    if kind == "degenerate":
        lam = np.repeat([5.0, 2.0, 2.0, -1.0], -(-P // 4))[:P]
    elif kind == "indefinite":
        lam = rng.normal(size=P) * 10
    else:
        lam = np.exp(rng.normal(size=P) * 3)
    Hm = (Q * lam) @ Q.T
    eig = L.HessianEigenVectors(Hm)
    vals = np.array([e[1] for e in eig])
    V = np.array([e[0] for e in eig])
    scale = np.max(np.abs(lam))
    assert np.all(np.diff(vals) <= 0)
    assert np.max(np.abs(vals - np.sort(lam)[::-1])) <= 1e-12 * scale
    assert np.max(np.abs(V @ V.T - np.eye(P))) <= 1e-12
    assert max(np.linalg.norm(Hm @ V[k] - vals[k] * V[k]) for k in range(P)) <= 1e-12 * scale


def test_fit_drivers_refuse_unfinished_likelihood():
    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    L.init(wave_list((1, 2)), 1)
    with pytest.raises(capi.RpwaError) as e:
        PL.pwaNloptFit(L, seed=1)
    assert e.value.code == capi.RPWA_ERR_STATE
    with pytest.raises(capi.RpwaError):
        PL.pwaFit(L, seed=1, memory=0)


def test_batch_errors_without_gpu():
    from rootpwa_amd import batch as B
    from rootpwa_amd import capi
    with pytest.raises(capi.RpwaError) as e:
        B.BatchFit(97, 4)  # more waves than the kernel holds
    assert e.value.code == capi.RPWA_ERR_INVALID
    with pytest.raises(capi.RpwaError):
        B.BatchFit(8, 65)  # more starts than one workgroup's columns
    if capi.device_count() == 0:
        with pytest.raises(capi.RpwaError) as e:
            B.BatchFit(8, 4)
        assert "no GPU" in str(e.value)


def test_settings_and_accessors():
    """Modifiers and accessors of pwaLikelihood.h:154,195-196,208-215 with the
    reference's defaults (pwaLikelihood.cc:145-149)."""
    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    assert L.normalizedAmpsUsed() and L.priorType() == PL.pwaLikelihood.FLAT and L.cauchyWidth() == 0.5
    L.useNormalizedAmps(False)
    L.setPriorType(PL.pwaLikelihood.HALF_CAUCHY)
    assert L.cauchyWidth() == 0.5          # setPriorType keeps the width
    L.setCauchyWidth(0.25)
    assert (L.normalizedAmpsUsed(), L.priorType(), L.cauchyWidth()) == (False, PL.pwaLikelihood.HALF_CAUCHY, 0.25)
    L.setPriorType(PL.pwaLikelihood.FLAT)
    assert L.priorType() == PL.pwaLikelihood.FLAT and L.cauchyWidth() == 0.25
    assert L.cudaEnabled()
    L.enableCuda(False)
    assert not L.cudaEnabled()
    L.enableCuda()
    PL.pwaLikelihood.setQuiet(True)
    PL.pwaLikelihood.setQuiet(False)
    L.init([("a", -1, 0.0), ("b", -1, 1.0), ("c", +1, 0.0), ("d", +1, 2.0)], 2, massBinCenter=1.5)
    pars = L.parameters()
    assert len(pars) == L.nmbPars() and pars[0] == L.parameter(0)
    assert [(p.waveName(), p.rank(), p.threshold(), p.fixed(), p.realPart()) for p in pars] == [
        ("a", 0, 0.0, False, True), ("b", 0, 1.0, False, True), ("b", 0, 1.0, False, False),
        ("c", 0, 0.0, False, True), ("d", 0, 2.0, True, True), ("d", 0, 2.0, True, False),
        ("b", 1, 1.0, False, True), ("d", 1, 2.0, True, True),
        ("flat", 0, 0.0, False, True)]
    text = str(L)
    assert "number of function parameters ........... %d" % L.nmbPars() in text and "V1_d_RE" in text
    assert L.printFuncInfo() == ""          # no calls yet
    with pytest.raises(capi.RpwaError):
        L.CovarianceMatrix(np.zeros(L.nmbPars()))   # evaluation before finishInit


def test_build_prod_amp_arrays():
    """buildProdAmpArrays (pwaLikelihood.cc:1624-1661) against copyFromParArray,
    whose map the oracle pins, and against the parameter names."""
    from rootpwa_amd import likelihood as PL
    nW, rank = (3, 5), 3
    L = PL.pwaLikelihood()
    L.init(wave_list(nW), rank)
    par = np.random.default_rng(4).normal(size=L.nmbPars())
    prod, idx, names = L.buildProdAmpArrays(par)
    po, flat = O.OracleLikelihood(nW, rank).copy_from_par(par)
    expect = [(r, refl, w) for r in range(rank) for refl in (0, 1) for w in range(nW[refl]) if w >= r]
    assert len(prod) == len(expect) == len(idx) == len(names)
    waves = wave_list(nW)
    pnames = [p.parName() for p in L.parameters()]
    for k, (r, refl, w) in enumerate(expect):
        assert prod[k] == po[r, refl, w]
        wave = waves[w if refl == 0 else nW[0] + w][0]
        assert names[k] == "V%d_%s" % (r, wave)
        assert pnames[idx[k][0]] == names[k] + "_RE"
        assert (idx[k][1] == -1) == (w == r)
        if idx[k][1] != -1:
            assert pnames[idx[k][1]] == names[k] + "_IM"
    prodF, idxF, namesF = L.buildProdAmpArrays(par, withFlat=True)
    assert namesF[-1] == "V_flat" and idxF[-1] == (L.nmbPars() - 1, -1) and prodF[-1] == flat
    assert np.array_equal(prodF[:-1], prod)


def test_covariance_from_hessian():
    """CovarianceMatrix(hessian) (pwaLikelihood.cc:927-979): the inverse on the
    free parameters, zero at the fixed ones."""
    from rootpwa_amd import likelihood as PL
    L = PL.pwaLikelihood()
    L.init([("a", -1, 0.0), ("b", -1, 3.0), ("c", +1, 0.0), ("d", +1, 0.0)], 1, massBinCenter=1.0)
    P = L.nmbPars()
    fixed = np.array([p.fixed() for p in L.parameters()])
    assert fixed.sum() == 2
    rng = np.random.default_rng(5)
    M = rng.normal(size=(P, P))
    Hm = M @ M.T + P * np.eye(P)
    C = L.CovarianceMatrix(Hm)
    free = np.flatnonzero(~fixed)
    assert np.allclose(C[np.ix_(free, free)] @ Hm[np.ix_(free, free)], np.eye(free.size), atol=1e-12)
    assert np.all(C[fixed] == 0) and np.all(C[:, fixed] == 0)
    with pytest.raises(ValueError):
        L.CovarianceMatrix(Hm[:-1, :-1])


def test_on_the_fly_binning_selection():
    """setOnTheFlyBinning (pwaLikelihood.cc:1391-1436): an entry is in the bin
    when lower <= v < upper for every variable; entries of several event files
    concatenate in file order."""
    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL
    m = np.array([0.9, 1.0, 1.05, 1.1, 1.2, np.nan, 1.0])
    t = np.array([0.2, 0.1, 0.5, 0.1, 0.1, 0.1, 0.3])
    bounds = {"mass": (1.0, 1.1), "tPrime": (0.1, 0.4)}
    L = PL.pwaLikelihood()
    L.init(wave_list((1, 2)), 1)
    L.setOnTheFlyBinning(bounds, [{"mass": m[:3], "tPrime": t[:3]}, {"mass": m[3:], "tPrime": t[3:]}])
    # 1.0 (lower edge) kept, 1.1 (upper edge) dropped, t 0.5 out, NaN compares false both ways: kept
    assert L.nmbEntriesInBin() == 3
    L2 = PL.pwaLikelihood()
    L2.init(wave_list((1, 2)), 1)
    L2.setOnTheFlyBinning({}, {"mass": m})
    assert L2.nmbEntriesInBin() == m.size
    with pytest.raises(KeyError):
        L2.setOnTheFlyBinning({"x": (0, 1)}, {"mass": m})
    # amplitudes must cover all entries; the amplitudes come after the integrals
    I = PL.ampIntegralMatrix(["neg00", "pos00", "pos01"])
    I.setSums(np.eye(3, dtype=np.complex128) * 7, 7)
    L.addNormIntegral(I)
    L.addAccIntegral(I)
    with pytest.raises(capi.RpwaError):
        L.addAmplitude("pos00", np.ones(3))
    L.addAmplitude("pos00", np.ones(m.size))
    assert L.nmbEvents() == 3
    with pytest.raises(capi.RpwaError):       # binning after the amplitudes
        L.setOnTheFlyBinning(bounds, {"mass": m, "tPrime": t})


def test_cpp_consumer_links_and_sets_up():
    """tests/cpp/pwalik_dropin.cc, a C++ program using rpwa::pwaLikelihood as
    ROOTPWA's drivers do, links against the library (built by
    __graft_entry__.build()); without a GPU it sets up and names the
    parameters, and an evaluation before finishInit throws."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "pwalik_dropin")
    assert os.path.exists(exe), "build() links tests/cpp/pwalik_dropin"
    r = subprocess.run([exe, "--host-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].split() == ["V0_neg00_RE", "V0_pos00_RE", "V0_pos01_RE", "V0_pos01_IM", "V1_pos01_RE",
                                "V_flat_RE"]
    assert lines[1] == "evaluation before finishInit throws"
