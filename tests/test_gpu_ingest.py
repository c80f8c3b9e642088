"""Amplitude ingest from device memory (pwaLikelihood::addAmplitudesDevice ->
rpwa_lik_load_decay_amps_device -> the on-device normalisation) against the
host path (addAmplitude, pwaLikelihood.cc:1112-1288), and the steady state of
the persistent event kernels: event lists long enough that every workgroup
processes several tiles (staged kernel) or several grid-stride iterations
(register tiles), against the single-thread oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _waves(nW):
    return [("neg%02d" % i, -1, 0.0) for i in range(nW[0])] + [("pos%02d" % i, +1, 0.0) for i in range(nW[1])]


def _pair(capi, PL, nW, rank, data, mc, mcAcc, precision, lanes):
    """(host-ingest likelihood, device-ingest likelihood) with one fixed tile."""
    waves = _waves(nW)
    names = [w[0] for w in waves]

    def integrals():
        # fresh matrices per likelihood: addAccIntegral sets the acceptance
        # matrix's event count (pwaLikelihood.cc:1100), as in the reference
        I = PL.ampIntegralMatrix(names)
        I.integrate(mc)
        Iacc = PL.ampIntegralMatrix(names)
        Iacc.integrate(mcAcc)
        return I, Iacc

    I, Iacc = integrals()
    Lh = PL.initLikelihood(waves, {names[k]: data[:, k] for k in range(len(names))}, I, accIntegral=Iacc,
                           rank=rank, precision=precision)
    d = capi.DeviceBuffer(data.nbytes)
    d.upload(np.ascontiguousarray(data, dtype=np.complex128))
    I, Iacc = integrals()
    Ld = PL.pwaLikelihood(precision)
    Ld.init(waves, rank)
    Ld.addNormIntegral(I)
    Ld.addAccIntegral(Iacc)
    Ld.addAmplitudesDevice(d.address, data.shape[0])
    Ld.finishInit()
    # one tile for both contexts, so that both reduce in the same order
    Lh.gpuContext().set_tile(lanes)
    Ld.gpuContext().set_tile(lanes)
    return Lh, Ld, d


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("lanes", [1, 16])
def test_device_ingest_bitwise_equals_host(gpu, precision, lanes):
    """The device path divides each column by sqrt(Re I_ii) with the rounding
    of the host's division: value and gradient are bitwise those of the host
    path, in FP64 and with complex64 storage; both agree with the oracle."""
    from rootpwa_amd import likelihood as PL
    nW, rank = (3, 9), 2
    ds = H.dataset(nW, rank, nData=2500, nMC=3000, seed=41)
    acc = ds["mc"][np.random.default_rng(3).random(3000) < 0.55]
    Lh, Ld, d = _pair(gpu, PL, nW, rank, ds["data"], ds["mc"], acc, precision, lanes)
    Lo = O.OracleLikelihood(nW, rank)
    Lo.add_integrals(ds["integral"], 3000, O.integral(acc), acc.shape[0])
    Lo.add_amplitudes(ds["data"])
    Lo.finish_init()
    rng = np.random.default_rng(5)
    for k in range(3):
        par = H.random_par(rng, Lo, 0.8)
        fh, gh = Lh.FdF(par)
        fd, gd = Ld.FdF(par)
        print("ingest precision", precision, "lanes", lanes, "point", k, fh, fd, flush=True)
        assert fh == fd and np.array_equal(gh, gd)
        fo, go = Lo.fdf(par)
        tol_f, tol_g = (1e-12, 1e-11) if precision == 0 else (1e-6, 1e-5)
        assert abs(fd - fo) <= tol_f * max(1.0, abs(fo))
        assert np.max(np.abs(gd - go)) <= tol_g * np.max(np.abs(go))
    d.free()


def test_device_ingest_zero_wave_rules(gpu):
    """A wave whose amplitudes and acceptance integral are all zero gets its
    production amplitudes fixed on both paths (pwaLikelihood.cc:1247-1281); a
    non-zero amplitude of a wave with a zero normalisation integral is refused
    on both paths."""
    from rootpwa_amd import likelihood as PL
    nW, rank = (2, 5), 1
    ds = H.dataset(nW, rank, nData=700, nMC=1500, seed=43)
    zero = 1 + 3  # refl +1 wave 3
    data = ds["data"].copy()
    mc = ds["mc"].copy()
    data[:, zero] = 0
    mc[:, zero] = 0
    Lh, Ld, d = _pair(gpu, PL, nW, rank, data, mc, mc, 0, 16)
    fixed_h = [p.fixed() for p in Lh.parameters()]
    fixed_d = [p.fixed() for p in Ld.parameters()]
    assert fixed_h == fixed_d and sum(fixed_d) == 2
    par = H.random_par(np.random.default_rng(9), O.OracleLikelihood(nW, rank), 0.8)
    fh, gh = Lh.FdF(par)
    fd, gd = Ld.FdF(par)
    assert fh == fd and np.array_equal(gh, gd)
    d.free()
    # non-zero amplitudes of a wave whose normalisation integral is zero
    data[5, zero] = 0.5 + 0.25j
    waves = _waves(nW)
    names = [w[0] for w in waves]
    I = PL.ampIntegralMatrix(names)
    I.integrate(mc)
    with pytest.raises(Exception):
        PL.initLikelihood(waves, {names[k]: data[:, k] for k in range(len(names))}, I, rank=rank)
    d = gpu.DeviceBuffer(data.nbytes)
    d.upload(data)
    Ld = PL.pwaLikelihood(0)
    Ld.init(waves, rank)
    Ld.addNormIntegral(I)
    Ld.addAccIntegral(I)
    with pytest.raises(Exception):
        Ld.addAmplitudesDevice(d.address, data.shape[0])
    assert gpu.wave_nonzero_device(d.address, data.shape[0], data.shape[1], zero)
    assert not gpu.wave_nonzero_device(d.address, 5, data.shape[1], zero)
    assert gpu.wave_nonzero_device(d.address, data.shape[0], data.shape[1], 0)
    d.free()


@pytest.fixture(scope="module")
def long_list():
    """200017 events x (2+5) waves: at least two tiles or grid-stride steps per
    workgroup for every variant (checked in the test)."""
    nW, rank, N = (2, 5), 2, 200017
    W = sum(nW)
    scales = O.gen_wave_scales(61, W)
    mc = O.gen_amps(61, 3, 0, 20000, scales)
    data = O.gen_amps(61, 4, 0, N, scales)
    Lo = O.OracleLikelihood(nW, rank)
    Lo.add_integrals(O.integral(mc), 20000)
    Lo.add_amplitudes(data)
    Lo.finish_init()
    return nW, rank, Lo, H.normalised_amps(Lo)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_steady_state_every_variant(gpu, long_list, precision):
    """Every workgroup of the persistent grid loops: the staged kernel stages
    the prefetched tile on its second and later iterations, sums G and the
    running product of sigma across tiles; the register tiles run several
    grid-stride iterations.  Against the single-thread oracle."""
    nW, rank, Lo, amps = long_list
    N = amps.shape[0]
    prod = H.random_prod(np.random.default_rng(17), rank, nW, 0.8)
    f_ref = Lo.event_loglik(prod, 0.3)
    d_ref, df_ref = Lo.event_deriv(prod, 0.3)
    tol_f, tol_d = (1e-12, 1e-10) if precision == 0 else (1e-6, 1e-5)
    ctx = gpu.LikelihoodContext(0, precision)
    ctx.load(amps, nW)
    variants = [("stage", e) for e in (64, 32, 16)] + [("lanes", l) for l in (8, 16, 32, 64)]
    variants += [("stage_mfma", e) for e in (64, 32)] + [("stage_wave", 64)]
    for kind, v in variants:
        try:
            if kind == "lanes":
                ctx.set_tile(v)
            else:
                ctx.set_stage_variant(v, {"stage": 0, "stage_mfma": 1, "stage_wave": 2}[kind])
        except gpu.RpwaError as e:
            assert e.code == gpu.RPWA_ERR_UNSUPPORTED
            assert kind not in ("stage_mfma", "stage_wave") or precision == 0
            continue
        f, d, df = ctx.fdf(prod, 0.3)
        info = ctx.info()
        nWg = int(info["workgroups"])
        perSweep = nWg * (v if kind != "lanes" else 256 // v)
        print("steady state precision", precision, kind, v, "workgroups", nWg, "sweeps", N / perSweep, flush=True)
        assert N >= 2 * perSweep
        assert abs(f - f_ref) <= tol_f * max(1.0, abs(f_ref))
        assert np.max(np.abs(d - d_ref)) <= tol_d * np.max(np.abs(d_ref))
        assert abs(df - df_ref) <= tol_d * max(1.0, abs(df_ref))
    ctx.close()
