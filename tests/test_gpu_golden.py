"""The GPU path against the golden vectors (60-digit decimal arithmetic,
tests/golden/make_golden.py) and on ragged event counts (tile and grid
boundaries of every kernel variant) against the oracle."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "likelihood_golden.json")
with open(GOLDEN) as _fh:
    CASES = json.load(_fh)["cases"]


def _cplx(a):
    a = np.asarray(a)
    return a[..., 0] + 1j * a[..., 1]


def wave_list(nW):
    return [("neg%02d" % i, -1, 0.0) for i in range(nW[0])] + [("pos%02d" % i, +1, 0.0) for i in range(nW[1])]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_golden_vectors_on_gpu(gpu, name, precision):
    """pwaLikelihood on the GPU (integral accumulated on the GPU from the golden
    MC amplitudes) reproduces the golden value and gradient."""
    from rootpwa_amd import likelihood as PL
    c = CASES[name]
    nW = tuple(c["nW"])
    waves = wave_list(nW)
    names = [w[0] for w in waves]
    I = PL.ampIntegralMatrix(names)
    I.integrate(_cplx(c["mc"]))
    sums, n = I.sums()
    assert n == c["mc_events"] and H.rel_err(sums, _cplx(c["integral_sums"])) <= 1e-13
    Iacc = None
    if "acc_sums" in c:
        # acceptance integral over the accepted MC subset, also accumulated on the GPU
        Iacc = PL.ampIntegralMatrix(names)
        Iacc.integrate(_cplx(c["mc"])[np.array(c["acc_index"], dtype=np.int64)])
        sums, n = Iacc.sums()
        assert n == c["acc_events"] and H.rel_err(sums, _cplx(c["acc_sums"])) <= 1e-13
    data = _cplx(c["data"])
    L = PL.initLikelihood(waves, {names[k]: data[:, k] for k in range(len(names))}, I, accIntegral=Iacc,
                          accEventsOverride=c.get("acc_override", 0), rank=c["rank"], precision=precision,
                          cauchyWidth=c["cauchy_width"], useNormalizedAmps=c.get("normalised", True))
    par = np.array(c["par"])
    f, g = L.FdF(par)
    gref = np.array(c["gradient"])
    # complex64 storage: amplitudes rounded to 24-bit mantissas
    tol_f, tol_g = (1e-12, 1e-12) if precision == 0 else (1e-6, 1e-5)
    print(name, precision, f, c["value"], flush=True)
    assert abs(f - c["value"]) <= tol_f * max(1.0, abs(c["value"]))
    assert np.max(np.abs(g - gref)) <= tol_g * np.max(np.abs(gref))
    assert L.DoEval(par) == f


@pytest.mark.parametrize("nEvents", [1, 2, 15, 16, 17, 63, 64, 65, 129, 4097])
def test_ragged_event_counts_every_variant(gpu, nEvents):
    """Event counts at and around the tile sizes (16/32/64 events of the staged
    kernel, lane groups of the register tiles) and below the number of
    workgroups: every variant, both precisions, against the oracle."""
    nW, rank = (3, 11), 2
    ds = H.dataset(nW, rank, nData=max(nEvents, 1), nMC=1000, seed=300 + nEvents)
    L = H.oracle_likelihood(ds)
    amps = H.normalised_amps(L)
    assert amps.shape[0] == nEvents
    prod = H.random_prod(np.random.default_rng(nEvents), rank, nW, 0.8)
    f_ref = L.event_loglik(prod, 0.3)
    d_ref, df_ref = L.event_deriv(prod, 0.3)
    for prec, tol_f, tol_d in ((0, 1e-12, 1e-11), (1, 1e-6, 1e-5)):
        ctx = gpu.LikelihoodContext(0, prec)
        ctx.load(amps, nW)
        for lanes in (1, 8, 16, 32, 64):
            try:
                ctx.set_tile(lanes)
            except gpu.RpwaError as e:
                assert e.code == gpu.RPWA_ERR_UNSUPPORTED
                continue
            print("events", nEvents, "precision", prec, "lanes", lanes, flush=True)
            f, d, df = ctx.fdf(prod, 0.3)
            assert abs(f - f_ref) <= tol_f * max(1.0, abs(f_ref))
            assert np.max(np.abs(d - d_ref)) <= tol_d * np.max(np.abs(d_ref))
            assert abs(df - df_ref) <= tol_d * max(1.0, abs(df_ref))


@pytest.mark.parametrize("nW", [(1, 1), (1, 2), (2, 3)])
@pytest.mark.parametrize("nEvents", [1, 64, 65, 200, 4097])
def test_small_rows_staged_kernel(gpu, nW, nEvents):
    """Rows of 2-5 16-byte units: a staged tile has fewer units than a
    workgroup has threads, so most threads' loads fall past the tile and must
    read inside it (its first unit) rather than past the amplitude array."""
    rank = 1 if nW == (1, 1) else 2
    ds = H.dataset(nW, rank, nData=nEvents, nMC=1000, seed=700 + nEvents + 10 * sum(nW))
    L = H.oracle_likelihood(ds)
    amps = H.normalised_amps(L)
    prod = H.random_prod(np.random.default_rng(nEvents + 1), rank, nW, 0.8)
    f_ref = L.event_loglik(prod, 0.3)
    d_ref, df_ref = L.event_deriv(prod, 0.3)
    for prec, tol_f, tol_d in ((0, 1e-12, 1e-11), (1, 1e-6, 1e-5)):
        ctx = gpu.LikelihoodContext(0, prec)
        ctx.load(amps, nW)
        ctx.set_tile(1)
        print("small rows", nW, "events", nEvents, "precision", prec, flush=True)
        f, d, df = ctx.fdf(prod, 0.3)
        assert abs(f - f_ref) <= tol_f * max(1.0, abs(f_ref))
        assert np.max(np.abs(d - d_ref)) <= tol_d * np.max(np.abs(d_ref))
        assert abs(df - df_ref) <= tol_d * max(1.0, abs(df_ref))


@pytest.mark.parametrize("nEvents", [1, 16, 17, 33, 4097])
def test_every_staged_tile_size(gpu, nEvents):
    """The LDS-staged kernel with tiles of 64, 32 and 16 events (each a
    candidate of the autotuner; with complex64 also the MFMA gradient at 64
    and 32) against the oracle, both precisions."""
    nW, rank = (8, 20), 2
    ds = H.dataset(nW, rank, nData=nEvents, nMC=1000, seed=900 + nEvents)
    L = H.oracle_likelihood(ds)
    amps = H.normalised_amps(L)
    prod = H.random_prod(np.random.default_rng(nEvents + 7), rank, nW, 0.8)
    f_ref = L.event_loglik(prod, 0.3)
    d_ref, df_ref = L.event_deriv(prod, 0.3)
    for prec, tol_f, tol_d in ((0, 1e-12, 1e-11), (1, 1e-6, 1e-5)):
        ctx = gpu.LikelihoodContext(0, prec)
        ctx.load(amps, nW)
        variants = [(e, 0) for e in (64, 32, 16)] + ([(64, 1), (32, 1), (64, 2)] if prec == 1 else [])
        for events, gm in variants:
            ctx.set_stage_variant(events, gm)
            print("staged tile", events, "mfma gradient", gm, "events", nEvents, "precision", prec, flush=True)
            f, d, df = ctx.fdf(prod, 0.3)
            assert abs(f - f_ref) <= tol_f * max(1.0, abs(f_ref))
            assert np.max(np.abs(d - d_ref)) <= tol_d * np.max(np.abs(d_ref))
            assert abs(df - df_ref) <= tol_d * max(1.0, abs(df_ref))
        with pytest.raises(gpu.RpwaError):
            ctx.set_stage_tile(48)
