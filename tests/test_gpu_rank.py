"""Ranks above 4 at the C ABI: the staged kernel in blocks of 4 ranks (sigma of
every block, then every block's gradient entries) against the oracle's event
sums, for every staged tile size, event counts below one tile up to several
tiles per workgroup, both precisions; bitwise determinism and equality of the
value-only and value+gradient calls."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[1, 65, 70001], ids=lambda n: "events%d" % n)
def sample(request):
    nW = (3, 11)
    N = request.param
    W = sum(nW)
    scales = O.gen_wave_scales(71, W)
    mc = O.gen_amps(71, 3, 0, 3000, scales)
    data = O.gen_amps(71, 4, 0, N, scales)
    Lo = {}
    for rank in (5, 6, 9):
        L = O.OracleLikelihood(nW, rank)
        L.add_integrals(O.integral(mc), 3000)
        L.add_amplitudes(data)
        L.finish_init()
        Lo[rank] = L
    return nW, Lo, H.normalised_amps(Lo[5])


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_high_rank_event_sums(gpu, sample, precision):
    nW, Lo, amps = sample
    tol_f, tol_d = (1e-12, 1e-10) if precision == 0 else (1e-6, 1e-5)
    for rank in (5, 6, 9):
        prod = H.random_prod(np.random.default_rng(rank), rank, nW, 0.7)
        f_ref = Lo[rank].event_loglik(prod, 0.3)
        d_ref, df_ref = Lo[rank].event_deriv(prod, 0.3)
        ctx = gpu.LikelihoodContext(0, precision)
        ctx.load(amps, nW)
        # the default (no tile fixed): the largest staged tile
        f, d, df = ctx.fdf(prod, 0.3)
        assert ctx.log_likelihood(prod, 0.3) == f
        f2, d2, df2 = ctx.fdf(prod, 0.3)
        assert f2 == f and np.array_equal(d2, d) and df2 == df
        variants = [(e, 0) for e in (64, 32, 16)] + ([(64, 1), (32, 1), (64, 2)] if precision == 1 else [])
        for events, gm in variants:
            ctx.set_stage_variant(events, gm)
            print("rank", rank, "precision", precision, "events", amps.shape[0], "tile", events, "mfma", gm,
                  flush=True)
            f, d, df = ctx.fdf(prod, 0.3)
            assert abs(f - f_ref) <= tol_f * max(1.0, abs(f_ref))
            assert np.max(np.abs(d - d_ref)) <= tol_d * np.max(np.abs(d_ref))
            assert abs(df - df_ref) <= tol_d * max(1.0, abs(df_ref))
            assert ctx.log_likelihood(prod, 0.3) == f
        # a lower rank on the same context afterwards still uses its own kernel
        p2 = prod[:2]
        f, d, df = ctx.fdf(p2, 0.3)
        L2 = O.OracleLikelihood(nW, 2)
        L2.add_integrals(np.eye(sum(nW)), 1)
        L2.add_amplitudes(amps)
        L2.finish_init()
        assert abs(f - L2.event_loglik(p2, 0.3)) <= tol_f * max(1.0, abs(f))
        ctx.close()
