"""GPU parity of the event kernels (the cuda::likelihoodInterface replacement)
against the oracle's event part, plus the bitwise contracts."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

# (refl -1 waves, refl +1 waves), rank; (30, 170) makes the LDS-staged kernel use 16-event tiles
SHAPES = [((0, 16), 1), ((3, 5), 2), ((8, 80), 2), ((7, 34), 2), ((1, 1), 1), ((5, 0), 1), ((4, 9), 3),
          ((30, 170), 2)]


@pytest.fixture(scope="module", params=SHAPES, ids=lambda p: "W%d+%d_r%d" % (p[0][0], p[0][1], p[1]))
def case(request):
    nW, rank = request.param
    ds = H.dataset(nW, rank, nData=777, nMC=3000, seed=11 + sum(nW) + rank)
    L = H.oracle_likelihood(ds)
    return ds, L, H.normalised_amps(L)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_event_part_matches_oracle(gpu, case, prec):
    ds, L, amps = case
    nW, rank = ds["nW"], ds["rank"]
    print("case", nW, rank, "prec", prec, flush=True)
    ctx = gpu.LikelihoodContext(0, prec)
    ctx.load(amps, nW)
    rng = np.random.default_rng(5)
    prod = H.random_prod(rng, rank, nW, 0.8)
    flat = 0.3
    f_ref = L.event_loglik(prod, flat)
    d_ref, df_ref = L.event_deriv(prod, flat)
    f, d, df = ctx.fdf(prod, flat)
    tol_f, tol_d = (1e-12, 1e-10) if prec == 0 else (1e-6, 1e-5)
    assert abs(f - f_ref) <= tol_f * abs(f_ref)
    assert H.rel_err(d, d_ref) <= tol_d
    assert abs(df - df_ref) <= tol_d * abs(df_ref)
    # value-only and gradient-only entry points agree bitwise with the fused one
    assert ctx.log_likelihood(prod, flat) == f
    d2, df2 = ctx.log_likelihood_deriv(prod, flat)
    assert np.array_equal(d2, d) and df2 == df


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_sign_flip_and_determinism(gpu, case, prec):
    ds, L, amps = case
    nW, rank = ds["nW"], ds["rank"]
    ctx = gpu.LikelihoodContext(0, prec)
    ctx.load(amps, nW)
    rng = np.random.default_rng(9)
    prod = H.random_prod(rng, rank, nW, 1.1)
    f0 = ctx.log_likelihood(prod, 0.2)
    for _ in range(3):
        assert ctx.log_likelihood(prod, 0.2) == f0
    for r in range(rank):
        for s in range(2):
            p2 = prod.copy()
            p2[r, s] *= -1
            assert ctx.log_likelihood(p2, 0.2) == f0
            assert ctx.log_likelihood(prod, -0.2) == f0


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_every_tile_shape_agrees(gpu, case, prec):
    """Every kernel variant (lanes per event 8..64 of the register-tile kernels,
    1 = the LDS-staged kernel) against the oracle, with the bitwise contracts
    (value-only == fused, sign flip of a (rank, sector) block) per variant."""
    ds, L, amps = case
    nW, rank = ds["nW"], ds["rank"]
    ctx = gpu.LikelihoodContext(0, prec)
    ctx.load(amps, nW)
    prod = H.random_prod(np.random.default_rng(3), rank, nW, 0.9)
    f_ref = L.event_loglik(prod, 0.1)
    d_ref, df_ref = L.event_deriv(prod, 0.1)
    tol_f, tol_d = (1e-12, 1e-10) if prec == 0 else (1e-6, 1e-5)
    for lanes in (1, 8, 16, 32, 64):
        try:
            ctx.set_tile(lanes)
        except gpu.RpwaError as e:
            assert e.code == gpu.RPWA_ERR_UNSUPPORTED
            continue
        print("tile", lanes, "lanes per event", flush=True)
        f, d, df = ctx.fdf(prod, 0.1)
        assert abs(f - f_ref) <= tol_f * abs(f_ref), lanes
        assert H.rel_err(d, d_ref) <= tol_d, lanes
        assert abs(df - df_ref) <= tol_d * abs(df_ref), lanes
        assert ctx.log_likelihood(prod, 0.1) == f
        p2 = prod.copy()
        p2[rank - 1, 1] *= -1
        assert ctx.log_likelihood(p2, 0.1) == f


@pytest.mark.parametrize("events,variant", [(64, 1), (32, 1), (64, 2)], ids=["g64", "g32", "wave64"])
def test_stage_gradient_on_mfma(gpu, case, events, variant):
    """The staged kernel with its gradient phase on FP32 MFMA (variant 1:
    complex64 rows of at most 128 waves) or with S on FP64 MFMA and the
    gradient on FP32 MFMA, one wave per 16 events (variant 2: rows of at most
    96 waves, ranks 1-2; rank 3 falls back to variant 0) against the oracle and
    against the FP64-VALU variant of the same tile: the value within the
    complex64 bound (bitwise equal for variant 1 when the grids agree), the
    gradient within 1e-6 of its largest entry; sign flip of a (rank, sector)
    block and the value-only call bitwise."""
    ds, L, amps = case
    nW, rank = ds["nW"], ds["rank"]
    ctx = gpu.LikelihoodContext(0, 1)
    ctx.load(amps, nW)
    prod = H.random_prod(np.random.default_rng(31), rank, nW, 0.9)
    wpad = sum(nW) + (nW[0] % 2) + (nW[1] % 2)
    try:
        ctx.set_stage_variant(events, variant)
    except gpu.RpwaError as e:
        assert e.code == gpu.RPWA_ERR_UNSUPPORTED and wpad > {1: 128, 2: 96}[variant]
        ctx.close()
        return
    print("mfma variant", variant, nW, "rank", rank, "tile", events, flush=True)
    assert ctx.kernel_variant(rank)["gradient_mfma"] == (variant if variant == 1 or rank <= 2 else 0)
    f, d, df = ctx.fdf(prod, 0.1)
    wg = ctx.kernel_variant(rank)["workgroups"]
    f_ref = L.event_loglik(prod, 0.1)
    d_ref, df_ref = L.event_deriv(prod, 0.1)
    assert abs(f - f_ref) <= 1e-6 * abs(f_ref)
    assert H.rel_err(d, d_ref) <= 1e-5
    assert abs(df - df_ref) <= 1e-5 * abs(df_ref)
    assert ctx.log_likelihood(prod, 0.1) == f
    d2, df2 = ctx.log_likelihood_deriv(prod, 0.1)
    assert np.array_equal(d2, d) and df2 == df
    p2 = prod.copy()
    p2[rank - 1, 1] *= -1
    assert ctx.log_likelihood(p2, 0.1) == f
    ctx.set_stage_variant(events, 0)
    fv, dv, dfv = ctx.fdf(prod, 0.1)
    if variant == 1 and ctx.kernel_variant(rank)["workgroups"] == wg:
        assert fv == f and dfv == df
    assert abs(fv - f) <= 1e-12 * abs(f)
    assert np.max(np.abs(d - dv)) <= 1e-6 * np.max(np.abs(dv))
    ctx.close()


def test_stage_gradient_on_mfma_rejected(gpu):
    """FP64 amplitudes, 16-event tiles and rows of more than 128 waves do not
    take the MFMA gradient; the wave-tile variant needs 64-event tiles and rows
    of at most 96 waves; variants other than 0-2 are invalid."""
    rng = np.random.default_rng(5)
    a = rng.normal(size=(100, 20)) + 1j * rng.normal(size=(100, 20))
    ctx = gpu.LikelihoodContext(0, 0)
    ctx.load(a, (4, 16))
    with pytest.raises(gpu.RpwaError):
        ctx.set_stage_variant(64, True)
    ctx.close()
    ctx = gpu.LikelihoodContext(0, 1)
    ctx.load(a, (4, 16))
    with pytest.raises(gpu.RpwaError):
        ctx.set_stage_variant(16, True)
    with pytest.raises(gpu.RpwaError):
        ctx.set_stage_variant(32, 2)
    with pytest.raises(gpu.RpwaError):
        ctx.set_stage_variant(64, 3)
    ctx.set_stage_variant(32, True)
    ctx.set_stage_variant(64, 2)
    ctx.close()
    c = rng.normal(size=(100, 100)) + 1j * rng.normal(size=(100, 100))
    ctx = gpu.LikelihoodContext(0, 1)
    ctx.load(c, (2, 98))  # 50 units: a 64-event tile exceeds the staging loads, 32 events fit
    ctx.set_stage_variant(32, 1)
    with pytest.raises(gpu.RpwaError):
        ctx.set_stage_variant(64, 2)
    ctx.close()
    b = rng.normal(size=(100, 130)) + 1j * rng.normal(size=(100, 130))
    ctx = gpu.LikelihoodContext(0, 1)
    ctx.load(b, (2, 128))
    with pytest.raises(gpu.RpwaError):
        ctx.set_stage_variant(64, True)
    ctx.close()


def test_reshuffled_layout(gpu):
    """The [2][Wmax][N] layout pwaLikelihood::finishInit builds for the GPU."""
    ds = H.dataset((3, 6), 2, 500, 2000, seed=4)
    L = H.oracle_likelihood(ds)
    amps = H.normalised_amps(L)
    nW = ds["nW"]
    wmax = max(nW)
    resh = np.zeros((2, wmax, amps.shape[0]), dtype=np.complex128)
    resh[0, :nW[0]] = amps[:, :nW[0]].T
    resh[1, :nW[1]] = amps[:, nW[0]:].T
    ctx = gpu.LikelihoodContext(0, 0)
    ctx.load(resh, nW, layout=gpu.LAYOUT_REFL_WAVE_EVT, nEvents=amps.shape[0])
    prod = H.random_prod(np.random.default_rng(1), 2, nW)
    assert abs(ctx.log_likelihood(prod, 0.5) - L.event_loglik(prod, 0.5)) <= 1e-12 * abs(L.event_loglik(prod, 0.5))


def test_integral_matches_oracle(gpu):
    rng = np.random.default_rng(2)
    # 57-64, 121-128 and 249-256 waves: the widest chunk buffers of each row padding
    for W, M in ((3, 1000), (16, 5000), (57, 900), (64, 1200), (88, 3001), (110, 700), (121, 300), (128, 400),
                 (256, 200)):
        print("integral W", W, "M", M, flush=True)
        a = rng.normal(size=(M, W)) + 1j * rng.normal(size=(M, W))
        ref = O.integral(a)
        got = gpu.integral(a)
        assert H.rel_err(got, ref) <= 1e-12, W
        w = rng.uniform(0.5, 2.0, size=M)
        assert H.rel_err(gpu.integral(a, w), O.integral(a, w)) <= 1e-12, W


def test_integral_several_chunks_per_workgroup(gpu):
    """Every workgroup of the persistent grid walks at least three event chunks (up to 64 events each), so
    the staged double buffer, the accumulation over chunks and the ragged last chunk are all exercised:
    5 column blocks (W=40), 11 (W=88), 12 (W=96: the pair hand-over in two rounds), and the windowed passes of
    wider rows: 16 blocks (W=128, two windows of 8) and 19 (W=150, three windows, the last one padded), with and
    without weights."""
    rng = np.random.default_rng(21)
    n_wg = gpu.device_info(0)["cus"]  # one workgroup per CU
    for W in (40, 88, 96, 128, 150):
        M = 3 * 64 * n_wg + 17
        print("integral W", W, "M", M, flush=True)
        a = rng.normal(size=(M, W)) + 1j * rng.normal(size=(M, W))
        a *= np.linspace(0.3, 1.7, W)[None, :]
        assert H.rel_err(gpu.integral(a), O.integral(a)) <= 1e-12, W
        w = rng.uniform(0.5, 2.0, size=M)
        assert H.rel_err(gpu.integral(a, w), O.integral(a, w)) <= 1e-12, W


def test_integral_host_array_in_pieces(gpu):
    """rpwa_integral streams a host array through the device in pieces of 256 MB: three pieces here (the last one
    ragged), so the kernels add to the partial tiles of the earlier pieces and the weight sums accumulate; 88 waves
    (grouped kernel) and 150 waves (windowed passes), with importance weights (ampIntegralMatrix.cc:477-483:
    sums renormalised by the mean weight 1/w) and without. Reference: the product A^T conj(A) in numpy."""
    rng = np.random.default_rng(33)
    for W in (88, 150):
        piece = (256 << 20) // (8 * 2 * W)  # events per streamed piece
        M = 2 * piece + 12345
        print("host integral W", W, "M", M, flush=True)
        a = rng.normal(size=(M, W)) + 1j * rng.normal(size=(M, W))
        a *= np.linspace(0.2, 1.8, W)[None, :]
        ref = a.T @ a.conj()
        assert H.rel_err(gpu.integral(a), ref) <= 1e-12, W
        w = rng.uniform(0.5, 2.0, size=M)
        wt = 1.0 / w
        refw = (a * wt[:, None]).T @ a.conj() * (M / wt.sum())
        assert H.rel_err(gpu.integral(a, w), refw) <= 1e-12, W
        del a


def test_integral_incoherent_sub_amplitudes(gpu):
    """ampIntegralMatrix with two incoherent sub-amplitudes per wave
    (ampIntegralMatrix.cc:432-466), host and device input."""
    from rootpwa_amd import likelihood as PL
    rng = np.random.default_rng(21)
    S, M, W = 2, 2500, 12
    a = rng.normal(size=(S, M, W)) + 1j * rng.normal(size=(S, M, W))
    w = rng.uniform(0.5, 2.0, size=M)
    names = ["w%02d" % i for i in range(W)]
    for weights in (None, w):
        m = PL.ampIntegralMatrix(names)
        m.integrate(a, weights)
        sums, n = m.sums()
        assert n == M and H.rel_err(sums, O.integral(a, weights)) <= 1e-12
    buf = gpu.DeviceBuffer(a.nbytes)
    buf.upload(a)
    wbuf = gpu.DeviceBuffer(w.nbytes)
    wbuf.upload(w)
    m = PL.ampIntegralMatrix(names)
    m.integrateDevice(buf.address, M, wbuf.address, nmbSubAmps=S)
    sums, n = m.sums()
    assert n == M and H.rel_err(sums, O.integral(a, w)) <= 1e-12


def test_generator_matches_cpu_restatement(gpu):
    seed, W = 77, 12
    scales = O.gen_wave_scales(seed, W)
    assert np.array_equal(scales, gpu.gen_wave_scales(seed, W))
    buf = gpu.DeviceBuffer(500 * W * 16)
    gpu.gen_amps_device(buf.address, seed, 3, 1000, 500, scales)
    got = buf.download((500, W), np.complex128)
    ref = O.gen_amps(seed, 3, 1000, 500, scales)
    assert H.rel_err(got, ref) <= 1e-13
    nW, rank = (2, 5), 2
    prod = O.gen_prod_amps(seed, 2, rank, nW)
    assert np.array_equal(prod, gpu.gen_prod_amps(seed, 2, rank, nW))
    out = gpu.DeviceBuffer(300 * 7 * 16)
    pool, smax = gpu.gen_accept_reject_device(out.address, seed, 4, 300, nW, rank, prod, scales)
    data, idx, pool_c, smax_c = O.gen_accept_reject(seed, 4, 300, nW, rank, prod, scales)
    assert pool == pool_c
    assert abs(smax - smax_c) <= 1e-12 * smax_c
    assert H.rel_err(out.download((300, 7), np.complex128), data) <= 1e-13


@pytest.mark.parametrize("rank,prec", [(2, 0), (2, 1), (6, 1)], ids=["r2_fp64", "r2_mixed", "r6_mixed_blocks"])
def test_kernel_timing_leaves_results_unchanged(gpu, rank, prec):
    """Timing events are bound to the kernel launches: results with timing on
    are bitwise those with it off (the completion flag path), and the split
    times cover every evaluation, the reduction much shorter than the event
    kernels at 50k events x 8+40 waves."""
    nW = (8, 40)
    rng = np.random.default_rng(77 + rank)
    N = 50_000
    amps = (rng.normal(size=(N, sum(nW))) + 1j * rng.normal(size=(N, sum(nW))))
    print("timing rank", rank, "prec", prec, flush=True)
    ctx = gpu.LikelihoodContext(0, prec)
    ctx.load(amps, nW)
    prod = H.random_prod(rng, rank, nW, 0.7)
    ctx.set_timing(False)
    ref = ctx.fdf(prod, 0.1)
    ctx.reset_kernel_time()
    ctx.set_timing(True)
    for _ in range(4):
        f, d, df = ctx.fdf(prod, 0.1)
        assert f == ref[0] and np.array_equal(d, ref[1]) and df == ref[2]
    ev, red, n = ctx.kernel_times()
    assert n == 4
    assert ev > 0 and red > 0 and red < ev
    assert abs(ctx.kernel_time() - (ev + red)) <= 1e-12 * (ev + red)
    ctx.set_timing(False)
    f, d, df = ctx.fdf(prod, 0.1)
    assert f == ref[0] and np.array_equal(d, ref[1])
    assert ctx.kernel_times()[2] == 4
    # timing every third evaluation: 9 evaluations, 3 of them timed, results unchanged
    ctx.reset_kernel_time()
    ctx.set_timing(3)
    for _ in range(9):
        f, d, df = ctx.fdf(prod, 0.1)
        assert f == ref[0] and np.array_equal(d, ref[1]) and df == ref[2]
    ev3, red3, n3 = ctx.kernel_times()
    assert n3 == 3 and ev3 > 0 and red3 > 0
    ctx.set_timing(False)
    ctx.close()
