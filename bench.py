#!/usr/bin/env python3
"""Benchmark of the likelihood hot path on one MI355X.

Workload (the headline configuration): 2,000,000 accept-rejected data events x
88 waves (8 with reflectivity -1, 80 with +1), rank 2, normalisation integral
from a 5,000,000-event MC pool, all generated on the GPU with the synthetic
amplitude recipe.  A step is one FdF (value + full gradient, the NLopt call)
of rpwa::pwaLikelihood with complex64 amplitudes (mixed precision: FP64
arithmetic, FP64 compensated event sum); the same number of FdF calls is timed
on the all-FP64 drop-in (complex128 amplitudes) for comparison.  Every result
is checked against the CPU oracle at the full size.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("lnL+grad evals/s, 2M ev×88 waves (mixed vs FP64, |ΔlnL|<1e-3); fits/hr")
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6
FP32_MFMA_PEAK_TFS = 157.3
SEED = 20261202 + 1
# kernel timing in the timed FdF loop: HIP events on the launches of every 8th evaluation
TIMING_PERIOD = 8
SEED_BATCH = 20261202 + 3
STREAM_TRUE_T, STREAM_MC, STREAM_DATA, STREAM_START = 2, 3, 4, 5


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def wave_list(nW):
    return [("neg%02d" % i, -1, 0.0) for i in range(nW[0])] + [("pos%02d" % i, +1, 0.0) for i in range(nW[1])]


def measured_traffic(N, W):
    """HBM bytes per launch of the event kernels from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel),
    valid for the default workload only."""
    path = os.path.join(ROOT, "profiles", "event_kernel_2Mx88_pmc.txt")
    if (N, W) != (2_000_000, 88) or not os.path.exists(path):
        return {}
    for line in open(path):
        if line.startswith("TRAFFIC_BYTES_PER_LAUNCH"):
            return {k: float(v) for k, v in (kv.split("=") for kv in line.split()[1:])}
    return {}


def cpu_model():
    """Host CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return "%s, %d usable cores" % (model, len(os.sched_getaffinity(0)))


def c0_leg(args, log):
    """The CPU-runnable configuration: 16 waves in one sector, rank 1, 200k MC
    events, 100k accept-rejected data events; one FP64 FdF at a random point on
    the GPU (rate over repeated calls) and in the single-thread oracle (median of
    5 after a warm-up, full event count), with the value and gradient compared."""
    from oracle import oracle as O
    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL
    nW, rank, N, M = (0, 16), 1, 100_000, 200_000
    W = sum(nW)
    seed = 20261202 + 0
    scales = capi.gen_wave_scales(seed, W)
    truth = capi.gen_prod_amps(seed, STREAM_TRUE_T, rank, nW)
    d_mc = capi.DeviceBuffer(M * W * 16)
    capi.gen_amps_device(d_mc.address, seed, STREAM_MC, 0, M, scales)
    sums = capi.integral_device(d_mc.address, M, W)
    d_mc.free()
    d_data = capi.DeviceBuffer(N * W * 16)
    capi.gen_accept_reject_device(d_data.address, seed, STREAM_DATA, N, nW, rank, truth, scales)
    waves = wave_list(nW)
    intmat = PL.ampIntegralMatrix([w[0] for w in waves])
    intmat.setSums(sums, M)
    L = PL.pwaLikelihood(capi.PREC_FP64)
    L.init(waves, rank)
    L.addNormIntegral(intmat)
    L.addAccIntegral(intmat)
    L.addAmplitudesDevice(d_data.address, N)
    L.finishInit()
    par = np.random.default_rng(seed + STREAM_START).normal(size=L.nmbPars()) * np.sqrt(N / W)
    for _ in range(10):
        f, g = L.FdF(par)
    n = 500
    t = time.perf_counter()
    for _ in range(n):
        f, g = L.FdF(par)
    gpu_rate = n / (time.perf_counter() - t)
    # the same calls from C++ (the native drop-in path, no Python binding per call)
    sec, f_native, _ = L.timeFdF(par, 2000)
    native_rate = 2000 / sec
    assert f_native == f
    data = d_data.download((N, W), np.complex128)
    d_data.free()
    Lo = O.OracleLikelihood(nW, rank)
    Lo.add_integrals(sums, M)
    Lo.add_amplitudes(data)
    Lo.finish_init()
    f_ref, g_ref = Lo.fdf(par)
    ts = []
    for _ in range(5):
        t = time.perf_counter()
        Lo.fdf(par)
        ts.append(time.perf_counter() - t)
    cpu_s = float(np.median(ts))
    out = {"workload": "C0: 100k accept-rejected events x 16 waves (one sector), rank 1, 200k-event MC; FP64 FdF",
           "parameters": int(L.nmbPars()), "gpu_evals_per_s": gpu_rate, "gpu_evals_per_s_native": native_rate,
           "cpu_evals_per_s": 1.0 / cpu_s,
           "cpu_cores": 1, "cpu_kind": "port (single-thread C oracle, median of 5 after a warm-up)",
           "rel_lnL": abs(f - f_ref) / abs(f_ref),
           "grad_rel": float(np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref)))}
    log("C0:", out)
    return out


def kernel_label(info, precision, rank):
    """Name of the event-kernel variant the context's autotuner kept."""
    if info["lanes"] == 1:
        if info.get("stage_variant") == 2:
            return ("wave_tile_kernel<%s, rank %d, %d events per tile> (LDS-staged; each wave: S of 16 events on "
                    "FP64 MFMA, sigma, gradient on FP32 MFMA)" % (precision, rank, info["slots"]))
        if info.get("stage_variant") == 1:
            return ("stage_kernel<%s, rank %d, %d events per tile> (LDS-staged, gradient on FP32 MFMA)"
                    % (precision, rank, info["slots"]))
        return "stage_kernel<%s, rank %d, %d events per tile> (LDS-staged, fused lnL + gradient)" % (
            precision, rank, info["slots"])
    return "event_kernel<%s, %d lanes, %d slots, rank %d> (register tiles, fused lnL + gradient)" % (
        precision, info["lanes"], info["slots"], rank)


def start_values(rng, nPars, nEvents, W):
    """Random start at the scale of the fit (|T|^2 ~ N/W per wave)."""
    return rng.normal(size=nPars) * np.sqrt(nEvents / W)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # the event kernel settles under sustained load only after several hundred
    # evaluations (0.1-0.2 s; it runs 3-5% slower before that)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--events", type=int, default=2_000_000)
    ap.add_argument("--mc-events", type=int, default=5_000_000)
    ap.add_argument("--cpu-sample", type=int, default=400_000, help="events of the single-thread CPU baseline")
    ap.add_argument("--cpu-evals", type=int, default=5)
    ap.add_argument("--oracle-threads", type=int, default=16)
    ap.add_argument("--skip-oracle", action="store_true", help="skip the full-size oracle check (profiling runs)")
    ap.add_argument("--skip-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--skip-batch", action="store_true", help="skip the batched-fit leg")
    ap.add_argument("--skip-hessian", action="store_true", help="skip the analytic-Hessian leg")
    ap.add_argument("--skip-c0", action="store_true", help="skip the small CPU-runnable configuration")
    ap.add_argument("--skip-fit", action="store_true", help="skip the full-size fits from one start")
    ap.add_argument("--only-batch", action="store_true",
                    help="profiling runs: only the batched leg, printed as its own JSON line")
    ap.add_argument("--initializer", action="store_true",
                    help="learned-initializer protocol only (training set of --init-train bins, best of 64 "
                         "starts; evaluation on --init-test held-out bins), printed as its own JSON line")
    ap.add_argument("--skip-variants", action="store_true",
                    help="skip the side-by-side timing of the complex64 staged-kernel variants")
    ap.add_argument("--skip-initializer", action="store_true",
                    help="skip the learned-initializer protocol (about 3.5 minutes of the default run)")
    ap.add_argument("--init-train", type=int, default=2000)
    ap.add_argument("--init-test", type=int, default=200)
    ap.add_argument("--init-epochs", type=int, default=400)
    ap.add_argument("--batch-fits", type=int, default=5,
                    help="timed batched fit runs (different seeds; the mean time is reported)")
    ap.add_argument("--batch-bins", type=int, default=48)
    ap.add_argument("--batch-starts", type=int, default=64)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards over events with one "
              "fixed-order exchange of ~350 doubles per evaluation; no multi-GPU figure is produced); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from rootpwa_amd import capi
    from rootpwa_amd import likelihood as PL

    if capi.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = capi.device_info(0)
    log("device", dev)
    if args.only_batch:
        print(json.dumps({"batched": batched_leg(args, log)}), flush=True)
        return
    if args.initializer:
        print(json.dumps({"initializer": initializer_leg(args, log)}), flush=True)
        return

    read_gbs, copy_gbs = capi.bandwidth_probe(4 << 30, 10)
    log("bandwidth probe: read %.0f GB/s, copy %.0f GB/s (4 GiB)" % (read_gbs, copy_gbs))

    nW = (8, 80)
    W = sum(nW)
    rank = 2
    N, M = args.events, args.mc_events
    waves = wave_list(nW)
    names = [w[0] for w in waves]

    # ---- workload generation on the GPU --------------------------------------
    t0 = time.perf_counter()
    scales = capi.gen_wave_scales(SEED, W)
    truth = capi.gen_prod_amps(SEED, STREAM_TRUE_T, rank, nW)
    d_mc = capi.DeviceBuffer(M * W * 16)
    capi.gen_amps_device(d_mc.address, SEED, STREAM_MC, 0, M, scales)
    log("MC pool generated: %d events (%.1f s)" % (M, time.perf_counter() - t0))
    # normalisation integral on FP64 MFMA: the first calls run while the clock
    # ramps up under the sustained matrix load (about 3.3 ms, then 2.7-2.8 ms)
    for _ in range(5):
        sums = capi.integral_device(d_mc.address, M, W)
    capi.synchronize()
    t1 = time.perf_counter()
    n_int = 10
    for _ in range(n_int):
        sums = capi.integral_device(d_mc.address, M, W)
    t_int = (time.perf_counter() - t1) / n_int
    int_flops = 8.0 * M * W * (W + 1) / 2
    log("integral: %.3f ms per call, %.1f TFLOP/s algorithmic" % (1e3 * t_int, int_flops / t_int / 1e12))
    d_mc.free()
    d_data = capi.DeviceBuffer(N * W * 16)
    t1 = time.perf_counter()
    pool, smax = capi.gen_accept_reject_device(d_data.address, SEED, STREAM_DATA, N, nW, rank, truth, scales)
    log("data: %d events accepted from a pool of %d (%.1f s)" % (N, pool, time.perf_counter() - t1))

    intmat = PL.ampIntegralMatrix(names)
    intmat.setSums(sums, M)

    def make(precision):
        L = PL.pwaLikelihood(precision)
        L.init(waves, rank)
        L.addNormIntegral(intmat)
        L.addAccIntegral(intmat)
        L.addAmplitudesDevice(d_data.address, N)
        L.finishInit()
        return L

    L_mix = make(capi.PREC_MIXED)
    L_64 = make(capi.PREC_FP64)
    nPars = L_mix.nmbPars()
    rng = np.random.default_rng(SEED + STREAM_START)
    par = start_values(rng, nPars, N, W)

    # ---- timed FdF evaluations ------------------------------------------------
    def timed(L, label):
        ctx = L.gpuContext()
        # HIP events bound to the launches of every TIMING_PERIOD-th evaluation in the timed
        # region: the kernel durations of the roofline (events on every launch add ~15 us of
        # host time to each evaluation; the bench reports that rate as well)
        ctx.set_timing(TIMING_PERIOD)
        for _ in range(args.warmup):
            L.FdF(par)
        capi.synchronize()
        ctx.reset_kernel_time()
        t = time.perf_counter()
        for _ in range(args.steps):
            f, g = L.FdF(par)
        capi.synchronize()
        el = time.perf_counter() - t
        ev, red, n = ctx.kernel_times()
        info = ctx.info()
        info["stage_variant"] = ctx.kernel_variant(rank)["gradient_mfma"]
        rate = args.steps / el
        k_avg = ev / max(n, 1)
        log("%s: %.1f evals/s (%.3f ms per FdF; event kernel %.3f ms, reduce %.3f ms; tile %d lanes x %d slots, "
            "%d workgroups)" % (label, rate, 1e3 * el / args.steps, 1e3 * k_avg, 1e3 * red / max(n, 1),
                                info["lanes"], info["slots"], info["workgroups"]))
        # the Minuit2 pattern: DoEval then Gradient at the same point
        n_pair = max(1, args.steps // 2)
        capi.synchronize()
        t = time.perf_counter()
        for _ in range(n_pair):
            L.DoEval(par)
            L.Gradient(par)
        el_pair = time.perf_counter() - t
        log("%s: %.1f DoEval+Gradient pairs/s" % (label, n_pair / el_pair))
        # the library's default mode (no events; completion read from the reduction's flag)
        ctx.set_timing(False)
        capi.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            L.FdF(par)
        rate_default = args.steps / (time.perf_counter() - t)
        log("%s: %.1f evals/s without kernel events" % (label, rate_default))
        # and with the events on the launches of every evaluation
        ctx.set_timing(1)
        capi.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            L.FdF(par)
        rate_every = args.steps / (time.perf_counter() - t)
        ctx.set_timing(False)
        log("%s: %.1f evals/s with kernel events on every evaluation" % (label, rate_every))
        return dict(rate=rate, ms=1e3 * el / args.steps, kernel_s=k_avg, reduce_s=red / max(n, 1), f=f, g=g,
                    info=info, pair_rate=n_pair / el_pair, rate_without_events=rate_default, timed_launches=n,
                    rate_every_timed=rate_every)

    r_mix = timed(L_mix, "mixed (complex64)")
    r_64 = timed(L_64, "fp64 (complex128)")
    variants = None if args.skip_variants else variants_leg(d_data, N, nW, rank, sums, M, truth, args.steps, log)
    hess = None if args.skip_hessian else hessian_leg(L_64, L_mix, par, N, W, rank, log)
    fits = None if args.skip_fit else fit_leg(L_64, L_mix, log)

    # ---- checks against the oracle at full size ---------------------------------
    check = {}
    if not args.skip_oracle:
        from oracle import oracle as O
        t = time.perf_counter()
        data = d_data.download((N, W), np.complex128)
        Lo = O.OracleLikelihood(nW, rank)
        Lo.add_integrals(sums, M)
        Lo.add_amplitudes(data)
        Lo.finish_init()
        f_ref, g_ref = Lo.fdf(par, threads=args.oracle_threads)
        del data
        gscale = float(np.max(np.abs(g_ref)))
        check = dict(
            lnL_oracle=f_ref,
            dlnL_mixed=abs(r_mix["f"] - f_ref), dlnL_fp64=abs(r_64["f"] - f_ref),
            rel_lnL_fp64=abs(r_64["f"] - f_ref) / abs(f_ref),
            grad_rel_mixed=float(np.max(np.abs(r_mix["g"] - g_ref))) / gscale,
            grad_rel_fp64=float(np.max(np.abs(r_64["g"] - g_ref))) / gscale)
        if fits:
            # the FP64 fit's minimum against the oracle: lnL relative, gradient relative to the start's scale
            pm = fits["fp64"]["par"]
            fm_ref, gm_ref = Lo.fdf(pm, threads=args.oracle_threads)
            fm, gm = L_64.FdF(pm)
            fits["fp64"]["oracle_at_minimum"] = dict(
                rel_lnL=abs(fm - fm_ref) / abs(fm_ref),
                grad_err_over_start_scale=float(np.max(np.abs(gm - gm_ref))) / gscale,
                grad_max_abs_oracle=float(np.max(np.abs(gm_ref))))
            log("fit minimum vs oracle:", fits["fp64"]["oracle_at_minimum"])
        log("oracle check (%.1f s, %d threads): %s" % (time.perf_counter() - t, args.oracle_threads, check))
        del Lo

    # ---- CPU baseline: the oracle, single thread, on a bounded sample ------------
    cpu = cpu_all = None
    if not args.skip_cpu:
        from oracle import oracle as O
        S = min(args.cpu_sample, N)
        sub = d_data.download((S, W), np.complex128)
        Lc = O.OracleLikelihood(nW, rank)
        Lc.add_integrals(sums, M)
        Lc.add_amplitudes(sub)
        Lc.finish_init()
        del sub
        Lc.fdf(par)  # warm-up

        def median_time(fn, n):
            ts = []
            for _ in range(n):
                t = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t)
            return float(np.median(ts))
        el = median_time(lambda: Lc.fdf(par), args.cpu_evals)
        el_pair = median_time(lambda: (Lc.doeval(par), Lc.gradient(par)), args.cpu_evals)
        threads = args.oracle_threads
        el_mt = median_time(lambda: Lc.fdf(par, threads=threads), args.cpu_evals)
        cpu = dict(value=(1.0 / el) * S / N, unit="evals/s", cores=1, kind="port",
                   sample="median of %d FdF calls (after one warm-up) of the single-thread C oracle (restatement "
                          "of pwaLikelihood::FdF, Kahan sums, per-event allocation) on the first %d of the %d "
                          "events, %.3f s each; evals/s scaled by %d/%d to the full event count"
                          % (args.cpu_evals, S, N, el, S, N),
                   minuit_pattern_evals_per_s=(1.0 / el_pair) * S / N,
                   cpu_model=cpu_model())
        cpu_all = dict(value=(1.0 / el_mt) * S / N, unit="evals/s", cores=threads, kind="port",
                       sample="the same oracle FdF with the events split into %d contiguous ranges on %d threads, "
                              "partials merged in fixed order (median of %d, %.4f s each on %d events)"
                              % (threads, threads, args.cpu_evals, el_mt, S))
        log("cpu baseline:", cpu, "all cores:", cpu_all)

    c0 = None if args.skip_c0 else c0_leg(args, log)
    batched = None if args.skip_batch else batched_leg(args, log)
    init_rep = None if args.skip_initializer else initializer_leg(args, log)

    bytes_mix = float(N) * W * 8
    bytes_64 = float(N) * W * 16
    ach_mix = bytes_mix / r_mix["kernel_s"] / 1e9
    ach_64 = bytes_64 / r_64["kernel_s"] / 1e9
    traffic = measured_traffic(N, W)
    out = {
        "metric": METRIC,
        "value": r_mix["rate"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r_mix["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "amplitude_storage": "complex64",
        "data": "synthetic (GPU-generated amplitudes, accept-reject pseudo data)",
        "config": {
            "workload": "2M accept-rejected events x 88 waves (8 refl -1 + 80 refl +1), rank 2, "
                        "5M-event MC normalisation; one pwaLikelihood::FdF per step, complex64 amplitudes",
            "events": N, "waves": W, "waves_refl": list(nW), "rank": rank, "mc_events": M, "parameters": nPars,
            "data_pool": pool, "seed": SEED, "device": dev["arch"], "cus": dev["cus"],
        },
        "roofline": {"bound": "hbm", "achieved": ach_mix, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_mix / HBM_PEAK_GBS, "traffic": traffic.get("mixed"),
                     "kernel": kernel_label(r_mix["info"], "mixed", rank),
                     "kernel_ms": 1e3 * r_mix["kernel_s"], "bytes_per_launch": bytes_mix,
                     "measured_read_GBs": read_gbs, "frac_of_measured_read": ach_mix / read_gbs},
        "fp64": {"evals_per_s": r_64["rate"], "ms_per_eval": r_64["ms"], "kernel_ms": 1e3 * r_64["kernel_s"],
                 "kernel": kernel_label(r_64["info"], "fp64", rank),
                 "achieved_GBs": ach_64, "frac": ach_64 / HBM_PEAK_GBS, "frac_of_measured_read": ach_64 / read_gbs,
                 "traffic": traffic.get("fp64")},
        "mixed_over_fp64": r_mix["rate"] / r_64["rate"],
        # kernel timing of the timed loop (value, kernel_ms): HIP events bound to the
        # launches of every TIMING_PERIOD-th evaluation; the same FdF loop with timing off
        # (the library's default: no events, completion read from the reduction's flag)
        # and with events on every evaluation
        "kernel_timing": {"every_nth_evaluation": TIMING_PERIOD,
                          "timed_launches": {"mixed": r_mix["timed_launches"], "fp64": r_64["timed_launches"]}},
        "without_kernel_events": {"mixed_evals_per_s": r_mix["rate_without_events"],
                                  "fp64_evals_per_s": r_64["rate_without_events"]},
        "with_kernel_events_every_evaluation": {"mixed_evals_per_s": r_mix["rate_every_timed"],
                                                "fp64_evals_per_s": r_64["rate_every_timed"],
                                                "mixed_over_fp64": r_mix["rate_every_timed"] / r_64["rate_every_timed"]},
        "integral": {"ms": 1e3 * t_int, "tflops_algorithmic": int_flops / t_int / 1e12,
                     "frac_fp64_mfma": int_flops / t_int / 1e12 / FP64_MFMA_PEAK_TFS, "mc_events": M},
        "bandwidth_probe": {"read_GBs": read_gbs, "copy_GBs": copy_gbs, "bytes": 4 << 30},
        "check": check,
        "cpu_baseline": cpu,
        "cpu_all_cores": cpu_all,
        "minuit_pattern": {"mixed_evals_per_s": r_mix["pair_rate"], "fp64_evals_per_s": r_64["pair_rate"],
                           "note": "one DoEval + Gradient pair per eval at the same point (Minuit2 pattern)"},
        "c0": c0,
        "batched": batched,
        "initializer": init_rep,
        "stage_variants": variants,
        "hessian": hess,
        "fits": None if fits is None else {k: {kk: vv for kk, vv in v.items() if kk != "par"} if isinstance(v, dict)
                                           else v for k, v in fits.items()},
    }
    if cpu:
        out["vs_cpu"] = r_mix["rate"] / cpu["value"]
    print(json.dumps(out), flush=True)


def variants_leg(d_data, N, nW, rank, sums, M, truth, steps, log):
    """The complex64 LDS-staged kernel variants side by side on the same
    amplitudes, in one process (kernel times differ by several per cent from
    box to box): tiles of 64 and 32 events, gradient phase on FP64 VALU or on
    FP32 MFMA, and the wave-tile variant (S on FP64 MFMA, gradient on FP32
    MFMA, one wave per 16 events); event-kernel time (HIP events bound to the
    launches) and FdF rate of the C ABI, lnL and gradient against the first."""
    from rootpwa_amd import capi
    norm = np.sqrt(np.real(np.diag(sums)) / M)
    prod = np.asarray(truth) * 1.1
    ctxs = {}
    names = {0: "valu", 1: "mfma_gradient", 2: "wave_tile_mfma"}
    for events, variant in ((64, 0), (64, 1), (64, 2), (32, 0), (32, 1)):
        c = capi.LikelihoodContext(0, capi.PREC_MIXED)
        c.load_device(d_data.address, N, nW, norm)
        c.set_stage_variant(events, variant)
        c.set_timing(True)
        ctxs["%d_events_%s" % (events, names[variant])] = c
    ref = None
    out = {}
    for rep in range(3):
        for name, c in ctxs.items():
            for _ in range(3):
                f, d, df = c.fdf(prod, 0.1)
            if ref is None:
                ref = (f, d)
            c.reset_kernel_time()
            capi.synchronize()
            t = time.perf_counter()
            for _ in range(steps):
                c.fdf(prod, 0.1)
            el = (time.perf_counter() - t) / steps
            ev, red, n = c.kernel_times()
            o = out.setdefault(name, {"kernel_ms": [], "evals_per_s": []})
            o["kernel_ms"].append(1e3 * ev / n)
            o["evals_per_s"].append(1.0 / el)
            o["lnL_minus_first"] = f - ref[0]
            o["grad_vs_first"] = float(np.max(np.abs(d - ref[1])) / np.max(np.abs(ref[1])))
    for name, o in out.items():
        o["kernel_ms"] = float(np.median(o["kernel_ms"]))
        o["evals_per_s"] = float(np.median(o["evals_per_s"]))
        log("staged variant %s: kernel %.4f ms, %.1f evals/s" % (name, o["kernel_ms"], o["evals_per_s"]))
    for c in ctxs.values():
        c.close()
    return out


def fit_leg(L_64, L_mix, log):
    """A Minuit-pattern fit (pwaFit: DoEval, then Gradient) of the FP64 likelihood from one random
    start at full size, with the Hessian check and covariance; and the NLopt-pattern fit (pwaNloptFit:
    FdF) of both the FP64 and the complex64 likelihood from the start that driver draws for the same
    seed, to show that complex64 storage reaches the same minimum."""
    from rootpwa_amd import likelihood as PL
    out = {}
    for key, L, fn in (("fp64", L_64, "pwaFit"), ("fp64_nlopt", L_64, "pwaNloptFit"), ("mixed", L_mix, "pwaNloptFit")):
        L.gpuContext().set_timing(True)
        L.gpuContext().reset_kernel_time()
        t = time.perf_counter()
        # each driver draws its random start values from the seed, as the reference's drivers do; the
        # NLopt pair runs to tight tolerances so that its minima compare the precisions, not the stopping
        tight = dict(xtolRel=1e-10, ftolAbs=1e-9) if fn == "pwaNloptFit" else {}
        r = (PL.pwaFit if fn == "pwaFit" else PL.pwaNloptFit)(L, seed=SEED, checkHessian=True, **tight)
        wall = time.perf_counter() - t
        ev, red, n = L.gpuContext().kernel_times()
        # the minimisation alone: no Hessian check, no covariance
        t = time.perf_counter()
        r0 = (PL.pwaFit if fn == "pwaFit" else PL.pwaNloptFit)(L, seed=SEED, saveSpace=True, **tight)
        wall_min = time.perf_counter() - t
        out[key] = dict(driver=fn, options=tight or "driver defaults", wall_s=wall, minimisation_s=wall_min, lnL=r.logLikelihood(),
                        converged=r.converged(), edm=r.edm if np.isfinite(r.edm) else None,
                        evaluations=r.nmbEvaluations, iterations=r.nmbIterations,
                        kernel_passes=n, kernel_s=ev + red, covariance=r.covMatrixValid(),
                        non_positive_eigenvalues=r.nmbNonPositiveEigenvalues, par=r.fitParameters,
                        same_minimum_without_hessian=r0.logLikelihood() == r.logLikelihood())
        log("%s fit (%s): %.2f s (minimisation alone %.2f s), lnL %.6f, converged %s, %d evaluations, "
            "%d kernel passes" % (key, fn, wall, wall_min, r.logLikelihood(), r.converged(), r.nmbEvaluations, n))
    # the same driver and start on both precisions: the difference of the minima is the precision's alone
    # (pwaFit and pwaNloptFit stop at different tolerances from different starts)
    f_at = L_64.DoEval(out["mixed"]["par"])
    out["mixed_vs_fp64_same_driver"] = dict(
        dlnL_minima=abs(out["mixed"]["lnL"] - out["fp64_nlopt"]["lnL"]),
        fp64_lnL_at_mixed_minimum_minus_fp64_minimum=f_at - out["fp64_nlopt"]["lnL"],
        mixed_lnL_error_at_its_minimum=abs(out["mixed"]["lnL"] - f_at),
        max_parameter_difference=float(np.max(np.abs(out["mixed"]["par"] - out["fp64_nlopt"]["par"]))))
    log("complex64 vs FP64 minimum (pwaNloptFit, same start):", out["mixed_vs_fp64_same_driver"])
    return out


def hessian_leg(L_64, L_mix, par, N, W, rank, log):
    """Analytic Hessian of -lnL (pwaLikelihood::Hessian) at the benchmark point on
    the full workload, FP64 and mixed: wall time of one call (after one untimed
    call), and a size-independent check: H v against the central difference of
    the analytic gradient along a random direction v (FdF at par +- eps v)."""
    from rootpwa_amd import capi
    out = {}
    rng = np.random.default_rng(SEED + 7)
    v = rng.normal(size=par.size)
    v /= np.linalg.norm(v)
    for label, L in (("fp64", L_64), ("mixed", L_mix)):
        L.Hessian(par)
        capi.synchronize()
        t = time.perf_counter()
        H = L.Hessian(par)
        capi.synchronize()
        el = time.perf_counter() - t
        eps = 1e-3 * np.sqrt(N / W)
        gp = L.FdF(par + eps * v)[1]
        gm = L.FdF(par - eps * v)[1]
        fd = (gp - gm) / (2 * eps)
        hv = H @ v
        out[label] = {"ms": 1e3 * el, "asymmetry": float(np.max(np.abs(H - H.T)) / np.max(np.abs(H))),
                      "hv_vs_gradient_difference": float(np.max(np.abs(hv - fd)) / np.max(np.abs(fd)))}
        log("hessian %s: %.1f ms per call (%d x %d), |Hv - dg|/|dg| = %.2e" % (
            label, 1e3 * el, par.size, par.size, out[label]["hv_vs_gradient_difference"]))
    out["parameters"] = int(par.size)
    out["events"] = int(N)
    out["rank"] = rank
    return out


def initializer_leg(args, log):
    """Learned initializer: an MLP (3 x 512 GELU, FP32, trained by the library's
    kernels) from per-bin data moments (88 diagonal + 256 off-diagonal Re/Im)
    to start values; trained on best-of-64 fits of --init-train generated bins
    (the batched workload's recipe), evaluated on --init-test held-out bins:
    prediction + 7 random starts vs 64 random starts."""
    from rootpwa_amd import initializer as I
    from rootpwa_amd import synth
    W = 88
    seed_train, seed_test = SEED_BATCH + 100, SEED_BATCH + 200
    train_sizes = synth.bin_sizes(seed_train, args.init_train)
    test_sizes = synth.bin_sizes(seed_test, args.init_test)
    t = time.perf_counter()
    data = I.fitted_bins(seed_train, train_sizes, W, nStarts=64, fitSeed=11, log=log)
    t_data = time.perf_counter() - t
    init = I.Initializer(W, nPairs=256, nHidden=512, nHiddenLayers=3, seed=7)
    t = time.perf_counter()
    hist = init.train(data["moments"], data["best"], data["nEvents"], epochs=args.init_epochs, batchSize=64,
                      lr=3e-4)
    t_train = time.perf_counter() - t
    log("initializer trained: %d bins, loss %.4f -> %.4f, validation %.4f (%.1f s)"
        % (len(train_sizes), hist[0], hist[-1], init.validationLoss, t_train))
    rep = I.evaluate(init, seed_test, test_sizes, W, nRandom=64, nMixed=8, fitSeed=13, log=log)
    per_bin = rep.pop("per_bin")
    rep.update({"train_bins": len(train_sizes), "training_data_s": t_data, "training_s": t_train,
                "epochs": args.init_epochs, "validation_loss": float(init.validationLoss),
                "train_loss_first_last": [float(hist[0]), float(hist[-1])],
                "network": "775 (600 moments + 175 spectral) -> 3 x 512 GELU -> 176 (correction to the spectral estimate), FP32, Adam lr 3e-4, batch 64",
                "seeds": {"train": seed_train, "test": seed_test}})
    log("initializer evaluation:", rep)
    return rep


def batched_leg(args, log):
    """Batched fits: 48 bins with data counts log-uniform in [20k, 400k], MC pools
    of 5x, 88 waves in one sector, rank 1, 64 random starts per bin (3072 fits),
    complex64 amplitudes, FP32-MFMA contraction; one timed rpwa_batch_fit."""
    from rootpwa_amd import batch as B
    from rootpwa_amd import synth
    W, S = 88, args.batch_starts
    sizes = synth.bin_sizes(SEED_BATCH, args.batch_bins)
    bf = B.BatchFit(W, S)
    t = time.perf_counter()
    synth.fill_batch(bf, SEED_BATCH, sizes, log=log)
    gen_s = time.perf_counter() - t
    bf.finish()
    ev_ms = bf.timeEventKernel(10)
    flops = float(sizes.sum()) * W * S * 16
    ach = flops / (ev_ms * 1e-3) / 1e12
    log("batched: %d bins, %d events, event kernel %.3f ms = %.1f TFLOP/s algorithmic (%.1f%% of FP32 MFMA peak)"
        % (len(sizes), int(sizes.sum()), ev_ms, ach, 100 * ach / FP32_MFMA_PEAK_TFS))
    bf.fit(seed=1)  # untimed: graph instantiation and warm caches
    # a batched fit's wall time is set by its slowest fit (its iterations run
    # until every fit has converged), which moves with the start seed and with
    # rounding-level changes of the inputs: the figure is taken over several
    # batched fits with different start seeds
    runs = max(1, args.batch_fits)
    walls, iters, ne_runs = [], [], []
    for rep in range(runs):
        t = time.perf_counter()
        par, val, st, ne = bf.fit(seed=2 + rep)
        walls.append(time.perf_counter() - t)
        iters.append(int(bf.info()["lastEvaluations"]))
        ne_runs.append(ne)
    el = float(np.mean(walls))
    info = bf.info()
    ne_all = np.concatenate([x.ravel() for x in ne_runs])
    nfits = len(sizes) * S
    best = val.min(axis=1)
    within = (val <= best[:, None] + 0.5).sum(axis=1)
    u, c = np.unique(st, return_counts=True)
    out = {
        "fits_per_hr": nfits / el * 3600.0, "wall_s": el, "fits": nfits, "bins": len(sizes), "starts": S,
        "waves": W, "events": int(sizes.sum()), "batched_fits_timed": runs, "start_seeds": [2 + r for r in range(runs)],
        "wall_s_per_batched_fit": walls, "iterations": iters,
        "evals_per_fit_mean": float(ne_all.mean()), "evals_per_fit_max": int(ne_all.max()),
        "status": {str(k): int(v) for k, v in zip(u, c)},
        "starts_within_0.5_of_best": {"mean": float(within.mean()), "min": int(within.min())},
        "generation_s": gen_s, "seed": SEED_BATCH,
        "roofline": {"bound": "mfma", "achieved": ach, "peak": FP32_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                     "frac": ach / FP32_MFMA_PEAK_TFS, "kernel": "batch_event_kernel (persistent, FP32 MFMA)",
                     "kernel_ms": ev_ms, "flops_per_launch": flops, "workgroups": info["workgroups"]},
    }
    log("batched fits: %.3g fits/hr (%.3f s per batched fit of %d fits, mean of %d start seeds; iterations %s; "
        "%.1f evaluations per fit)" % (out["fits_per_hr"], el, nfits, runs, iters, ne_all.mean()))
    if not args.skip_oracle:
        # the batch's lnL and gradient against the oracle in the smallest and the
        # largest bin, at the timed fit's start point of the best start and at its
        # minimum (production amplitudes rounded to FP32, as the batch contracts
        # them; data regenerated from the same counters).  The gradient error is
        # relative to the gradient's size at the start: at the minimum the
        # gradient itself vanishes.
        from oracle import oracle as O

        def fp32_pars(x):
            x = x.copy()
            x[..., :-1] = x[..., :-1].astype(np.float32).astype(np.float64)
            return x
        pts = fp32_pars(par)
        p0 = fp32_pars(bf.startValues(2 + runs - 1))
        v, g = bf.eval(pts)
        v0, g0 = bf.eval(p0)
        checks = []
        for b in (int(np.argmin(sizes)), int(np.argmax(sizes))):
            n = int(sizes[b])
            d_data, sums, M, _ = synth.generate_bin(synth.bin_seed(SEED_BATCH, b), W, n)
            data = d_data.download((n, W), np.complex128)
            d_data.free()
            Lo = O.OracleLikelihood((0, W), 1)
            Lo.add_integrals(sums, M)
            Lo.add_amplitudes(data)
            Lo.finish_init()
            del data
            s_best = int(np.argmin(val[b]))
            f0_ref, g0_ref = Lo.fdf(p0[b, s_best], threads=args.oracle_threads)
            f_ref, g_ref = Lo.fdf(pts[b, s_best], threads=args.oracle_threads)
            gscale = float(np.max(np.abs(g0_ref)))
            checks.append({"bin": b, "events": n,
                           "start": {"dlnL": abs(v0[b, s_best] - f0_ref),
                                     "grad_rel": float(np.max(np.abs(g0[b, s_best] - g0_ref))) / gscale},
                           "minimum": {"dlnL": abs(v[b, s_best] - f_ref),
                                       "grad_err_over_start_scale":
                                           float(np.max(np.abs(g[b, s_best] - g_ref))) / gscale,
                                       "grad_max_abs": float(np.max(np.abs(g_ref)))}})
            del Lo
        out["check"] = checks
        log("batched oracle check:", checks)
    if not args.skip_cpu:
        # the oracle's single-thread FdF on one bin's likelihood (rank 1, 88 waves), timed
        # on a sample and scaled to the mean bin size and the mean evaluations per fit
        from oracle import oracle as O
        n = 40_000
        d_data, sums, M, _ = synth.generate_bin(synth.bin_seed(SEED_BATCH, 999), W, n)
        data = d_data.download((n, W), np.complex128)
        d_data.free()
        Lc = O.OracleLikelihood((0, W), 1)
        Lc.add_integrals(sums, M)
        Lc.add_amplitudes(data)
        Lc.finish_init()
        p = np.random.default_rng(3).normal(size=Lc.nPars) * np.sqrt(n / W)
        Lc.fdf(p)
        reps = 3
        t = time.perf_counter()
        for _ in range(reps):
            Lc.fdf(p)
        s_eval = (time.perf_counter() - t) / reps * (float(sizes.mean()) / n)
        cpu_fits_hr = 3600.0 / (s_eval * float(ne.mean()))
        out["cpu_baseline"] = {
            "value": cpu_fits_hr, "unit": "fits/hr", "cores": 1, "kind": "port",
            "sample": "%d FdF calls of the single-thread C oracle on a %d-event bin (88 waves, rank 1), "
                      "scaled to the mean bin size %d and %.1f evaluations per fit" % (reps, n, int(sizes.mean()),
                                                                                     ne.mean())}
        out["vs_cpu"] = out["fits_per_hr"] / cpu_fits_hr
        log("batched cpu baseline: %.3g fits/hr" % cpu_fits_hr)
    return out


if __name__ == "__main__":
    main()
