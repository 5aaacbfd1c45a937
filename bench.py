#!/usr/bin/env python3
"""Throughput of the all-pairs normalized cross-correlation on one MI355X.

Workload: N=4096 events x 3 channels x 2048
float32 samples, seeded noise + Ricker events, FFT period P=4096, full lag
range (-2047..2048), channel coefficients averaged per lag.  One step = one
pass of the hot path over the resident batch: batched R2C of all 12288
traces, the fused pair kernel over all N(N+1)/2 pairs (conj-multiply +
channel stack + C2R + max/argmax, min/argmin), and the FP64 near-tie
resolver.  Inputs are resident in HBM before the timed region.

Prints one JSON line (stdout).  value = unique pairs N(N-1)/2 per second.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "pairs/sec N=4096 L=2048: FP32-FFT vs MFMA-bf16x3; serve p50/p99 ms"
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
LAG_WINDOW = 512  # the lag-window forms are compared at +-512


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fft_flops_per_pair(C: int, P: int) -> float:
    """Algorithmic work of the FFT form per pair: C conj-multiply-accumulates
    (8 flops per bin) plus one stacked C2R (2.5 P log2 P)."""
    B = P // 2 + 1
    return C * 8.0 * B + 2.5 * P * math.log2(P)


def fp64_reference(ev, P, n_pairs, seed, windows):
    """FP64 numpy reference on seeded packed pairs (i <= j): the channel-
    averaged normalized period-P circular correlation (the FFT form's
    definition; equal to the linear one on every lag with |l| <= P - L), its
    maximum above 0 and minimum below 0 with the first (smallest) lag, over
    each window (lo, hi).  Returns the packed indices and per window
    {pos, lag_pos, neg, lag_neg, gap_pos, gap_neg}."""
    N, C, L = ev.shape
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.integers(0, N, size=n_pairs)
    b = rng.integers(0, N, size=n_pairs)
    pi, pj = np.minimum(a, b), np.maximum(a, b)
    idx = pi.astype(np.int64) * N - pi.astype(np.int64) * (pi - 1) // 2 + (pj - pi)
    x = ev.astype(np.float64)
    norms = np.linalg.norm(x, axis=2)                      # [N, C]
    used = np.unique(np.concatenate([pi, pj]))
    slot = np.full(N, -1, np.int64)
    slot[used] = np.arange(used.size)
    X = np.fft.rfft(x[used], P, axis=2) / norms[used][:, :, None]
    r = np.fft.irfft((X[slot[pi]] * np.conj(X[slot[pj]])).sum(axis=1), P, axis=1) / C
    out = {}
    for lo, hi in windows:
        lags = np.arange(lo, hi + 1)
        cur = r[:, lags % P]
        srt = np.sort(cur, axis=1)
        kp, kn = cur.argmax(axis=1), cur.argmin(axis=1)
        vp, vn = cur[np.arange(n_pairs), kp], cur[np.arange(n_pairs), kn]
        out[(lo, hi)] = {"pos": np.where(vp > 0, vp, 0.0), "lag_pos": np.where(vp > 0, lags[kp], 0),
                         "neg": np.where(vn < 0, vn, 0.0), "lag_neg": np.where(vn < 0, lags[kn], 0),
                         "gap_pos": srt[:, -1] - srt[:, -2], "gap_neg": srt[:, 1] - srt[:, 0]}
    return idx, out


def vs_fp64(outs, idx, ref):
    """Coefficient error and lag agreement of packed outputs (vp, lp, vn, ln)
    against the FP64 reference on the sampled pairs."""
    vp, lp, vn, ln = outs
    dp, dn = lp[idx] != ref["lag_pos"], ln[idx] != ref["lag_neg"]
    agree = ~dp & ~dn
    sep = (ref["gap_pos"] >= 1e-5) & (ref["gap_neg"] >= 1e-5)
    gaps = np.concatenate([ref["gap_pos"][dp], ref["gap_neg"][dn]])
    return {"pairs": int(idx.size),
            "max_abs_dcoef": float(max(np.abs(vp[idx] - ref["pos"]).max(), np.abs(vn[idx] - ref["neg"]).max())),
            "lag_agreement": float(agree.mean()),
            "fp64_near_ties_below_1e-5": int((~sep).sum()),
            # pairs whose top two FP64 values differ by at least 1e-5 on both sides
            "lag_agreement_separated": float(agree[sep].mean()) if sep.any() else None,
            # largest FP64 top-two gap on a side whose lag differs (None: none differ)
            "disagreeing_side_max_fp64_gap": float(gaps.max()) if gaps.size else None}


def time_plan(plan, tr, outs, stream, steps, warmup):
    from seismic_amd.device import synchronize
    for _ in range(warmup):
        plan.execute(tr, out=outs, stream=stream)
    stream.synchronize()
    plan.enable_timing(True)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute(tr, out=outs, stream=stream)
    stream.synchronize()
    synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, plan.timing(), plan.stats()


def lag_forms(ev, tr, stream, steps, warmup, f64=None):
    """MFMA lag-window forms and the FFT form on the same +-LAG_WINDOW window;
    f64 = (packed indices, FP64 reference on that window) adds each form's
    error against FP64."""
    from seismic_amd import engine
    N, C, L = ev.shape
    w = LAG_WINDOW
    out = {}
    for name, form in (("bf16x3", engine.FORM_LAG_BF16X3), ("fp16", engine.FORM_LAG_FP16)):
        plan = engine.AllPairsPlan(N, C, L, 0, -w, w, threshold=0.0, form=form)
        outs = plan.alloc_outputs()
        ms, tm, st = time_plan(plan, tr, outs, stream, steps, warmup)
        k = 3 if name == "bf16x3" else 1
        T = (N + 31) // 32
        n_tiles = T * (T + 1) // 2
        lag_first = -w
        while (lag_first - 1) % 8:
            lag_first -= 1
        n_lags_issued = 8 * ((w - lag_first + 1 + 7) // 8)
        l16 = (L + 15) // 16 * 16
        issued = n_tiles * 1024 * k * 2.0 * n_lags_issued * l16 * C
        useful = N * (N - 1) / 2 * 2.0 * ((2 * w + 1) * L - w * (w + 1)) * C
        kern_s = tm["ms_pairs"] / 1e3
        out[name] = {
            "value": round(N * (N - 1) / 2 / (ms / 1e3), 1), "unit": "pairs/s",
            "ms_per_step": round(ms, 3), "lags": [-w, w],
            "stages_ms": {k2: round(v, 4) for k2, v in tm.items() if k2 != "executes"},
            "near_ties_per_step": st["near_ties"],
            "roofline": {"bound": "mfma", "achieved": round(issued / kern_s / 1e12, 2),
                         "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(issued / kern_s / 1e12 / PEAK_BF16_TFLOPS, 4),
                         "issued_flops_per_launch": issued,
                         "useful_tflops": round(useful / kern_s / 1e12, 2)},
            "_outs": [o.numpy() for o in outs],
        }
        plan.close()
    plan = engine.AllPairsPlan(N, C, L, 4096, -w, w, threshold=0.0)
    outs = plan.alloc_outputs()
    ms, tm, st = time_plan(plan, tr, outs, stream, steps, warmup)
    out["fft_same_window"] = {"value": round(N * (N - 1) / 2 / (ms / 1e3), 1), "unit": "pairs/s",
                              "ms_per_step": round(ms, 3), "_outs": [o.numpy() for o in outs]}
    plan.close()
    # agreement with the FFT form on the same window (both against each other)
    if f64 is not None:
        for name in ("bf16x3", "fp16", "fft_same_window"):
            out[name]["vs_fp64"] = vs_fp64(out[name]["_outs"], *f64)
    ref = out["fft_same_window"].pop("_outs")
    for name in ("bf16x3", "fp16"):
        o = out[name].pop("_outs")
        out[name]["max_abs_dcoef_vs_fft_f32"] = float(max(np.abs(o[0] - ref[0]).max(),
                                                          np.abs(o[2] - ref[2]).max()))
        out[name]["lag_agreement_vs_fft_f32"] = float(((o[1] == ref[1]) & (o[3] == ref[3])).mean())
    out["ratio_bf16x3_over_fft"] = round(out["bf16x3"]["value"] / out["fft_same_window"]["value"], 4)
    return out


def serving_leg(n_catalog: int, stream_requests: int, fixed_requests: int, fixed_rates,
                alone_requests: int = 20):
    """Serving form: a 3-channel 2048-sample catalog resident in HBM; requests
    of 1..256 events with lengths U[512, 4096] at Poisson 200 req/s, top-10
    matches each; then 64-event requests one at a time (latency of a request
    alone on the GPU) and at a sweep of Poisson rates (latency against load).
    A leg with 0 requests is skipped."""
    from seismic_amd import serving, synth
    log(f"bench: serving catalog {n_catalog} x 3 x 2048")
    cat, _ = synth.make_events(n_catalog, 3, 2048, seed=777)
    t0 = time.perf_counter()
    srv = serving.Server(cat, max_query_len=4096, top_k=10, max_batch_events=256, n_slots=4)
    t_load = time.perf_counter() - t0
    del cat
    # warm-up: one request per period bucket
    warm = serving.make_requests(4, 3, (512, 4096), (1, 64), seed=5, pool_events=64)
    for i in [srv.submit(e, ln) for e, ln in warm]:
        srv.wait(i)
    out = {"catalog_events": n_catalog, "channels": 3, "catalog_samples": 2048, "top_k": 10,
           "slots": srv.stats()["slots"], "slot_bytes": srv.stats()["slot_bytes"],
           "catalog_load_s": round(t_load, 3)}

    def leg(reqs, rate, seed):
        srv.reset_stats()
        r = serving.run_poisson(srv, reqs, rate, seed)
        st = srv.stats()
        lat = r["latency_ms"]
        ev = int(r["sizes"].sum())
        return {"requests": len(reqs), "events": ev, "offered_req_per_s": rate,
                "completed_req_per_s": round(len(reqs) / r["wall_s"], 2),
                "events_per_s": round(ev / r["wall_s"], 1),
                "pairs_per_s": round(ev * n_catalog / r["wall_s"], 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "max_ms": round(float(lat.max()), 2),
                "mean_requests_per_batch": round(st["mean_requests_per_batch"], 2),
                "wall_s": round(r["wall_s"], 3)}

    if stream_requests > 0:
        reqs = serving.make_requests(stream_requests, 3, (512, 4096), (1, 256), seed=11)
        log(f"bench: serving stream, {stream_requests} requests at 200 req/s")
        out["stream_1_256_events"] = leg(reqs, 200.0, 12)
    if alone_requests > 0:
        reqs = serving.make_requests(alone_requests, 3, (512, 4096), (64, 64), seed=13)
        log(f"bench: serving 64-event requests one at a time, {alone_requests}")
        lat = []
        for e, ln in reqs:
            t0 = time.perf_counter()
            srv.wait(srv.submit(e, ln))
            lat.append((time.perf_counter() - t0) * 1e3)
        lat = np.array(lat)
        out["requests_64_events_alone"] = {"requests": alone_requests,
                                           "p50_ms": round(float(np.percentile(lat, 50)), 2),
                                           "max_ms": round(float(lat.max()), 2)}
    if fixed_requests > 0:
        reqs = serving.make_requests(fixed_requests, 3, (512, 4096), (64, 64), seed=13)
        sweep = []
        for rate in fixed_rates:
            log(f"bench: serving 64-event requests, {fixed_requests} at {rate} req/s")
            sweep.append(leg(reqs, rate, 14))
        out["requests_64_events"] = sweep
        ok = [r["offered_req_per_s"] for r in sweep if r["p99_ms"] < 50.0]
        out["p99_below_50ms_up_to_req_per_s"] = max(ok) if ok else None
    srv.close()
    return out


def tiled_leg(n_events: int, block: int):
    """Tiled N x N form at the largest-run trace shape: 1 channel, L=4096,
    window +-1024, bf16x3, log-uniform SNR [0.3, 30], 10% shifted duplicates;
    float16 coefficient + int16 lag blocks streamed to host memory."""
    from seismic_amd import synth, tiled
    from seismic_amd.device import DeviceArray
    L, w = 4096, 1024
    log(f"bench: tiled N={n_events} L={L} +-{w} block {block}")
    ev, _ = synth.make_events(n_events, 1, L, seed=65536, snr=(0.3, 30.0), log_snr=True, dup_frac=0.1)
    tr = DeviceArray.from_numpy(ev)
    del ev
    run = tiled.TiledRun(n_events, 1, L, -w, w, block=block)
    e = run.edge
    bufs = [np.empty((e, e), np.float16), np.empty((e, e), np.int16)]
    t0 = time.perf_counter()
    run.start(tr)
    nb = 0
    t_first = None
    while run.next_block(bufs) is not None:
        nb += 1
        if t_first is None:
            t_first = time.perf_counter() - t0
    wall = time.perf_counter() - t0
    st = run.stats()
    run.close()
    tr.free()
    pairs = n_events * (n_events + 1) // 2
    rate = pairs / wall
    issued = 3 * 2 * (2 * w + 1) * L  # bf16x3 MFMA flops per pair
    n5 = 65536
    return {"n_events": n_events, "samples": L, "lags": [-w, w], "form": "bf16x3", "block": block,
            "blocks": nb, "pairs": pairs, "wall_s": round(wall, 3),
            "first_block_s": round(t_first or 0.0, 3), "pairs_per_s": round(rate, 1),
            "mfma_tflops_issued": round(rate * issued / 1e12, 1),
            "output_bytes_to_host": pairs * 4, "near_ties": st["near_ties"],
            "tie_overflow": st["tie_overflow"],
            "projected_s_n65536": round(n5 * (n5 + 1) / 2 / rate, 1)}


def cpu_baseline(ev: np.ndarray, P: int, S: int, sample_events: int, threads: int):
    from oracle import oracle
    n = min(sample_events, ev.shape[0])
    C, L = ev.shape[1], ev.shape[2]
    padded = np.zeros((n, C, P), np.float32)
    padded[:, :, :L] = ev[:n]
    oracle.load()
    t0 = time.perf_counter()
    oracle.allpairs_stacked(padded, L, S, num_threads=threads, chunk_size=4)
    dt = time.perf_counter() - t0
    pairs = n * (n - 1) / 2
    return {"value": pairs / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"all {n * (n + 1) // 2} packed pairs of the first {n} events "
                      f"(C={C}, L={L}, P={P}, full lags), {dt:.2f} s wall, C/OpenMP "
                      f"restatement of Correlation_in_OMP with channel stacking, own FP32 FFT"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-events", type=int, default=4096)
    ap.add_argument("--cpu-sample-events", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lag-forms", action="store_true", help="skip the MFMA lag-window forms")
    ap.add_argument("--lag-steps", type=int, default=2)
    ap.add_argument("--fp64-pairs", type=int, default=2000,
                    help="seeded pairs checked against the FP64 numpy reference, every form")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-tiled", action="store_true", help="skip the tiled N x N leg")
    ap.add_argument("--tiled-n", type=int, default=16384,
                    help="events of the tiled leg (65536 = the full largest run)")
    ap.add_argument("--tiled-block", type=int, default=8192)
    ap.add_argument("--serve-catalog", type=int, default=32768)
    ap.add_argument("--serve-stream-requests", type=int, default=64)
    ap.add_argument("--serve-64-requests", type=int, default=200)
    ap.add_argument("--serve-alone-requests", type=int, default=20,
                    help="64-event requests submitted one at a time (latency alone on the GPU)")
    ap.add_argument("--serve-64-rates", type=str, default="5,10,20,30",
                    help="comma-separated Poisson rates (req/s) of the 64-event requests")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (scales PMC byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} (WORLD_SIZE={world}) is not "
              "supported: the all-pairs job runs on a single GPU", file=sys.stderr)
        sys.exit(2)

    from seismic_amd import engine, synth
    from seismic_amd.device import DeviceArray, Stream, synchronize

    N, C, L, P = args.n_events, 3, 2048, 4096
    lag_lo, lag_hi = -P // 2 + 1, P // 2
    log(f"bench: generating N={N} C={C} L={L}")
    ev, _ = synth.make_events(N, C, L, seed=20240601)
    tr = DeviceArray.from_numpy(ev)
    plan = engine.AllPairsPlan(N, C, L, P, lag_lo, lag_hi, threshold=0.0)
    outs = plan.alloc_outputs()
    stream = Stream()

    if args.copy_calibration:
        from seismic_amd._lib import load_engine
        lib = load_engine()
        a = DeviceArray(64 * 1024 * 1024, np.float32).zero_()
        b = DeviceArray(64 * 1024 * 1024, np.float32)
        import ctypes
        for _ in range(3):
            lib.sxc_copy_kernel(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), a.nbytes, None)
        synchronize()
        log("bench: copy calibration 3 x 256 MiB done")

    ms, tm, st = time_plan(plan, tr, outs, stream, args.steps, args.warmup)
    unique = N * (N - 1) / 2
    packed = N * (N + 1) / 2
    value = unique / (ms / 1e3)
    log(f"bench: {ms:.3f} ms/step, pair kernel {tm['ms_pairs']:.3f} ms, forward "
        f"{tm['ms_forward']:.3f} ms, resolve {tm['ms_resolve']:.3f} ms")

    fpp = fft_flops_per_pair(C, P)
    achieved = packed * fpp / (tm["ms_pairs"] / 1e3) / 1e12
    traffic = None
    prof = os.path.join(HERE, "profiles", "pairs_fft_n4096_c3_p4096_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic, "kernel": "k_pairs<32,full>",
                "compute": "FP32 vector ALU (no MFMA); peak is MI355X FP32 157.3 TFLOP/s",
                "flops_per_pair": fpp, "pairs_per_launch": int(packed),
                "kernel_ms": round(tm["ms_pairs"], 4)}
    if traffic:
        # profiled HBM bytes of one launch over this run's kernel time
        gbs = traffic / (tm["ms_pairs"] / 1e3) / 1e9
        roofline["hbm_gb_per_s"] = round(gbs, 1)
        roofline["hbm_frac"] = round(gbs / PEAK_HBM_GBS, 4)

    # accuracy: every form against an FP64 numpy reference on seeded pairs
    log(f"bench: FP64 reference on {args.fp64_pairs} sampled pairs")
    w = LAG_WINDOW
    idx64, ref64 = fp64_reference(ev, P, args.fp64_pairs, 64, [(lag_lo, lag_hi), (-w, w)])
    accuracy = {"reference": "numpy FP64, channel-averaged normalized correlation, seeded packed pairs",
                "fft_f32_full_lags": vs_fp64([o.numpy() for o in outs], idx64, ref64[(lag_lo, lag_hi)])}

    lag = None
    if not args.no_lag_forms:
        log(f"bench: lag-window forms at +-{LAG_WINDOW}")
        lag = lag_forms(ev, tr, stream, args.lag_steps, 1, (idx64, ref64[(-w, w)]))
        for name in ("bf16x3", "fp16", "fft_same_window"):
            accuracy[f"{name}_w{w}"] = lag[name]["vs_fp64"]
        log(f"bench: bf16x3 {lag['bf16x3']['ms_per_step']} ms, fp16 {lag['fp16']['ms_per_step']} ms, "
            f"fft {lag['fft_same_window']['ms_per_step']} ms")

    serve = None
    if not args.no_serving:
        serve = serving_leg(args.serve_catalog, args.serve_stream_requests, args.serve_64_requests,
                            [float(r) for r in args.serve_64_rates.split(",") if r],
                            args.serve_alone_requests)
        log(f"bench: serving {json.dumps(serve)}")

    tiles = None
    if not args.no_tiled:
        tiles = tiled_leg(args.tiled_n, args.tiled_block)
        log(f"bench: tiled {json.dumps(tiles)}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        log(f"bench: CPU baseline on {threads} threads, {args.cpu_sample_events} events")
        cpu = cpu_baseline(ev, P, P // 2, args.cpu_sample_events, threads)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded band-passed noise + 5 Hz Ricker, SNR U[0.5,10])",
        "config": {"workload": "all-pairs NCC, FP32-FFT form", "form": "fft_f32",
                   "n_events": N, "channels": C, "samples": L, "padded_size": P,
                   "lags": [lag_lo, lag_hi], "unique_pairs": int(unique),
                   "packed_pairs_computed": int(packed)},
        "stages_ms": {k: round(v, 4) for k, v in tm.items() if k != "executes"},
        "near_ties_per_step": st["near_ties"],
        "accuracy_vs_fp64": accuracy,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mfma_bf16x3": lag,
        "serving": serve,
        "tiled_nxn": tiles,
    }
    print(json.dumps(line), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
