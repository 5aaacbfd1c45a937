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


# Sources whose compiled code is the FP32-FFT pair kernel: a traffic profile
# is valid for the build it was measured on only.
PAIR_KERNEL_SOURCES = ("seismic_amd/csrc/xc_kernels.hip", "seismic_amd/csrc/xc_fft.hpp",
                       "seismic_amd/csrc/xc_kernels.h")


def pair_kernel_source_sha256() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in PAIR_KERNEL_SOURCES:
        with open(os.path.join(HERE, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def profiled_traffic():
    """HBM bytes per launch of the pair kernel from the committed counter
    profile, or (None, reason) when that profile was taken on other sources."""
    prof = os.path.join(HERE, "profiles", "pairs_fft_n4096_c3_p4096_traffic.json")
    if not os.path.exists(prof):
        return None, "no traffic profile"
    with open(prof) as f:
        t = json.load(f)
    if t.get("kernel_source_sha256") != pair_kernel_source_sha256():
        return None, "traffic profile was taken on other kernel sources"
    return t, None


def host_cpu():
    """CPU model, logical CPUs of the host and the CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_allowed": allowed,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


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


def serving_leg(n_catalog: int, stream_requests: int, fixed_requests: int, fixed_rates, budget_ms: float,
                alone_requests: int = 20):
    """Serving form: a 3-channel 2048-sample catalog resident in HBM; requests
    of 1..256 events with lengths U[512, 4096] at Poisson 200 req/s, top-10
    matches each; then 64-event requests one at a time (latency of a request
    alone on the GPU) and at a sweep of Poisson rates (latency against load).
    Deadline admission sheds requests predicted to finish later than
    budget_ms after arrival (0: off); every leg reports offered, admitted,
    rejected and completed rates, and latency percentiles of the admitted
    requests.  A leg with 0 requests is skipped."""
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
    st0 = srv.stats()
    out = {"catalog_events": n_catalog, "channels": 3, "catalog_samples": 2048, "top_k": 10,
           "slots": st0["slots"], "slot_bytes": st0["slot_bytes"], "catalog_load_s": round(t_load, 3),
           "latency_budget_ms": budget_ms,
           "admission": "deadline: a request predicted (modelled device work already admitted + its own "
                        "+ fixed cost) to finish more than the budget after arrival is shed at submit, "
                        "unless the device is idle",
           "model_tflops_at_create": round(st0["model_tflops"], 2),
           "model_overhead_ms_at_create": round(st0["model_overhead_ms"], 3)}

    def leg(reqs, rate, seed, budget):
        srv.set_latency_budget(budget)
        srv.reset_stats()
        r = serving.run_poisson(srv, reqs, rate, seed)
        st = srv.stats()
        ok = r["admitted"]
        lat = r["latency_ms"][ok]
        ev = int(r["sizes"][ok].sum())
        small = ok & (r["sizes"] <= 64)
        res = {"requests": len(reqs), "offered_req_per_s": rate,
               "offered_req_per_s_measured": round(len(reqs) / max(r["offered_s"], 1e-9), 2),
               "admitted": int(ok.sum()), "rejected": int((~ok).sum()),
               "rejected_frac": round(float((~ok).mean()), 4),
               "admitted_req_per_s": round(int(ok.sum()) / r["wall_s"], 2),
               "completed_req_per_s": round(int(st["requests"]) / r["wall_s"], 2),
               "events_per_s": round(ev / r["wall_s"], 1),
               "pairs_per_s": round(ev * n_catalog / r["wall_s"], 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 2) if lat.size else None,
               "p99_ms": round(float(np.percentile(lat, 99)), 2) if lat.size else None,
               "max_ms": round(float(lat.max()), 2) if lat.size else None,
               "mean_requests_per_batch": round(st["mean_requests_per_batch"], 2),
               "budget_ms": budget, "wall_s": round(r["wall_s"], 3),
               "model_tflops": round(st["model_tflops"], 2)}
        if small.any() and small.sum() != ok.sum():
            res["p99_ms_requests_le_64_events"] = round(float(np.percentile(r["latency_ms"][small], 99)), 2)
            by = {}
            for lo, hi in ((1, 64), (65, 128), (129, 192), (193, 256)):
                m = (r["sizes"] >= lo) & (r["sizes"] <= hi)
                lat_b = r["latency_ms"][m & ok]
                by[f"{lo}-{hi}"] = {"offered": int(m.sum()), "admitted": int((m & ok).sum()),
                                    "p50_ms": round(float(np.percentile(lat_b, 50)), 2) if lat_b.size else None,
                                    "p99_ms": round(float(np.percentile(lat_b, 99)), 2) if lat_b.size else None}
            res["by_request_events"] = by
        return res

    if stream_requests > 0:
        reqs = serving.make_requests(stream_requests, 3, (512, 4096), (1, 256), seed=11)
        log(f"bench: serving stream, {stream_requests} requests at 200 req/s, budget {budget_ms} ms")
        out["stream_1_256_events"] = leg(reqs, 200.0, 12, budget_ms)
        if budget_ms != 40.0:
            log(f"bench: serving stream, {stream_requests} requests at 200 req/s, budget 40 ms")
            out["stream_1_256_events_budget_40ms"] = leg(reqs, 200.0, 12, 40.0)
    if alone_requests > 0:
        reqs = serving.make_requests(alone_requests, 3, (512, 4096), (64, 64), seed=13)
        log(f"bench: serving 64-event requests one at a time, {alone_requests}")
        srv.set_latency_budget(0.0)
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
            log(f"bench: serving 64-event requests, {fixed_requests} at {rate} req/s, budget {budget_ms} ms")
            sweep.append(leg(reqs, rate, 14, budget_ms))
        out["requests_64_events"] = sweep
        ok = [r["offered_req_per_s"] for r in sweep if r["p99_ms"] is not None and r["p99_ms"] < 50.0]
        out["p99_below_50ms_up_to_req_per_s"] = max(ok) if ok else None
        top = max(fixed_rates)
        log(f"bench: serving 64-event requests at {top} req/s without admission")
        out["requests_64_events_no_admission"] = leg(reqs, top, 14, 0.0)
    srv.close()
    return out


def fp64_linear_pos(ev, pi, pj, lo, hi):
    """FP64 (numpy) channel-averaged linear NCC of the listed pairs over lags
    lo..hi: maximum above 0 with its first lag, and the top-two gap."""
    N, C, L = ev.shape
    P = 1 << int(2 * L - 1).bit_length()
    used = np.unique(np.concatenate([pi, pj]))
    slot = np.full(N, -1, np.int64)
    slot[used] = np.arange(used.size)
    x = ev[used].astype(np.float64)
    X = np.fft.rfft(x, P, axis=2) / np.linalg.norm(x, axis=2)[:, :, None]
    lags = np.arange(lo, hi + 1)
    pos, lag, gap = [], [], []
    for s0 in range(0, len(pi), 256):
        a, b = slot[pi[s0:s0 + 256]], slot[pj[s0:s0 + 256]]
        r = np.fft.irfft((X[a] * np.conj(X[b])).sum(axis=1), P, axis=1)[:, lags % P] / C
        k = r.argmax(axis=1)
        v = r[np.arange(len(a)), k]
        srt = np.sort(r, axis=1)
        pos.append(np.where(v > 0, v, 0.0))
        lag.append(np.where(v > 0, lags[k], 0))
        gap.append(srt[:, -1] - srt[:, -2])
    return np.concatenate(pos), np.concatenate(lag), np.concatenate(gap)


def tiled_leg(n_events: int, block: int, fp64_pairs: int):
    """Tiled N x N form at the largest single-GPU configuration: 1 channel,
    L=4096, window +-1024, bf16x3, log-uniform SNR [0.3, 30], 10% zero-filled
    time-shifted duplicates; float16 coefficient + int16 lag blocks streamed
    to host memory.  Each delivered block is checked as it arrives: every
    duplicate pair must peak at exactly its shift, and seeded pairs are
    compared with an FP64 numpy reference."""
    from seismic_amd import synth, tiled
    from seismic_amd.device import DeviceArray
    L, w = 4096, 1024
    log(f"bench: tiled N={n_events} L={L} +-{w} block {block}")
    ev, meta = synth.make_events(n_events, 1, L, seed=65536, snr=(0.3, 30.0), log_snr=True, dup_frac=0.1)
    rng = np.random.Generator(np.random.PCG64(5))
    a = rng.integers(0, n_events, size=fp64_pairs)
    b = rng.integers(0, n_events, size=fp64_pairs)
    pi, pj = np.minimum(a, b), np.maximum(a, b)
    ref_v, ref_l, ref_gap = fp64_linear_pos(ev, pi, pj, -w, w)
    dups = np.array(meta["dups"], dtype=np.int64).reshape(-1, 3)  # (src, dst, shift), src < dst
    tr = DeviceArray.from_numpy(ev)
    del ev
    run = tiled.TiledRun(n_events, 1, L, -w, w, block=block)
    e = run.edge
    bufs = [np.empty((e, e), np.float16), np.empty((e, e), np.int16)]
    got_v = np.full(fp64_pairs, np.nan)
    got_l = np.zeros(fp64_pairs, np.int64)
    dup_seen = dup_exact = 0
    dup_coef_min = 1.0
    t0 = time.perf_counter()
    run.start(tr)
    nb = 0
    t_first = None
    while True:
        res = run.next_block(bufs)
        if res is None:
            break
        info, coef, lagb = res
        nb += 1
        if t_first is None:
            t_first = time.perf_counter() - t0
        r0, c0, rows, cols = info["row0"], info["col0"], info["rows"], info["cols"]
        m = (pi >= r0) & (pi < r0 + rows) & (pj >= c0) & (pj < c0 + cols)
        got_v[m] = coef[pi[m] - r0, pj[m] - c0].astype(np.float64)
        got_l[m] = lagb[pi[m] - r0, pj[m] - c0]
        if dups.size:
            s_, d_, sh = dups[:, 0], dups[:, 1], dups[:, 2]
            md = (s_ >= r0) & (s_ < r0 + rows) & (d_ >= c0) & (d_ < c0 + cols)
            lg = lagb[s_[md] - r0, d_[md] - c0]
            dup_seen += int(md.sum())
            # x_dst[n] = x_src[n - shift]: the pair (src, dst) peaks at lag -shift
            dup_exact += int((lg == -sh[md]).sum())
            if md.any():
                dup_coef_min = min(dup_coef_min, float(coef[s_[md] - r0, d_[md] - c0].astype(np.float32).min()))
    wall = time.perf_counter() - t0
    st = run.stats()
    run.close()
    tr.free()
    pairs = n_events * (n_events + 1) // 2
    rate = pairs / wall
    issued = 3 * 2 * (2 * w + 1) * L  # bf16x3 MFMA flops per pair
    sure = ref_gap >= 1e-12
    live = ref_v > 1e-3
    out = {"n_events": n_events, "samples": L, "lags": [-w, w], "form": "bf16x3", "block": block,
           "blocks": nb, "pairs": pairs, "wall_s": round(wall, 3),
           "first_block_s": round(t_first or 0.0, 3), "pairs_per_s": round(rate, 1),
           "mfma_tflops_issued": round(rate * issued / 1e12, 1),
           "output_bytes_to_host": pairs * 4, "near_ties": st["near_ties"],
           "tie_overflow": st["tie_overflow"],
           "duplicates": {"pairs": len(dups), "checked": dup_seen, "lag_exact": dup_exact,
                          "min_coef": round(dup_coef_min, 4)},
           "vs_fp64": {"pairs": int(fp64_pairs), "checked": int(np.isfinite(got_v).sum()),
                       "max_abs_dcoef": float(np.nanmax(np.abs(got_v - ref_v))),
                       "dcoef_note": "float16 output: rounding up to 2^-11 relative",
                       "lag_agreement": float((got_l[live] == ref_l[live]).mean()),
                       "lag_agreement_fp64_gap_ge_1e-12": float((got_l[live & sure] == ref_l[live & sure]).mean())}}
    if n_events != 65536:
        n5 = 65536
        out["projected_s_n65536"] = round(n5 * (n5 + 1) / 2 / rate, 1)
    return out


def dropin_leg(N: int, L: int, calls: int):
    """The drop-in call as a reference caller sees it: correlationCUDA through
    ctypes (Correlation_in_CUDA/correlation_lib.py:86-97) on N events, 1
    channel, L samples, P = 1 << L.bit_length(), full lags (shift = L) --
    host-to-device copies of events and reversed events, compute, copies of
    the four packed outputs back (the device state of the shape is set up by
    the untimed first call) -- beside the device-resident plan of the same
    job (inputs already in HBM)."""
    import ctypes
    from seismic_amd import correlation_lib, engine, synth
    from seismic_amd.device import DeviceArray, Stream
    ev, _ = synth.make_events(N, 1, L, seed=4096)
    x = ev[:, 0]
    t0 = time.perf_counter()
    n, L_, P, padded, rev, outs = correlation_lib.prepare(x)
    t_prep = time.perf_counter() - t0
    lib = correlation_lib.initialise_library()
    lib.initialiseCUDA()
    fp = lambda a: a.ctypes.data_as(correlation_lib.c_float_p)
    ip = lambda a: a.ctypes.data_as(correlation_lib.c_int_p)
    args = (fp(padded), fp(rev), n, L_, L_, P, 1, ctypes.c_float(0.0), fp(outs[0]), ip(outs[1]),
            fp(outs[2]), ip(outs[3]))
    lib.correlationCUDA(*args)  # warm-up
    err = lib.seismic_last_error()
    if err:
        raise RuntimeError(err.decode())
    t0 = time.perf_counter()
    for _ in range(calls):
        lib.correlationCUDA(*args)
    ms_call = (time.perf_counter() - t0) * 1e3 / calls
    unique = N * (N - 1) / 2
    plan = engine.AllPairsPlan(N, 1, L_, P, -L_ + 1, L_)
    tr = DeviceArray.from_numpy(padded)
    rv = DeviceArray.from_numpy(rev)
    douts = plan.alloc_outputs()
    stream = Stream()
    plan.execute_reversed(tr, rv, out=douts, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        plan.execute_reversed(tr, rv, out=douts, stream=stream)
    stream.synchronize()
    ms_dev = (time.perf_counter() - t0) * 1e3 / calls
    plan.close()
    bytes_in = 2 * N * P * 4
    bytes_out = 4 * N * (N + 1) // 2 * 4
    return {"n_events": N, "channels": 1, "samples": L_, "padded_size": P, "shift": L_, "calls": calls,
            "ms_per_call": round(ms_call, 3), "pairs_per_s": round(unique / (ms_call / 1e3), 1),
            "device_resident_ms": round(ms_dev, 3),
            "device_resident_pairs_per_s": round(unique / (ms_dev / 1e3), 1),
            "host_to_device_bytes": bytes_in, "device_to_host_bytes": bytes_out,
            "binding_prep_s": round(t_prep, 3),
            "note": "ms_per_call is the whole correlationCUDA call after a first call of the same shape "
                    "(PCIe copies from and to the pageable numpy buffers through pinned staging, compute; "
                    "the plan and device buffers of the shape are kept from the first call); "
                    "binding_prep_s is the Python padding and flipping of compute_correlation, outside "
                    "the call"}


def cpu_baseline(ev: np.ndarray, P: int, S: int, sample_events: int, threads: int, host: dict):
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
                      f"restatement of Correlation_in_OMP with channel stacking, own FP32 FFT "
                      f"(FFTW3f is not installed), OpenMP dynamic schedule, chunk 4",
            "threads_rule": "every CPU this process may use (sched_getaffinity), capped by OMP_NUM_THREADS "
                            "when set (a host shared by several GPUs gives each its share)",
            **host}


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
    ap.add_argument("--tiled-n", type=int, default=65536,
                    help="events of the tiled leg (65536: the largest single-GPU configuration)")
    ap.add_argument("--tiled-fp64-pairs", type=int, default=2000)
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in end-to-end leg")
    ap.add_argument("--dropin-calls", type=int, default=3)
    ap.add_argument("--tiled-block", type=int, default=8192)
    ap.add_argument("--serve-catalog", type=int, default=32768)
    ap.add_argument("--serve-stream-requests", type=int, default=2000)
    ap.add_argument("--serve-64-requests", type=int, default=300)
    ap.add_argument("--serve-budget-ms", type=float, default=30.0,
                    help="latency budget of the deadline admission (0: admit everything)")
    ap.add_argument("--serve-alone-requests", type=int, default=20,
                    help="64-event requests submitted one at a time (latency alone on the GPU)")
    ap.add_argument("--serve-64-rates", type=str, default="10,20,30",
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
    tprof, tnote = profiled_traffic()
    traffic = tprof["hbm_bytes_per_launch"] if tprof else None
    roofline = {"bound": "fp32_valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic, "kernel": "k_pairs<32,full>",
                "compute": "FP32 vector ALU (no MFMA); peak is MI355X FP32 157.3 TFLOP/s",
                "flops_per_pair": fpp, "pairs_per_launch": int(packed),
                "kernel_ms": round(tm["ms_pairs"], 4),
                "kernel_source_sha256": pair_kernel_source_sha256()}
    if tprof:
        # profiled HBM bytes of one launch (same kernel sources) over this run's kernel time
        gbs = traffic / (tm["ms_pairs"] / 1e3) / 1e9
        roofline["hbm_gb_per_s"] = round(gbs, 1)
        roofline["hbm_frac"] = round(gbs / PEAK_HBM_GBS, 4)
        roofline["traffic_profile"] = {k: tprof.get(k) for k in ("kernel", "workload", "source", "profiled_ms",
                                                                 "fetch_scale", "write_scale")}
    else:
        roofline["traffic_note"] = tnote

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
                            [float(r) for r in args.serve_64_rates.split(",") if r], args.serve_budget_ms,
                            args.serve_alone_requests)
        log(f"bench: serving {json.dumps(serve)}")

    dropin = None
    if not args.no_dropin:
        log("bench: drop-in correlationCUDA end to end, N=4096 x 1 x 2048")
        dropin = dropin_leg(N, L, args.dropin_calls)
        log(f"bench: drop-in {json.dumps(dropin)}")

    tiles = None
    if not args.no_tiled:
        tiles = tiled_leg(args.tiled_n, args.tiled_block, args.tiled_fp64_pairs)
        log(f"bench: tiled {json.dumps(tiles)}")

    cpu = None
    if not args.no_cpu_baseline:
        host = host_cpu()
        threads = host["cpus_allowed"]
        if host["omp_num_threads_env"] and host["omp_num_threads_env"].isdigit():
            threads = min(threads, int(host["omp_num_threads_env"]))
        log(f"bench: CPU baseline on {threads} threads ({host['nproc']} logical CPUs, "
            f"{host['cpu_model']}), {args.cpu_sample_events} events")
        cpu = cpu_baseline(ev, P, P // 2, args.cpu_sample_events, threads, host)

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
        "dropin_end_to_end": dropin,
        "tiled_nxn": tiles,
    }
    print(json.dumps(line), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
