"""Serving loop: argument checks on the CPU, results against an FP64 numpy
oracle, batching independence and admission on the GPU."""
import numpy as np
import pytest

from seismic_amd import serving, synth


def ncc_f64(q, x):
    """Channel-averaged linear NCC of query q [C, Lq] against catalog event
    x [C, Lc] for lags -(Lc-1) .. Lq-1 (lag > 0: query delayed), FP64."""
    C, Lq = q.shape
    Lc = x.shape[1]
    n = 1
    while n < Lq + Lc - 1:
        n *= 2
    acc = np.zeros(Lq + Lc - 1)
    for c in range(C):
        a = q[c].astype(np.float64)
        b = x[c].astype(np.float64)
        na, nb = np.linalg.norm(a), np.linalg.norm(b)
        if na == 0 or nb == 0:
            continue
        # r[l] = sum_m a[m] b[m - l]; index l + (Lc - 1)
        r = np.fft.irfft(np.fft.rfft(a, n) * np.conj(np.fft.rfft(b, n)), n)
        full = np.concatenate([r[n - (Lc - 1):], r[:Lq]])
        acc += full / (na * nb)
    return acc / C


def oracle_topk(queries, lens, catalog, k):
    out = []
    Lc = catalog.shape[2]
    for qi in range(len(queries)):
        q = queries[qi][:, :lens[qi]]
        best = []
        for ci in range(catalog.shape[0]):
            r = ncc_f64(q, catalog[ci])
            m = int(np.argmax(r))
            v = r[m]
            srt = np.sort(r)
            gap = srt[-1] - srt[-2]
            best.append((v if v > 0 else 0.0, (m - (Lc - 1)) if v > 0 else 0, ci, gap))
        best.sort(key=lambda t: (-t[0], t[2]))
        out.append(best[:k])
    return out


def test_invalid_server_desc_rejected():
    cat = np.zeros((4, 1, 100), np.float32)
    with pytest.raises(RuntimeError, match="invalid server description"):
        serving.Server(cat, max_query_len=100, top_k=40)
    with pytest.raises(RuntimeError, match="memory budget"):
        serving.Server(cat, max_query_len=100, top_k=2, mem_budget=1000)
    # P >= max_query_len + catalog_len - 1: no supported period above 8192
    with pytest.raises(RuntimeError, match="FFT period above 8192"):
        serving.Server(np.zeros((4, 1, 5000), np.float32), max_query_len=3194, top_k=2)


@pytest.mark.gpu
def test_submit_checks_shapes():
    """The wrapper rejects events whose channel count or length list does not
    match (the C ABI cannot see buffer sizes) and unknown request ids; one
    channel may come as [n, stride]."""
    print("case serving submit shape checks", flush=True)
    cat, _ = synth.make_events(20, 2, 256, seed=31)
    srv = serving.Server(cat, max_query_len=300, top_k=2, max_batch_events=8, n_slots=1)
    try:
        q, _ = synth.make_events(3, 3, 300, seed=32)
        with pytest.raises(ValueError, match="events must be"):
            srv.submit(q, [300, 300, 300])
        with pytest.raises(ValueError, match="lengths for"):
            srv.submit(q[:, :2], [300, 300])
        with pytest.raises(RuntimeError, match="unknown"):
            srv.wait(12345)
        v, l, idx, _ = srv.wait(srv.submit(q[:, :2], [300, 200, 100]))
        assert v.shape == (3, 2) and (idx >= 0).all()
    finally:
        srv.close()
    cat1, _ = synth.make_events(20, 1, 256, seed=33)
    srv = serving.Server(cat1, max_query_len=300, top_k=2, max_batch_events=8, n_slots=1)
    try:
        q1 = np.zeros((2, 300), np.float32)
        q1[0, 10:266] = cat1[5, 0]
        q1[1, :256] = cat1[7, 0]
        v, l, idx, _ = srv.wait(srv.submit(q1, [300, 256]))
        assert idx[0, 0] == 5 and l[0, 0] == 10 and idx[1, 0] == 7 and l[1, 0] == 0
    finally:
        srv.close()


@pytest.mark.gpu
def test_serving_matches_f64():
    cat, _ = synth.make_events(150, 3, 512, seed=21)
    srv = serving.Server(cat, max_query_len=900, top_k=5, max_batch_events=32, n_slots=2)
    rng = np.random.default_rng(22)
    pool, _ = synth.make_events(12, 3, 900, seed=23)
    lens = rng.integers(100, 901, size=12).astype(np.int32)
    lens[0], lens[1] = 513, 514     # both sides of the P=1024 / P=2048 boundary
    q = pool.copy()
    for e in range(12):
        q[e, :, lens[e]:] = 0
    # a query that is an exact (shifted) copy of a catalog event
    q[2] = 0
    q[2, :, 30:30 + 512] = cat[77]
    lens[2] = 600
    rid = srv.submit(q, lens)
    v, l, idx, lat = srv.wait(rid)
    ref = oracle_topk(q, lens, cat, 5)
    worst = 0.0
    for e in range(12):
        rv = np.array([t[0] for t in ref[e]])
        worst = max(worst, np.abs(v[e] - rv).max())
        for r in range(5):
            sep = (r == 0 or rv[r - 1] - rv[r] > 1e-5) and (r == 4 or rv[r] - rv[r + 1] > 1e-5)
            if sep:
                assert idx[e, r] == ref[e][r][2], (e, r)
                if ref[e][r][3] > 1e-5:
                    assert l[e, r] == ref[e][r][1], (e, r)
    print(f"serving: max |dcoef| {worst:.3e}, latency {lat:.1f} ms")
    assert worst <= 1e-5
    assert idx[2, 0] == 77 and l[2, 0] == 30 and abs(v[2, 0] - 1) < 1e-5
    srv.close()


@pytest.mark.gpu
def test_serving_periods_and_wrapped_lags():
    """Query lengths on both sides of every period boundary of a 1200-sample
    catalog (P = 2048, 3072, 4096, 5120, 6144, 8192: the smallest P >= Lq +
    Lc - 1), and exact shifted catalog copies at lags beyond P/2 (reported by
    the kernel as lag - P and moved back), at or below -P/2 (reported as
    lag + P and moved back) and at negative lags."""
    cat, _ = synth.make_events(40, 2, 1200, seed=71)
    # event 8 carries its energy in its last 200 samples, so a query holding
    # only samples 1024.. of it still matches it first
    cat[8, :, :1000] *= 0.05
    srv = serving.Server(cat, max_query_len=6000, top_k=4, max_batch_events=24, n_slots=2)
    lens = np.array([1, 1000, 1873, 1874, 2897, 2898, 4945, 4946, 6000, 4900, 5500, 1800,
                     3921, 3922, 3900, 849, 850, 849], np.int32)
    q, _ = synth.make_events(len(lens), 2, 6000, seed=72)
    for e in range(len(lens)):
        q[e, :, lens[e]:] = 0
    # query -> (catalog event, lag): wrapped lags at P = 6144, 8192 and 5120
    # (above P/2) and at P = 2048 (-1024 = -P/2, reported by the kernel as +1024)
    copies = {9: (5, 3700), 10: (17, 4200), 11: (23, -300), 14: (31, 2650), 17: (8, -1024)}
    for e, (c, lag) in copies.items():
        q[e] = 0
        if lag >= 0:
            q[e, :, lag:lag + 1200] = cat[c]
        else:
            q[e, :, :1200 + lag] = cat[c][:, -lag:]
    rid = srv.submit(q, lens)
    v, l, idx, lat = srv.wait(rid)
    ref = oracle_topk(q, lens, cat, 4)
    worst = 0.0
    for e in range(len(lens)):
        rv = np.array([t[0] for t in ref[e]])
        worst = max(worst, np.abs(v[e] - rv).max())
        for r in range(4):
            sep = (r == 0 or rv[r - 1] - rv[r] > 1e-5) and (r == 3 or rv[r] - rv[r + 1] > 1e-5)
            if sep:
                assert idx[e, r] == ref[e][r][2], (e, r)
                if ref[e][r][3] > 1e-5:
                    assert l[e, r] == ref[e][r][1], (e, r, l[e, r], ref[e][r][1])
    print(f"serving periods: max |dcoef| {worst:.3e}, latency {lat:.1f} ms")
    assert worst <= 1e-5
    for e, (c, lag) in copies.items():
        assert idx[e, 0] == c and l[e, 0] == lag, (e, idx[e, 0], l[e, 0])
        # a partial copy (negative lag) overlaps samples -lag.. of the event only
        x = cat[c].astype(np.float64)
        want = 1.0 if lag >= 0 else np.mean(np.linalg.norm(x[:, -lag:], axis=1) / np.linalg.norm(x, axis=1))
        assert abs(v[e, 0] - want) < 1e-5, (e, v[e, 0], want)
    srv.close()


@pytest.mark.gpu
def test_serving_catalog_longer_than_queries():
    """Catalog events of 3000 samples, queries of 1..1000 (P = 3072 or 4096
    < 2 L_c): negative lags down to -(L_c - 1) come out right, -2500 at
    P = 4096 through the fix-up of a lag at or below -P/2 (exact partial
    copies of catalog events at lags -2500 and -10)."""
    cat, _ = synth.make_events(30, 1, 3000, seed=81)
    srv = serving.Server(cat, max_query_len=1000, top_k=3, max_batch_events=8, n_slots=1)
    lens = np.array([1000, 1000, 700, 1, 400], np.int32)
    q, _ = synth.make_events(len(lens), 1, 1000, seed=82)
    for e in range(len(lens)):
        q[e, :, lens[e]:] = 0
    copies = {0: (4, -2500), 1: (11, -10)}      # query -> (catalog event, lag)
    for e, (c, lag) in copies.items():
        q[e] = 0
        n = min(lens[e], 3000 + lag)
        q[e, :, :n] = cat[c][:, -lag:-lag + n]
    v, l, idx, lat = srv.wait(srv.submit(q, lens))
    ref = oracle_topk(q, lens, cat, 3)
    worst = 0.0
    for e in range(len(lens)):
        rv = np.array([t[0] for t in ref[e]])
        worst = max(worst, np.abs(v[e] - rv).max())
        if rv[0] - rv[1] > 1e-5:
            assert idx[e, 0] == ref[e][0][2]
            if ref[e][0][3] > 1e-5:
                assert l[e, 0] == ref[e][0][1], (e, l[e, 0], ref[e][0][1])
    print(f"serving long catalog: max |dcoef| {worst:.3e}")
    assert worst <= 1e-5
    for e, (c, lag) in copies.items():
        assert idx[e, 0] == c and l[e, 0] == lag, (e, idx[e, 0], l[e, 0])
    srv.close()


@pytest.mark.gpu
def test_result_same_alone_and_in_mix():
    cat, _ = synth.make_events(300, 3, 1024, seed=31)
    srv = serving.Server(cat, max_query_len=2000, top_k=10, max_batch_events=64, n_slots=3)
    reqs = serving.make_requests(24, 3, (200, 2000), (1, 40), seed=32, pool_events=64)
    ev, lens = reqs[7]
    alone = srv.wait(srv.submit(ev, lens))
    ids = [srv.submit(e, ln) for e, ln in reqs]          # all in flight together
    mixed = [srv.wait(i) for i in ids]
    st = srv.stats()
    for a, b in zip(alone[:3], mixed[7][:3]):
        assert np.array_equal(a, b)
    assert st["batches"] < 25          # some requests were batched together
    print("requests per batch", st["mean_requests_per_batch"], "slots", st["slots"])
    srv.close()


@pytest.mark.gpu
def test_memory_budget_limits_slots():
    cat, _ = synth.make_events(64, 1, 256, seed=41)
    one = serving.Server(cat, max_query_len=256, top_k=3, max_batch_events=8, n_slots=1)
    per_slot = one.stats()["slot_bytes"]
    one.close()
    assert per_slot > 0
    srv = serving.Server(cat, max_query_len=256, top_k=3, max_batch_events=8, n_slots=8,
                         mem_budget=int(2.5 * per_slot))
    assert srv.stats()["slots"] == 2
    reqs = serving.make_requests(10, 1, (50, 256), (1, 8), seed=42, pool_events=16)
    ids = [srv.submit(e, ln) for e, ln in reqs]
    for i in ids:
        srv.wait(i)
    assert srv.stats()["requests"] == 10
    with pytest.raises(RuntimeError, match="events"):
        srv.submit(np.zeros((9, 1, 256), np.float32), np.full(9, 100, np.int32))
    with pytest.raises(RuntimeError, match="length"):
        srv.submit(np.zeros((1, 1, 300), np.float32), np.array([300], np.int32))
    srv.close()


@pytest.mark.gpu
def test_concurrent_submitters():
    """Several host threads submit and wait at once; every result equals the
    same request served alone."""
    import threading
    cat, _ = synth.make_events(200, 2, 800, seed=51)
    srv = serving.Server(cat, max_query_len=1500, top_k=4, max_batch_events=48, n_slots=3)
    reqs = serving.make_requests(16, 2, (100, 1500), (1, 24), seed=52, pool_events=32)
    alone = [srv.wait(srv.submit(e, ln))[:3] for e, ln in reqs]
    got = [None] * len(reqs)

    def client(k0):
        for k in range(k0, len(reqs), 4):
            e, ln = reqs[k]
            got[k] = srv.wait(srv.submit(e, ln))[:3]
    th = [threading.Thread(target=client, args=(c,)) for c in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for a, b in zip(alone, got):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
    srv.close()


@pytest.mark.gpu
def test_serving_one_channel_demean_top32_length_extremes():
    """1 channel, FP64 mean removal, top_k = 32 (the maximum) and query
    lengths 1, catalog_len and max_query_len, against the FP64 oracle."""
    cat, _ = synth.make_events(60, 1, 300, seed=61)
    cat = cat + np.float32(2.0)                    # offsets the device must remove
    srv = serving.Server(cat, max_query_len=700, top_k=32, max_batch_events=8, n_slots=2, demean=True)
    q, _ = synth.make_events(3, 1, 700, seed=62)
    q = q + np.float32(-1.5)
    lens = np.array([1, 300, 700], np.int32)
    for e in range(3):
        q[e, :, lens[e]:] = 0
    v, l, idx, _ = srv.wait(srv.submit(q, lens))
    srv.close()
    cat0 = cat.astype(np.float64) - cat.astype(np.float64).mean(-1, keepdims=True)
    for e in range(3):
        x = q[e, :, :lens[e]].astype(np.float64)
        x = x - x.mean(-1, keepdims=True)
        ref = oracle_topk(x[None].astype(np.float32), [lens[e]], cat0.astype(np.float32), 32)[0]
        rv = np.array([t[0] for t in ref])
        assert np.abs(v[e] - rv).max() <= 1e-5, e
        if e == 0:   # a one-sample query has zero energy after mean removal: every match is 0
            assert np.all(v[e] == 0)


def test_latency_budget_argument_checks():
    """The admission setter refuses a null server and a negative budget
    without touching a device."""
    import ctypes
    from seismic_amd import _lib
    lib = _lib.load_engine()
    assert lib.sxc_server_set_latency_budget(None, ctypes.c_double(10.0)) == -1
    assert b"null" in lib.seismic_last_error()


@pytest.mark.gpu
def test_deadline_admission_sheds_and_keeps_results():
    """A burst submitted at once against a small latency budget: the server
    sheds what its work model predicts would finish late (Overloaded, not
    queued), serves the rest, and every admitted request's results are
    bitwise those of the same request served alone.  A request larger than
    the budget is still admitted onto an idle device, and budget 0 admits
    everything."""
    cat, _ = synth.make_events(2048, 3, 1024, seed=61)
    srv = serving.Server(cat, max_query_len=2000, top_k=5, max_batch_events=64, n_slots=3)
    st0 = srv.stats()
    print(f"model: {st0['model_tflops']:.1f} TFLOP/s, {st0['model_overhead_ms']:.2f} ms per request")
    assert st0["model_tflops"] > 0
    reqs = serving.make_requests(40, 3, (300, 2000), (16, 64), seed=62, pool_events=64)
    alone = [srv.wait(srv.submit(e, ln))[:3] for e, ln in reqs]
    srv.reset_stats()
    srv.set_latency_budget(4.0)
    ids = []
    for e, ln in reqs:
        try:
            ids.append(srv.submit(e, ln))
        except serving.Overloaded as ex:
            assert "budget" in str(ex)
            ids.append(None)
    got = [srv.wait(i)[:3] if i is not None else None for i in ids]
    st = srv.stats()
    shed = sum(i is None for i in ids)
    print(f"burst of 40: {shed} shed, {40 - shed} served; stats offered {st['offered']} rejected {st['rejected']}")
    assert st["offered"] == 40 and st["rejected"] == shed and st["requests"] == 40 - shed
    assert 0 < shed < 40
    for k, g in enumerate(got):
        if g is not None:
            for a, b in zip(alone[k], g):
                assert np.array_equal(a, b)
    # idle device: a request predicted above the budget is still served
    srv.set_latency_budget(0.01)
    big = reqs[0]
    r = srv.wait(srv.submit(*big))
    for a, b in zip(alone[0], r[:3]):
        assert np.array_equal(a, b)
    # admission off: the whole burst is queued
    srv.set_latency_budget(0.0)
    srv.reset_stats()
    ids = [srv.submit(e, ln) for e, ln in reqs]
    for i in ids:
        srv.wait(i)
    assert srv.stats()["rejected"] == 0 and srv.stats()["requests"] == 40
    with pytest.raises(RuntimeError, match="budget_ms"):
        srv.set_latency_budget(-1.0)
    srv.close()
