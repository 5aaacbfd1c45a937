"""GPU parity of the FP32-FFT form against the oracle, through the C ABI.

Bar (stated per test):
  coefficients: |gpu - fp64| <= 1e-5 absolute (FP64 = numpy FFT in float64,
                itself pinned to the direct FP64 oracle in test_oracle.py);
  lags: equal to the oracle's C/OMP restatement except where the FP64 gap
        between the top two values is below TIE_TOL, and equal to the FP64
        argmax wherever that gap is at least GPU_EXACT_GAP (the GPU settles
        near ties in FP64 itself, so only exact ties are exempt).
"""
import numpy as np
import pytest

from seismic_amd import correlation_lib, multicuda_lib, synth
from tests import np_ref

pytestmark = pytest.mark.gpu

COEF_TOL = 1e-5
TIE_TOL = 1e-5
GPU_EXACT_GAP = 1e-12


def _run_dropin(ev2d, shift, threshold=0.0):
    return correlation_lib.compute_correlation(ev2d, shift, 1, threshold)


def _check(gpu, ref, oracle_lags=None, name=""):
    vp, lp, vn, ln = gpu
    dp = np.abs(vp - ref["pos"]).max()
    dn = np.abs(vn - ref["neg"]).max()
    print(f"{name}: max |dcoef| pos {dp:.3e} neg {dn:.3e}; FP64 gaps < 1e-5: "
          f"{int((ref['gapp'] < 1e-5).sum())} pos, {int((ref['gapn'] < 1e-5).sum())} neg of {len(vp)}")
    assert dp <= COEF_TOL and dn <= COEF_TOL, f"{name}: max coef err pos {dp:.2e} neg {dn:.2e}"
    # GPU lags are the FP64 argmax unless the FP64 gap is tiny
    sure_p = ref["gapp"] >= GPU_EXACT_GAP
    sure_n = ref["gapn"] >= GPU_EXACT_GAP
    bad_p = np.nonzero(sure_p & (lp != ref["lagp"]))[0]
    bad_n = np.nonzero(sure_n & (ln != ref["lagn"]))[0]
    assert bad_p.size == 0, f"{name}: {bad_p.size} pos lags differ from FP64, first {bad_p[:5]}"
    assert bad_n.size == 0, f"{name}: {bad_n.size} neg lags differ from FP64, first {bad_n[:5]}"
    if oracle_lags is not None:
        olp, oln = oracle_lags
        mp = (lp != olp) & (ref["gapp"] >= TIE_TOL)
        mn = (ln != oln) & (ref["gapn"] >= TIE_TOL)
        assert not mp.any() and not mn.any(), f"{name}: lag mismatch vs C/OMP oracle outside ties"
    return dp, dn


def test_n256_l1024_dropin_full_window(oracle):
    """N=256, 1 channel, L=1024 (P=2048), full lag range (shift 1024)."""
    ev, _ = synth.make_events(256, 1, 1024, seed=101)
    x = ev[:, 0]
    print("case N=256 L=1024 shift=1024", flush=True)
    gpu = _run_dropin(x, 1024)
    padded, rev, L, P = synth.pad_like_binding(x)
    ref = np_ref.reference_f64(padded, rev, L, 1024)
    ovp, olp, ovn, oln, _, _ = oracle.correlation_cpp(padded, rev, L, 1024)
    dp, dn = _check(gpu, ref, (olp, oln), "N=256 L=1024")
    print(f"N=256 L=1024 max|dcoef| pos {dp:.2e} neg {dn:.2e}")


@pytest.mark.parametrize("L,shift,thr", [(1024, 256, 0.2), (1500, 548, 0.0), (700, 300, 0.1),
                                         (2048, 2048, 0.0), (2048, 100, 0.3), (4096, 4096, 0.0),
                                         (4096, 512, 0.25), (600, 40, 0.0)])
def test_dropin_sizes_windows(oracle, L, shift, thr):
    n = 40 if L <= 2048 else 20
    ev, _ = synth.make_events(n, 1, L, seed=L + shift)
    x = ev[:, 0]
    print(f"case sizes L={L} shift={shift} thr={thr}", flush=True)
    gpu = _run_dropin(x, shift, thr)
    padded, rev, L_, P = synth.pad_like_binding(x)
    ref = np_ref.reference_f64(padded, rev, L_, shift, threshold=thr)
    ovp, olp, ovn, oln, _, _ = oracle.correlation_cpp(padded, rev, L_, shift, threshold=thr)
    _check(gpu, ref, (olp, oln), f"L={L} shift={shift}")


def test_known_answers_gpu():
    rng = np.random.default_rng(5)
    L = 1024
    base = np.zeros(L, np.float32)
    base[200:800] = rng.standard_normal(600)
    ev = np.stack([base, np.roll(base, 37), -base, base, np.zeros(L, np.float32), np.roll(base, -120)])
    vp, lp, vn, ln = _run_dropin(ev, 1024, 0.0)
    n = len(ev)
    idx = lambda i, j: i * n - i * (i - 1) // 2 + (j - i)
    assert abs(vp[idx(0, 0)] - 1) < 1e-5 and lp[idx(0, 0)] == 0
    assert abs(vp[idx(0, 1)] - 1) < 1e-5 and lp[idx(0, 1)] == -37
    assert abs(vn[idx(0, 2)] + 1) < 1e-5 and ln[idx(0, 2)] == 0
    assert abs(vp[idx(0, 3)] - 1) < 1e-5 and lp[idx(0, 3)] == 0
    assert abs(vp[idx(0, 5)] - 1) < 1e-5 and lp[idx(0, 5)] == 120
    assert abs(vp[idx(1, 5)] - 1) < 1e-5 and lp[idx(1, 5)] == 157
    # zero-energy event: (0, 0) everywhere it takes part
    for j in range(n):
        k = idx(min(j, 4), max(j, 4))
        assert vp[k] == 0 and lp[k] == 0 and vn[k] == 0 and ln[k] == 0
    # threshold above every value
    vp2, lp2, vn2, ln2 = _run_dropin(ev, 1024, 1.5)
    assert not vp2.any() and not lp2.any() and not vn2.any() and not ln2.any()


def test_exact_ties_take_first_lag(oracle):
    """The autocorrelation of integer-valued noise is exactly symmetric in
    FP64 (every partial sum of the oracle's direct sums is an exact integer),
    so its minimum at -l and +l is an exact tie, and a period-64 cosine has
    near-equal maxima every 64 samples: the FP64 tie-break must return the
    smaller lag on the exact tie, as the reference scan's first occurrence,
    and the FP64 argmax otherwise."""
    L = 1024
    t = np.arange(L)
    cosine = np.cos(2 * np.pi * t / 64.0).astype(np.float32)
    rng = np.random.default_rng(9)
    noise = rng.integers(-8, 9, size=L).astype(np.float32)
    ev = np.stack([cosine, noise])
    vp, lp, vn, ln = _run_dropin(ev, 1024, 0.0)
    pi, pj = oracle.packed_pairs(2)
    exact = oracle.pairs_f64(ev, pi, pj, -1023, 1024, P=2048, circular=True)
    k = 2  # pair (1, 1)
    assert exact["gap_neg"][k] == 0.0, "integer autocorrelation: exact FP64 tie at -l and +l"
    assert exact["lag_neg"][k] < 0
    assert ln[k] == exact["lag_neg"][k]
    assert lp[0] == 0 and abs(vp[0] - 1) < 1e-5
    padded, rev, L_, P_ = synth.pad_like_binding(ev)
    ref = np_ref.reference_f64(padded, rev, L_, 1024)
    _check((vp, lp, vn, ln), ref, None, "cosine and integer noise")


def test_multicuda_matches_cuda():
    ev, _ = synth.make_events(64, 1, 2048, seed=7)
    x = ev[:, 0]
    a = _run_dropin(x, 512, 0.1)
    multicuda_lib.initialiseRunTime(4)
    b = multicuda_lib.compute_correlation(x, 512, 4, 0.1)
    for u, v in zip(a, b[:4]):
        assert np.array_equal(u, v)


def test_execute_rejects_small_outputs():
    """Caller-provided output buffers shorter than N(N+1)/2 entries, or of the
    wrong type, are refused before any launch."""
    from seismic_amd import engine
    from seismic_amd.device import DeviceArray
    print("case execute output checks", flush=True)
    ev, _ = synth.make_events(40, 1, 500, seed=9)
    plan = engine.AllPairsPlan(40, 1, 500, 1024, -100, 100)
    tr = DeviceArray.from_numpy(ev)
    n = plan.n_pairs
    bufs = [DeviceArray(n - 1, np.float32), DeviceArray(n, np.int32),
            DeviceArray(n, np.float32), DeviceArray(n, np.int32)]
    try:
        with pytest.raises(ValueError, match="outputs must be"):
            plan.execute(tr, out=bufs)
        bad_type = [DeviceArray(n, np.float32)] * 4
        with pytest.raises(ValueError, match="outputs must be"):
            plan.execute(tr, out=bad_type)
        bad_type[0].free()
        ok = plan.alloc_outputs()
        plan.execute(tr, out=ok)
        assert ok[0].numpy().shape == (n,)
        for o in ok:
            o.free()
    finally:
        for b in bufs:
            b.free()
        tr.free()
        plan.close()


def test_plans_on_concurrent_streams():
    """Three plans (two FFT periods and the bf16x3 lag form) launched on three
    streams at once give bitwise the results of sequential runs on the
    default stream: plans share no device state."""
    from seismic_amd import engine
    from seismic_amd.device import DeviceArray, Stream, synchronize
    print("case plans on 3 concurrent streams", flush=True)
    cfgs = [(200, 1, 900, 2048, -500, 500, engine.FORM_FFT_F32),
            (150, 3, 1500, 4096, -1200, 1300, engine.FORM_FFT_F32),
            (180, 2, 700, 0, -100, 100, engine.FORM_LAG_BF16X3)]
    plans, trs, seq = [], [], []
    for k, (n, c, L, P, lo, hi, form) in enumerate(cfgs):
        ev, _ = synth.make_events(n, c, L, seed=40 + k)
        plans.append(engine.AllPairsPlan(n, c, L, P, lo, hi, form=form))
        trs.append(DeviceArray.from_numpy(ev))
    try:
        for pl, tr in zip(plans, trs):
            outs = pl.execute(tr)
            synchronize()
            seq.append([o.numpy() for o in outs])
            for o in outs:
                o.free()
        streams = [Stream() for _ in plans]
        outs = [pl.alloc_outputs() for pl in plans]
        for _ in range(2):
            for pl, tr, o, st in zip(plans, trs, outs, streams):
                pl.execute(tr, out=o, stream=st)
        for st in streams:
            st.synchronize()
        for k in range(len(plans)):
            for a, b in zip(seq[k], [o.numpy() for o in outs[k]]):
                assert np.array_equal(a, b), k
        for st, oo in zip(streams, outs):
            st.destroy()
            for o in oo:
                o.free()
    finally:
        for pl, tr in zip(plans, trs):
            pl.close()
            tr.free()


def test_initialise_runtime():
    correlation_lib.initialiseRunTime()


def test_extended_multichannel(oracle):
    """3 channels averaged per lag, P=4096 full window, device-resident plan."""
    from seismic_amd import engine
    from seismic_amd.device import DeviceArray, synchronize
    n, C, L, P = 32, 3, 2048, 4096
    ev, _ = synth.make_events(n, C, L, seed=11)
    print("case multichannel N=32 C=3 L=2048", flush=True)
    plan = engine.AllPairsPlan(n, C, L, P, -P // 2 + 1, P // 2)
    tr = DeviceArray.from_numpy(ev)
    outs = plan.execute(tr)
    synchronize()
    vp, lp, vn, ln = [t.numpy() for t in outs]
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, -P // 2 + 1, P // 2, P=P, circular=True)
    print(f"multichannel: max |dcoef| pos {np.abs(vp - ref['pos']).max():.3e} "
          f"neg {np.abs(vn - ref['neg']).max():.3e}")
    assert np.abs(vp - ref["pos"]).max() <= COEF_TOL
    assert np.abs(vn - ref["neg"]).max() <= COEF_TOL
    sure_p = ref["gap_pos"] >= GPU_EXACT_GAP
    sure_n = ref["gap_neg"] >= GPU_EXACT_GAP
    assert np.array_equal(lp[sure_p], ref["lag_pos"][sure_p])
    assert np.array_equal(ln[sure_n], ref["lag_neg"][sure_n])
    st = plan.stats()
    assert st["tie_overflow"] == 0
    plan.close()


def test_extended_demean_and_short_rows(oracle):
    """FP64 mean removal on device and rows shorter than the period."""
    from seismic_amd import engine
    n, L, P = 30, 900, 2048
    ev, _ = synth.make_events(n, 1, L, seed=12)
    ev = ev + np.float32(3.0)  # offset the device must remove
    vp, lp, vn, ln = engine.allpairs_host(ev, P, -400, 400, demean=True)
    ref_in = (ev.astype(np.float64) - ev.astype(np.float64).mean(-1, keepdims=True)).astype(np.float32)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ref_in, pi, pj, -400, 400, P=P, circular=True)
    assert np.abs(vp - ref["pos"]).max() <= COEF_TOL
    assert np.abs(vn - ref["neg"]).max() <= COEF_TOL
    sure = ref["gap_pos"] >= GPU_EXACT_GAP
    assert np.array_equal(lp[sure], ref["lag_pos"][sure])


@pytest.mark.parametrize("C,full", [(1, False), (2, False), (3, False), (3, True), (4, False)])
def test_extended_p8192_channel_counts(oracle, C, full):
    """P=8192 with 1..4 channels: 8-wave kernel with full twiddle tables
    (C=1), with compact tables (C=2, 3) and with rows read from global
    memory (C=4)."""
    from seismic_amd import engine
    n, L, P = 16, 1500, 8192
    ev, _ = synth.make_events(n, C, L, seed=60 + C)
    lo, hi = (-P // 2 + 1, P // 2) if full else (-1600, 1600)
    print(f"case P=8192 C={C} window [{lo}, {hi}]", flush=True)
    vp, lp, vn, ln = engine.allpairs_host(ev, P, lo, hi)
    pi, pj = oracle.packed_pairs(n)
    # lags beyond +-L hold exact zeros, which never win the strict scan
    ref = oracle.pairs_f64(ev, pi, pj, -1600, 1600, P=P, circular=True)
    dp, dn = np.abs(vp - ref["pos"]).max(), np.abs(vn - ref["neg"]).max()
    print(f"P=8192 C={C}: max |dcoef| pos {dp:.3e} neg {dn:.3e}")
    assert dp <= COEF_TOL and dn <= COEF_TOL
    sure_p = ref["gap_pos"] >= GPU_EXACT_GAP
    sure_n = ref["gap_neg"] >= GPU_EXACT_GAP
    assert np.array_equal(lp[sure_p], ref["lag_pos"][sure_p])
    assert np.array_equal(ln[sure_n], ref["lag_neg"][sure_n])


@pytest.mark.parametrize("P,C,full", [(3072, 1, True), (3072, 3, False), (3072, 8, False),
                                      (5120, 1, True), (5120, 3, False), (5120, 5, False),
                                      (5120, 8, False),
                                      (6144, 1, True), (6144, 3, False), (6144, 4, False),
                                      (6144, 5, False)])
def test_extended_mixed_radix_periods(oracle, P, C, full):
    """Periods 3 * 2^k and 5 * 2^k (a radix-3 or radix-5 stage in the
    register DFT), traces longer than P/2 so the circular wrap matters, over
    the kernel variants: 16 waves (P=3072 with 1 and 3 channels), 12 waves
    (P=3072 with 8, P=5120 with 1 and 3, P=6144 with 1, 3 and 4), 8 waves
    (P=5120 with 5), rows from global memory (P=5120 with 8, P=6144 with
    5 channels)."""
    from seismic_amd import engine
    n, L = 12, P - 200
    ev, _ = synth.make_events(n, C, L, seed=90 + C + P // 1024)
    lo, hi = (-P // 2 + 1, P // 2) if full else (-1000, 1400)
    print(f"case P={P} C={C} L={L} window [{lo}, {hi}]", flush=True)
    vp, lp, vn, ln = engine.allpairs_host(ev, P, lo, hi)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, lo, hi, P=P, circular=True)
    dp, dn = np.abs(vp - ref["pos"]).max(), np.abs(vn - ref["neg"]).max()
    print(f"P={P} C={C}: max |dcoef| pos {dp:.3e} neg {dn:.3e}")
    assert dp <= COEF_TOL and dn <= COEF_TOL
    sure_p = ref["gap_pos"] >= GPU_EXACT_GAP
    sure_n = ref["gap_neg"] >= GPU_EXACT_GAP
    assert np.array_equal(lp[sure_p], ref["lag_pos"][sure_p])
    assert np.array_equal(ln[sure_n], ref["lag_neg"][sure_n])


@pytest.mark.parametrize("P,C,full", [(1024, 8, True), (2048, 8, False), (4096, 5, True), (4096, 5, False),
                                      (4096, 6, False), (4096, 7, True), (4096, 8, False)])
def test_extended_pow2_kernel_variants(oracle, P, C, full):
    """Power-of-two periods over the pair kernel's variants, chosen by what
    the staged row spectra leave of the 160 KiB of LDS: 16 waves (P <= 2048
    up to 8 channels, P=4096 up to 5), 12 waves (P=4096 with 6), 8 waves
    (P=4096 with 7), rows from global memory (P=4096 with 8); traces longer
    than P/2 so the circular wrap matters."""
    from seismic_amd import engine
    n, L = 12, P - 200
    ev, _ = synth.make_events(n, C, L, seed=70 + C + P // 1024)
    lo, hi = (-P // 2 + 1, P // 2) if full else (-300, 500)
    print(f"case P={P} C={C} L={L} window [{lo}, {hi}]", flush=True)
    vp, lp, vn, ln = engine.allpairs_host(ev, P, lo, hi)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, lo, hi, P=P, circular=True)
    dp, dn = np.abs(vp - ref["pos"]).max(), np.abs(vn - ref["neg"]).max()
    print(f"P={P} C={C}: max |dcoef| pos {dp:.3e} neg {dn:.3e}")
    assert dp <= COEF_TOL and dn <= COEF_TOL
    sure_p = ref["gap_pos"] >= GPU_EXACT_GAP
    sure_n = ref["gap_neg"] >= GPU_EXACT_GAP
    assert np.array_equal(lp[sure_p], ref["lag_pos"][sure_p])
    assert np.array_equal(ln[sure_n], ref["lag_neg"][sure_n])


def test_dropin_padded_3072(oracle):
    """The raw correlationCUDA entry point with paddedSize = 3072 (a period a
    cuFFT plan accepts; the reference's caller only passes powers of two):
    period-3072 circular correlation, lags -shift+1 .. shift."""
    import ctypes
    n, L, P, shift = 24, 1500, 3072, 1500
    ev, _ = synth.make_events(n, 1, L, seed=3072)
    rows = np.zeros((n, P), np.float32)
    rows[:, :L] = ev[:, 0]
    rev = np.zeros((n, P), np.float32)
    rev[:, :L] = ev[:, 0, ::-1]  # the binding's reversed rows (correlation_lib.py:61-62)
    m = n * (n + 1) // 2
    vp, vn = np.zeros(m, np.float32), np.zeros(m, np.float32)
    lp, ln = np.zeros(m, np.int32), np.zeros(m, np.int32)
    lib = correlation_lib.initialise_library()
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    print("case drop-in N=24 L=1500 paddedSize=3072", flush=True)
    lib.correlationCUDA(fp(rows), fp(rev), n, L, shift, P, 1, 0.0, fp(vp), ip(lp), fp(vn), ip(ln))
    assert lib.seismic_last_error() == b""
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, -shift + 1, shift, P=P, circular=True)
    dp, dn = np.abs(vp - ref["pos"]).max(), np.abs(vn - ref["neg"]).max()
    print(f"drop-in P=3072: max |dcoef| pos {dp:.3e} neg {dn:.3e}")
    assert dp <= COEF_TOL and dn <= COEF_TOL
    sure_p, sure_n = ref["gap_pos"] >= GPU_EXACT_GAP, ref["gap_neg"] >= GPU_EXACT_GAP
    assert np.array_equal(lp[sure_p], ref["lag_pos"][sure_p])
    assert np.array_equal(ln[sure_n], ref["lag_neg"][sure_n])


def test_edge_one_and_two_events(oracle):
    """N=1 (diagonal only) and N=2 through the drop-in."""
    ev, _ = synth.make_events(2, 1, 1024, seed=71)
    for n in (1, 2):
        x = ev[:n, 0]
        print(f"case edge N={n}", flush=True)
        vp, lp, vn, ln = _run_dropin(x, 512)
        padded, rev, L, P = synth.pad_like_binding(x)
        ref = np_ref.reference_f64(padded, rev, L, 512)
        assert vp.shape == (n * (n + 1) // 2,)
        assert np.abs(vp - ref["pos"]).max() <= COEF_TOL and np.abs(vn - ref["neg"]).max() <= COEF_TOL
        assert abs(vp[0] - 1.0) < 1e-5 and lp[0] == 0   # autocorrelation peak at lag 0


def test_edge_zero_energy_events():
    """All-zero traces give (0, 0) with every partner; other pairs unchanged."""
    from seismic_amd import engine
    n, C, L, P = 12, 2, 1000, 2048
    ev, _ = synth.make_events(n, C, L, seed=72)
    ev0 = ev.copy()
    ev0[3] = 0.0
    ev0[7, 1] = 0.0     # one silent channel: that channel contributes 0
    a = engine.allpairs_host(ev0, P, -P // 2 + 1, P // 2)
    iu = np.triu_indices(n)
    for k, (i, j) in enumerate(zip(*iu)):
        if i == 3 or j == 3:
            assert a[0][k] == 0 and a[1][k] == 0 and a[2][k] == 0 and a[3][k] == 0
    # channel 0 of event 7 unchanged: the pair (7, 7) peaks at 1/2 (one of two channels)
    k77 = 7 * n - 7 * 6 // 2
    assert abs(a[0][k77] - 0.5) < 1e-5 and a[1][k77] == 0


@pytest.mark.parametrize("form", [0, 1, 2])
def test_edge_single_lag_window(oracle, form):
    """lag_lo == lag_hi: the value at that one lag (or 0 when not beyond 0)."""
    from seismic_amd import engine
    n, L, lag = 40, 600, 17
    ev, _ = synth.make_events(n, 1, L, seed=73)
    P = 2048 if form == 0 else 0
    print(f"case single lag form={form}", flush=True)
    vp, lp, vn, ln = engine.allpairs_host(ev, P, lag, lag, form=form)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, lag, lag, P=P if P else 0, circular=form == 0)
    tol = 2e-3 if form == 2 else 1e-5
    assert np.abs(vp - ref["pos"]).max() <= tol and np.abs(vn - ref["neg"]).max() <= tol
    assert set(np.unique(lp)) <= {0, lag} and set(np.unique(ln)) <= {0, lag}
    clear_p = ref["pos"] > 1e-4          # value at the lag clearly positive
    clear_n = ref["neg"] < -1e-4         # clearly negative
    assert (lp[clear_p] == lag).all() and (lp[clear_n] == 0).all()
    assert (ln[clear_n] == lag).all() and (ln[clear_p] == 0).all()


@pytest.mark.parametrize("form", [0, 1, 2])
def test_edge_eight_channels_ragged(oracle, form):
    """C=8 (the maximum; the FFT form reads rows from global memory) and a
    ragged N=45 (not a multiple of the 32-event tile)."""
    from seismic_amd import engine
    n, C, L = 45, 8, 1024
    ev, _ = synth.make_events(n, C, L, seed=74)
    P = 4096 if form == 0 else 0
    lo, hi = -300, 300
    print(f"case C=8 form={form}", flush=True)
    vp, lp, vn, ln = engine.allpairs_host(ev, P, lo, hi, form=form)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, lo, hi, P=P if P else 0, circular=form == 0)
    tol, gap = {0: (1e-5, 2e-6), 1: (1e-5, 4e-5), 2: (2e-3, 1e-3)}[form]
    assert np.abs(vp - ref["pos"]).max() <= tol and np.abs(vn - ref["neg"]).max() <= tol
    sure = ref["gap_pos"] >= gap
    assert np.array_equal(lp[sure], ref["lag_pos"][sure])


def test_dropin_concurrent_host_threads():
    """The drop-in call is reentrant: four host threads at once get the same
    results as one at a time."""
    import threading
    evs = [synth.make_events(40, 1, 1500, seed=90 + t)[0][:, 0] for t in range(4)]
    alone = [_run_dropin(e, 400, 0.1) for e in evs]
    got = [None] * 4

    def work(t):
        got[t] = _run_dropin(evs[t], 400, 0.1)
    print("case drop-in 4 threads", flush=True)
    th = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for a, b in zip(alone, got):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


def test_dropin_kept_device_state_is_reset(monkeypatch):
    """Host-buffer calls keep the device state of their shape between calls
    (plan, input and output buffers).  A reused job must give exactly what a
    fresh one gives: the same shape called on other data, on other data with
    a threshold, and interleaved with another shape, against the same calls
    with the cache off (SEISMIC_XCORR_CACHE_MB=0: every job freed at the end
    of its call).  Below-threshold entries and histograms (OMP ABI) must not
    carry over from the previous call either."""
    from seismic_amd import omp_lib
    evs = [synth.make_events(60, 1, 1500, seed=300 + t)[0][:, 0] for t in range(3)]
    other = synth.make_events(33, 1, 700, seed=310)[0][:, 0]

    def run_all():
        out = []
        for k, e in enumerate(evs):
            out.append(_run_dropin(e, 300, 0.3 if k == 1 else 0.0))
            out.append(_run_dropin(other, 200, 0.0))
            out.append(omp_lib.compute_correlation(e, 300, 1, 1, 0.2, 16))
        return out

    print("case drop-in kept device state", flush=True)
    monkeypatch.setenv("SEISMIC_XCORR_CACHE_MB", "0")
    fresh = run_all()
    monkeypatch.delenv("SEISMIC_XCORR_CACHE_MB")
    kept = run_all()
    kept2 = run_all()
    for a, b, c in zip(fresh, kept, kept2):
        for u, v, w in zip(a, b, c):
            assert np.array_equal(u, v) and np.array_equal(u, w)
