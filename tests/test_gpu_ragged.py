"""Events of uneven length through the three drop-in libraries.

The reference binding pads every event to P and flips each one over its own
length (Correlation_in_CUDA/correlation_lib.py:59-62), while the scan is
centred on event 0's length L (correlation_lib.py:45, :89;
Correlation_in_OMP/correlation_c.cpp:102-111) and the norms are taken over
the first L samples (:195-204).  For an event j of length L_j != L the
reference's value at reported lag l is therefore the correlation at true lag
l + L - L_j, and its coefficient can exceed 1 when L_j > L.  The drop-ins
transform the caller's reversed rows and form X_i R_j exactly like the
reference, so they must reproduce all of this.

Bars: coefficients within 1e-5 of the reference algorithm in FP64 (numpy,
tests/np_ref.py: literally rfft(events) * rfft(reversed), unnormalised
irfft, scan corr_t[L + i]) and of the oracle's C/OMP restatement; lags equal
to the FP64 argmax wherever the FP64 gap between the top two values is at
least 1e-12 (the GPU settles near ties with FP64 sums of the same
convolution), and to the oracle's wherever that gap is at least 1e-5;
histograms of the OMP ABI as in test_gpu_omp.py."""
import numpy as np
import pytest

from seismic_amd import correlation_lib, engine, multicuda_lib, omp_lib, synth
from seismic_amd.device import DeviceArray, synchronize
from tests import np_ref

pytestmark = pytest.mark.gpu

EXACT_GAP = 1e-12


def _ragged(lens, seed):
    """Object array of traces of the given lengths (noise + Ricker, cut)."""
    ev, _ = synth.make_events(len(lens), 1, max(lens), seed=seed)
    out = np.empty(len(lens), dtype=object)
    for i, n in enumerate(lens):
        out[i] = np.ascontiguousarray(ev[i, 0, :n])
    return out


def _lens(n, L0, lo, hi, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, size=n)
    lens[0] = L0
    # a few exactly at the bounds: shorter and longer than event 0, and P itself
    lens[1], lens[2], lens[3] = lo, hi, L0
    return [int(v) for v in lens]


def _classes(v, n):
    d = np.float32(1.0 / n)
    q = (np.abs(v).astype(np.float32) / d).astype(np.int32)
    return np.where(np.abs(v) >= 1, n - 1, np.minimum(q, n - 1))


def _check_lags(name, lp, ln, ref, thr):
    near_thr = (np.abs(ref["rawpos"] - thr) < 1e-5) | (np.abs(ref["rawneg"] + thr) < 1e-5)
    sure_p = (ref["gapp"] >= EXACT_GAP) & ~near_thr
    sure_n = (ref["gapn"] >= EXACT_GAP) & ~near_thr
    bad_p = np.nonzero(sure_p & (lp != ref["lagp"]))[0]
    bad_n = np.nonzero(sure_n & (ln != ref["lagn"]))[0]
    print(f"{name}: FP64 gaps < 2e-6: {int((ref['gapp'] < 2e-6).sum())} pos "
          f"{int((ref['gapn'] < 2e-6).sum())} neg; lag mismatches {bad_p.size} / {bad_n.size}")
    assert bad_p.size == 0, f"{name}: pos lags differ from FP64 at {bad_p[:5]}"
    assert bad_n.size == 0, f"{name}: neg lags differ from FP64 at {bad_n[:5]}"
    return near_thr


CASES = [  # (n, L0, shortest, longest, shift, threshold)
    (48, 1500, 700, 2048, 548, 0.0),
    (48, 1500, 400, 2048, 300, 0.1),
    (40, 1100, 300, 2048, 948, 0.0),
    (40, 700, 200, 1024, 300, 0.05),
    (32, 3000, 1000, 4096, 1096, 0.0),
]


@pytest.mark.parametrize("n,L0,lo,hi,shift,thr", CASES)
def test_ragged_dropins_vs_reference(oracle, n, L0, lo, hi, shift, thr):
    lens = _lens(n, L0, lo, hi, seed=n + L0 + shift)
    ev = _ragged(lens, seed=L0 + shift)
    print(f"case ragged n={n} L0={L0} lengths [{min(lens)}, {max(lens)}] shift={shift} thr={thr}", flush=True)
    padded, rev, L, P = synth.pad_like_binding(ev)
    ref = np_ref.reference_f64(padded, rev, L, shift, threshold=thr)
    ovp, olp, ovn, oln, _, _ = oracle.correlation_cpp(padded, rev, L, shift, threshold=thr)
    # the length of event j moves its lag by L - L_j: check this really is a ragged case
    assert max(lens) > L > min(lens)

    cuda = correlation_lib.compute_correlation(ev, shift, 1, thr)
    multi = multicuda_lib.compute_correlation(ev, shift, 4, thr)[:4]
    for a, b in zip(cuda, multi):
        assert np.array_equal(a, b), "MultiCUDA drop-in differs from the single-GPU one"
    vp, lp, vn, ln = cuda
    dp, dn = np.abs(vp - ref["pos"]), np.abs(vn - ref["neg"])
    near_thr = _check_lags("correlationCUDA", lp, ln, ref, thr)
    print(f"ragged: max |dcoef| vs FP64 pos {dp[~near_thr].max():.2e} neg {dn[~near_thr].max():.2e}")
    assert dp[~near_thr].max() <= 1e-5 and dn[~near_thr].max() <= 1e-5
    # against the C/OMP restatement (FP32 with its own FFT)
    assert np.abs(vp - ovp)[~near_thr].max() <= 1e-5 and np.abs(vn - ovn)[~near_thr].max() <= 1e-5
    sp = (ref["gapp"] >= 1e-5) & ~near_thr
    sn = (ref["gapn"] >= 1e-5) & ~near_thr
    assert np.array_equal(lp[sp], olp[sp]) and np.array_equal(ln[sn], oln[sn])

    # OMP ABI with histograms
    nc = 50
    ovp2, olp2, ovn2, oln2, ohp, ohn = oracle.correlation_cpp(padded, rev, L, shift, threshold=thr,
                                                              n_classes_hist=nc)
    wp, wlp, wn, wln, hp, hn = omp_lib.compute_correlation(ev, shift, 8, 4, thr, nc)
    _check_lags("correlationCPP", wlp, wln, ref, thr)
    assert np.abs(wp - ref["pos"])[~near_thr].max() <= 1e-5 and np.abs(wn - ref["neg"])[~near_thr].max() <= 1e-5
    npairs = n * (n + 1) // 2
    assert hp.sum() == npairs and hn.sum() == npairs
    if thr == 0.0:
        assert np.array_equal(hp, np.bincount(_classes(wp, nc), minlength=nc))
        assert np.array_equal(hn, np.bincount(_classes(wn, nc), minlength=nc))
    d = 1.0 / nc
    for h, oh, v in ((hp, ohp, ref["rawpos"]), (hn, ohn, -ref["rawneg"])):
        frac = np.abs(v) / d
        edge = (np.minimum(frac - np.floor(frac), np.ceil(frac) - frac) * d < 1e-5) | (np.abs(np.abs(v) - 1) < 1e-5)
        moved = np.abs(np.cumsum(h.astype(np.int64) - oh)).max()
        assert moved <= int(edge.sum()), (moved, int(edge.sum()))


def test_ragged_equal_lengths_match_plain_execute():
    """When every event has length L the reversed-row execute equals the
    forward-only execute to FP32 rounding, lags identical off near ties."""
    n, L, P = 64, 2000, 2048
    ev, _ = synth.make_events(n, 1, L, seed=77)
    x = ev[:, 0]
    padded, rev, L_, P_ = synth.pad_like_binding(x)
    plan = engine.AllPairsPlan(n, 1, L, P, -500, 500)
    tr = DeviceArray.from_numpy(padded)
    rv = DeviceArray.from_numpy(rev)
    oa = plan.execute(tr)
    synchronize()
    a = [o.numpy() for o in oa]
    ob = plan.execute_reversed(tr, rv)
    synchronize()
    b = [o.numpy() for o in ob]
    assert np.abs(a[0] - b[0]).max() < 1e-6 and np.abs(a[2] - b[2]).max() < 1e-6
    ref = np_ref.reference_f64(padded, rev, L_, 500)
    sp, sn = ref["gapp"] >= EXACT_GAP, ref["gapn"] >= EXACT_GAP
    assert np.array_equal(a[1][sp], b[1][sp]) and np.array_equal(a[3][sn], b[3][sn])


def test_reversed_rows_taken_literally(oracle):
    """The reversed rows are used as given, as the reference does: rows that
    are not the flip of their event (here another event's flip, and a flip
    shifted by 5 samples) give the reference algorithm's numbers."""
    n, L = 12, 900
    ev, _ = synth.make_events(n, 1, L, seed=12)
    padded, rev, L_, P = synth.pad_like_binding(ev[:, 0])
    rev2 = rev.copy()
    rev2[3] = rev[7]
    rev2[5] = np.roll(rev[5], 5)
    ref = np_ref.reference_f64(padded, rev2, L_, 124)
    lib = correlation_lib.initialise_library()
    ne = n * (n + 1) // 2
    outs = [np.zeros(ne, np.float32), np.zeros(ne, np.int32), np.zeros(ne, np.float32), np.zeros(ne, np.int32)]
    fp = lambda a: a.ctypes.data_as(correlation_lib.c_float_p)
    ip = lambda a: a.ctypes.data_as(correlation_lib.c_int_p)
    lib.correlationCUDA(fp(padded), fp(rev2), n, L_, 124, P, 1, 0.0, fp(outs[0]), ip(outs[1]), fp(outs[2]),
                        ip(outs[3]))
    assert lib.seismic_last_error() == b""
    assert np.abs(outs[0] - ref["pos"]).max() <= 1e-5 and np.abs(outs[2] - ref["neg"]).max() <= 1e-5
    _check_lags("literal reversed rows", outs[1], outs[3], ref, 0.0)


def test_reversed_multichannel_plan():
    """sxc_plan_execute_reversed with 3 channels and uneven lengths: the
    channel average of the reference's per-channel X_i R_j at lag l."""
    n, C, P = 20, 3, 2048
    L = 1400
    rng = np.random.default_rng(4)
    lens = rng.integers(600, 2049, size=n)
    lens[0] = L
    base, _ = synth.make_events(n, C, 2048, seed=31)
    tr = np.zeros((n, C, P), np.float32)
    rv = np.zeros((n, C, P), np.float32)
    for i, m in enumerate(lens):
        tr[i, :, :m] = base[i, :, :m]
        rv[i, :, :m] = base[i, :, :m][:, ::-1]
    lo, hi = -300, 301
    plan = engine.AllPairsPlan(n, C, L, P, lo, hi)
    dtr, drv = DeviceArray.from_numpy(tr), DeviceArray.from_numpy(rv)
    oo = plan.execute_reversed(dtr, drv)
    synchronize()
    out = [o.numpy() for o in oo]
    # FP64: per channel irfft(X R) / (norms over the first L samples), averaged
    X = np.fft.rfft(tr.astype(np.float64), axis=-1)
    Rv = np.fft.rfft(rv.astype(np.float64), axis=-1)
    nr = np.sqrt((tr[:, :, :L].astype(np.float64) ** 2).sum(-1))
    lags = np.arange(lo, hi + 1)
    k = 0
    for i in range(n):
        for j in range(i, n):
            ct = np.fft.irfft(X[i] * Rv[j], n=P, axis=-1)  # [C, P]
            w = (ct[:, (L - 1 + lags) % P] / (nr[i] * nr[j])[:, None]).mean(0)
            a = int(np.argmax(w))
            srt = np.sort(w)[::-1]
            assert abs(out[0][k] - w[a]) <= 1e-5
            if srt[0] - srt[1] >= EXACT_GAP:
                assert out[1][k] == lags[a], (i, j, out[1][k], lags[a])
            k += 1


def test_cuda_abi_window_default_and_compat(oracle, monkeypatch):
    """shift = 250 (the original script's value).  By default the CUDA ABIs
    scan lags -249 .. 250 like the C/OMP path; with
    SEISMIC_XCORR_WINDOW=cuda they scan the CUDA references' whole 128-lag
    strides, -127 .. 128 (Correlation_in_CUDA/correlation_cuda.cu:114), and a
    shift below 128 scans nothing.  The OMP ABI is not affected."""
    n, L = 30, 1500  # P = 2048: shift 250 needs L - 250 >= 0 and L + 250 <= P
    ev, _ = synth.make_events(n, 1, L, seed=250)
    x = ev[:, 0]
    padded, rev, L_, P = synth.pad_like_binding(x)
    monkeypatch.delenv("SEISMIC_XCORR_WINDOW", raising=False)
    full = correlation_lib.compute_correlation(x, 250, 1, 0.0)
    ref250 = np_ref.reference_f64(padded, rev, L_, 250)
    assert np.abs(full[0] - ref250["pos"]).max() <= 1e-5
    _check_lags("shift 250", full[1], full[3], ref250, 0.0)
    assert np.abs(full[1]).max() > 128 or np.abs(full[3]).max() > 128
    monkeypatch.setenv("SEISMIC_XCORR_WINDOW", "cuda")
    ref128 = np_ref.reference_f64(padded, rev, L_, 128)
    for got in (correlation_lib.compute_correlation(x, 250, 1, 0.0),
                multicuda_lib.compute_correlation(x, 250, 2, 0.0)[:4]):
        assert np.abs(got[0] - ref128["pos"]).max() <= 1e-5 and np.abs(got[2] - ref128["neg"]).max() <= 1e-5
        _check_lags("shift 250, CUDA window", got[1], got[3], ref128, 0.0)
        assert got[1].min() >= -127 and got[1].max() <= 128
    empty = correlation_lib.compute_correlation(x, 100, 1, 0.0)
    assert not any(a.any() for a in empty)
    omp = omp_lib.compute_correlation(x, 250, 4, 4, 0.0, 20)
    assert np.abs(omp[0] - ref250["pos"]).max() <= 1e-5
