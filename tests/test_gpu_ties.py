"""The FP64 tie-break of every form, asserted inside its near-tie band.

Every form records the pairs whose best lag has a rival within its band
(FFT 2e-6, bf16x3 2e-5, fp16 2.5e-4) and re-evaluates the candidates as FP64
sums, taking the smaller lag on exact FP64 ties -- the reference scan's first
occurrence (Correlation_in_OMP/correlation_c.cpp:102-111) applied to exact
values.  So the reported lag must be the FP64 argmax of every pair whose FP64
top-two gap is at least 1e-12 (everything but exact ties, whose order FP64
rounding of different summation orders could decide), and on exact ties the
smallest of the tied lags.

Constructed cases pin what sampled pairs rarely show: several distinct FP64
peaks within 1e-6 of each other (more than two inside every band, so a
resolver that looked only at the two best FP32 lags would miss the winner),
and exact ties, built from integer-valued samples or isolated spikes so the
FP64 sums are exact."""
import numpy as np
import pytest

from seismic_amd import correlation_lib, engine, synth
from seismic_amd.device import DeviceArray, synchronize
from tests import np_ref

pytestmark = pytest.mark.gpu

EXACT_GAP = 1e-12


def _assert_fp64_lags(name, lp, ln, ref_lp, ref_ln, gap_p, gap_n, pos, neg, thr=0.0):
    live_p = pos > thr + 1e-5
    live_n = neg < -thr - 1e-5
    sp = live_p & (gap_p >= EXACT_GAP)
    sn = live_n & (gap_n >= EXACT_GAP)
    bad_p = np.nonzero(sp & (lp != ref_lp))[0]
    bad_n = np.nonzero(sn & (ln != ref_ln))[0]
    print(f"{name}: {int(sp.sum())} pos / {int(sn.sum())} neg checked; FP64 gap < band: "
          f"{int((live_p & (gap_p < 2e-6)).sum())} / {int((live_n & (gap_n < 2e-6)).sum())}; "
          f"mismatches {bad_p.size} / {bad_n.size}", flush=True)
    assert bad_p.size == 0, f"{name}: pos lag != FP64 argmax at {bad_p[:5]}"
    assert bad_n.size == 0, f"{name}: neg lag != FP64 argmax at {bad_n[:5]}"


def test_fft_dropin_n256_l1024_every_pair_is_fp64_argmax():
    """N=256, 1 channel, L=1024, full lags (shift 1024): all 32,896 pairs."""
    ev, _ = synth.make_events(256, 1, 1024, seed=101)
    x = ev[:, 0]
    print("case ties N=256 L=1024", flush=True)
    vp, lp, vn, ln = correlation_lib.compute_correlation(x, 1024, 1, 0.0)
    padded, rev, L, P = synth.pad_like_binding(x)
    ref = np_ref.reference_f64(padded, rev, L, 1024)
    assert np.abs(vp - ref["pos"]).max() <= 1e-5 and np.abs(vn - ref["neg"]).max() <= 1e-5
    _assert_fp64_lags("N=256 L=1024", lp, ln, ref["lagp"], ref["lagn"], ref["gapp"], ref["gapn"],
                      ref["pos"], ref["neg"])


def _fp64_stacked_circular(ev, pi, pj, P):
    """FP64 channel-averaged circular NCC of the listed pairs over every lag
    -P/2+1 .. P/2 (numpy FFT), with the reference scan rule and top-two gaps."""
    x = ev.astype(np.float64)
    C = x.shape[1]
    nr = np.sqrt((x * x).sum(-1))
    F = np.fft.rfft(x, n=P, axis=-1)
    lags = np.arange(-P // 2 + 1, P // 2 + 1)
    out = {k: [] for k in ("pos", "lagp", "gapp", "neg", "lagn", "gapn")}
    for s in range(0, len(pi), 256):
        a, b = pi[s:s + 256], pj[s:s + 256]
        y = F[a] * np.conj(F[b])
        c = np.fft.irfft(y, n=P, axis=-1) / (nr[a] * nr[b])[..., None]
        w = c.mean(1)[:, lags % P]
        pos, lagp, neg, lagn = np_ref.scan_window(w, P // 2)
        gp, gn = np_ref.second_gap(w, P // 2)
        for k, v in (("pos", pos), ("lagp", lagp), ("gapp", gp), ("neg", neg), ("lagn", lagn), ("gapn", gn)):
            out[k].append(v)
    return {k: np.concatenate(v) for k, v in out.items()}


def test_fft_n4096_c3_l2048_sampled_pairs_are_fp64_argmax():
    """The headline workload (N=4096, 3 channels, L=2048, P=4096, full lags):
    all 8.4M pairs on the GPU, 6,000 seeded pairs checked against FP64."""
    n, C, L, P = 4096, 3, 2048, 4096
    ev, _ = synth.make_events(n, C, L, seed=2048)
    print("case ties N=4096 C=3 L=2048", flush=True)
    plan = engine.AllPairsPlan(n, C, L, P, -P // 2 + 1, P // 2)
    tr = DeviceArray.from_numpy(ev)
    out = plan.execute(tr, in_len=L)
    synchronize()
    vp, lp, vn, ln = (o.numpy() for o in out)
    st = plan.stats()
    print(f"near ties resolved in FP64: {st['near_ties']} (overflow {st['tie_overflow']})")
    assert st["tie_overflow"] == 0
    rng = np.random.Generator(np.random.PCG64(11))
    pi = rng.integers(0, n, size=6000)
    pj = rng.integers(0, n, size=6000)
    pi, pj = np.minimum(pi, pj), np.maximum(pi, pj)
    idx = pi.astype(np.int64) * n - pi.astype(np.int64) * (pi - 1) // 2 + (pj - pi)
    ref = _fp64_stacked_circular(ev, pi, pj, P)
    assert np.abs(vp[idx] - ref["pos"]).max() <= 1e-5 and np.abs(vn[idx] - ref["neg"]).max() <= 1e-5
    _assert_fp64_lags("N=4096 C=3", lp[idx], ln[idx], ref["lagp"], ref["lagn"], ref["gapp"], ref["gapn"],
                      ref["pos"], ref["neg"])


def _peak_train(L, positions, amps, width=0):
    """Spikes (or short identical bursts) of the given amplitudes."""
    x = np.zeros(L, np.float32)
    rng = np.random.default_rng(0)
    burst = rng.integers(-3, 4, size=max(1, width)).astype(np.float32) if width else np.ones(1, np.float32)
    burst[0] = 4.0
    for p, a in zip(positions, amps):
        x[p:p + len(burst)] += np.float32(a) * burst
    return x


def _constructed_events(L):
    """Events whose correlations have many FP64 peaks within 1e-6 (distinct)
    or exactly equal, on both sides."""
    m = L // 2
    ev = []
    # 0: single spike (the column partner of the trains)
    ev.append(_peak_train(L, [m], [1.0]))
    # 1: five spikes, amplitudes within 1.2e-6 relative, the largest in the middle
    ev.append(_peak_train(L, [m - 300, m - 150, m + 20, m + 170, m + 290],
                          [1.0, 1.0 + 6e-7, 1.0 + 1.2e-6, 1.0 + 3e-7, 1.0 + 9e-7]))
    # 2: the same with negative amplitudes (neg side), largest magnitude first-but-one
    ev.append(_peak_train(L, [m - 250, m - 100, m + 60, m + 200],
                          [-(1.0 + 2e-7), -(1.0 + 1.1e-6), -(1.0 + 5e-7), -(1.0 + 8e-7)]))
    # 3: exact ties: three equal spikes (pos) and two equal negative ones
    ev.append(_peak_train(L, [m - 200, m + 100, m + 250, m - 50, m + 150], [1.0, 1.0, 1.0, -0.5, -0.5]))
    # 4: bursts of integer samples, amplitude steps 2^-21 (exact in float32)
    ev.append(_peak_train(L, [m - 240, m - 80, m + 90, m + 230],
                          [1.0, 1.0 + 2.0 ** -21, 1.0 + 3 * 2.0 ** -21, 1.0 + 2 * 2.0 ** -21], width=24))
    ev.append(_peak_train(L, [m], [1.0], width=24))
    # 6: integer-valued noise: its autocorrelation is exactly symmetric, so
    # the extrema at +-l are exact FP64 ties (smaller lag first)
    rng = np.random.default_rng(9)
    ev.append(rng.integers(-8, 9, size=L).astype(np.float32))
    return np.stack(ev)


@pytest.mark.parametrize("form,P", [(engine.FORM_FFT_F32, 2048), (engine.FORM_LAG_BF16X3, 0),
                                    (engine.FORM_LAG_FP16, 0)])
def test_constructed_near_and_exact_ties(oracle, form, P):
    L, w = 1024, 400
    ev = _constructed_events(L)
    n = len(ev)
    print(f"case constructed ties form={form}", flush=True)
    vp, lp, vn, ln = engine.allpairs_host(ev, P, -w, w, form=form)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, -w, w, P=P, circular=form == engine.FORM_FFT_F32)
    coef_tol = 2e-3 if form == engine.FORM_LAG_FP16 else 1e-5
    assert np.abs(vp - ref["pos"]).max() <= coef_tol and np.abs(vn - ref["neg"]).max() <= coef_tol
    _assert_fp64_lags(f"constructed form={form}", lp, ln, ref["lag_pos"], ref["lag_neg"], ref["gap_pos"],
                      ref["gap_neg"], ref["pos"], ref["neg"])
    # exact ties: the smallest tied lag (the oracle's first occurrence over increasing lags)
    ties_p = (ref["pos"] > 1e-5) & (ref["gap_pos"] == 0.0)
    ties_n = (ref["neg"] < -1e-5) & (ref["gap_neg"] == 0.0)
    print(f"exact FP64 ties: {int(ties_p.sum())} pos, {int(ties_n.sum())} neg")
    assert ties_p.sum() >= 2 and ties_n.sum() >= 2
    assert np.array_equal(lp[ties_p], ref["lag_pos"][ties_p])
    assert np.array_equal(ln[ties_n], ref["lag_neg"][ties_n])
    idx = lambda i, j: i * n - i * (i - 1) // 2 + (j - i)
    m = L // 2
    # the trains against the single spike: the winners are the FP64 ones
    assert lp[idx(0, 1)] == -20 and ln[idx(0, 2)] == 100   # lag l: x_0[n] x_1[n - l], x_1 peaks at m + 20
    assert lp[idx(0, 3)] == -250 and ln[idx(0, 3)] == -150  # exact ties -> smallest lag
    assert lp[idx(4, 5)] == 90
    k = idx(6, 6)
    assert ln[k] == ref["lag_neg"][k] and ln[k] < 0, "symmetric autocorrelation: the negative lag comes first"
    del m


def test_dropin_exact_ties_take_smallest_lag(oracle):
    """Through correlationCUDA: integer-valued samples make the noise
    autocorrelation's extrema at -l and +l exact FP64 ties (the oracle's
    direct FP64 sums are exact integers), and a spike pair gives equal maxima
    at two lags; the smaller lag wins both."""
    L = 1024
    rng = np.random.default_rng(9)
    noise = rng.integers(-8, 9, size=L).astype(np.float32)
    a = np.zeros(L, np.float32)
    a[300] = 1.0
    a[700] = 1.0
    b = np.zeros(L, np.float32)
    b[500] = 1.0
    ev = np.stack([noise, a, b])
    vp, lp, vn, ln = correlation_lib.compute_correlation(ev, 1024, 1, 0.0)
    pi, pj = oracle.packed_pairs(3)
    ref = oracle.pairs_f64(ev, pi, pj, -1023, 1024, P=2048, circular=True)
    # pair (0, 0): neg extremum at +-l, exact tie -> -l
    assert ref["gap_neg"][0] == 0.0 and ln[0] == ref["lag_neg"][0] and ln[0] < 0
    # pair (1, 2) (packed index 4): equal maxima at lags -200 and +200
    assert abs(vp[4] - np.sqrt(0.5)) < 1e-6 and lp[4] == -200
