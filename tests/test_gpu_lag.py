"""GPU parity of the lag-window MFMA forms against the FP64 direct oracle
(linear correlation over [lag_lo, lag_hi]), through the C ABI.

Bars: bf16x3 |dcoef| <= 1e-5, fp16 |dcoef| <= 2e-3; lags equal to the FP64
argmax wherever the FP64 gap between the top two values is at least 1e-12
(each form's near-tie band -- bf16x3 2e-5, fp16 2.5e-4 -- is re-evaluated in
FP64, every in-window lag when more than two lie inside it)."""
import numpy as np
import pytest

from seismic_amd import engine, synth

pytestmark = pytest.mark.gpu

FORMS = {"bf16x3": (engine.FORM_LAG_BF16X3, 1e-5, 1e-12), "fp16": (engine.FORM_LAG_FP16, 2e-3, 1e-12)}


def _run(ev, lo, hi, form, thr=0.0):
    n, c, L = ev.shape
    return engine.allpairs_host(ev, 0, lo, hi, threshold=thr, form=form)


@pytest.mark.parametrize("name", ["bf16x3", "fp16"])
@pytest.mark.parametrize("n,C,L,lo,hi,thr", [(64, 1, 1024, -100, 100, 0.0), (50, 1, 777, -37, 90, 0.1),
                                           (40, 3, 2048, -512, 512, 0.0), (33, 2, 300, -299, 299, 0.2)])
def test_lag_form_vs_f64(oracle, name, n, C, L, lo, hi, thr):
    form, coef_tol, gap = FORMS[name]
    ev, _ = synth.make_events(n, C, L, seed=n * 7 + L)
    print(f"case lag {name} n={n} C={C} L={L} [{lo},{hi}] thr={thr}", flush=True)
    vp, lp, vn, ln = _run(ev, lo, hi, form, thr)
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, lo, hi, circular=False)
    rp = np.where(ref["pos"] > thr, ref["pos"], 0.0)
    rn = np.where(ref["neg"] < -thr, ref["neg"], 0.0)
    dp, dn = np.abs(vp - rp).max(), np.abs(vn - rn).max()
    print(f"{name}: max |dcoef| pos {dp:.3e} neg {dn:.3e}")
    assert dp <= coef_tol and dn <= coef_tol
    kp = ref["pos"] > thr + 1e-3
    kn = ref["neg"] < -thr - 1e-3
    sure_p = kp & (ref["gap_pos"] >= gap)
    sure_n = kn & (ref["gap_neg"] >= gap)
    agree_p = (lp[kp] == ref["lag_pos"][kp]).mean() if kp.any() else 1.0
    print(f"{name}: lag agreement pos {agree_p:.5f}")
    assert np.array_equal(lp[sure_p], ref["lag_pos"][sure_p])
    assert np.array_equal(ln[sure_n], ref["lag_neg"][sure_n])


def test_lag_matches_fft_form(oracle):
    """With P >= L + max|lag| the FFT form's circular lags equal linear lags:
    the two forms must agree (values to 1e-5; lags on every pair whose top two
    FP64 values are not an exact tie -- e.g. the diagonal pairs' symmetric
    autocorrelation has equal values at -l and +l, where the summation order
    of each form's FP64 re-evaluation decides)."""
    n, C, L, P = 48, 3, 2048, 4096
    ev, _ = synth.make_events(n, C, L, seed=5)
    a = engine.allpairs_host(ev, P, -512, 512)
    b = engine.allpairs_host(ev, 0, -512, 512, form=engine.FORM_LAG_BF16X3)
    assert np.abs(a[0] - b[0]).max() < 2e-5 and np.abs(a[2] - b[2]).max() < 2e-5
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(ev, pi, pj, -512, 512, circular=False)
    sp = ref["gap_pos"] > 1e-9
    sn = ref["gap_neg"] > 1e-9
    assert np.array_equal(a[1][sp], b[1][sp]) and np.array_equal(a[3][sn], b[3][sn])
    assert (a[1] == b[1]).mean() > 0.99 and (a[3] == b[3]).mean() > 0.99


def test_lag_exact_shift_duplicates():
    """A zero-filled time-shifted copy peaks at exactly its shift with value
    1 (up to precision), in both precision forms."""
    rng = np.random.default_rng(3)
    L = 1024
    base = np.zeros((1, L), np.float32)
    base[0, 300:700] = rng.standard_normal(400)
    ev = np.stack([base, np.roll(base, 41, axis=1), np.roll(base, -77, axis=1)])
    for form, tol in ((engine.FORM_LAG_BF16X3, 1e-5), (engine.FORM_LAG_FP16, 2e-3)):
        vp, lp, vn, ln = engine.allpairs_host(ev, 0, -200, 200, form=form)
        # packed (0,1)=1, (0,2)=2, (1,2)=4
        assert lp[1] == -41 and lp[2] == 77 and lp[4] == 118
        assert abs(vp[1] - 1) < tol and abs(vp[2] - 1) < tol and abs(vp[0] - 1) < tol


def test_n8192_l4096_w512_forms_vs_f64_sampled(oracle):
    """Precision configuration at full size: N=8192, 1 channel, L=4096,
    window +-512, all 33.5M pairs in bf16x3, fp16 and the FP32-FFT form
    (P=8192, circular lags equal the linear ones since P >= L + 512);
    the FP64 oracle on 100,000 seeded pairs.  Prints the coefficient error and
    the lag agreement rate of every form."""
    n, L, w = 8192, 4096, 512
    ev, _ = synth.make_events(n, 1, L, seed=8192)
    rng = np.random.Generator(np.random.PCG64(3))
    k = 100000
    pi = rng.integers(0, n, size=k)
    pj = rng.integers(0, n, size=k)
    pi, pj = np.minimum(pi, pj).astype(np.int32), np.maximum(pi, pj).astype(np.int32)
    idx = pi.astype(np.int64) * n - pi.astype(np.int64) * (pi - 1) // 2 + (pj - pi)
    ref = oracle.pairs_f64(ev, pi, pj, -w, w, circular=False)
    keep = ref["pos"] > 1e-3
    for name, form, P, coef_tol, gap in (("bf16x3", engine.FORM_LAG_BF16X3, 0, 1e-5, 1e-12),
                                         ("fp16", engine.FORM_LAG_FP16, 0, 2e-3, 1e-12),
                                         ("fft_f32", engine.FORM_FFT_F32, 8192, 1e-5, 1e-12)):
        print(f"case N=8192 L=4096 +-512 {name}", flush=True)
        vp, lp, vn, ln = engine.allpairs_host(ev, P, -w, w, form=form)
        dp = np.abs(vp[idx] - ref["pos"]).max()
        dn = np.abs(vn[idx] - ref["neg"]).max()
        agree = (lp[idx][keep] == ref["lag_pos"][keep]).mean()
        print(f"N=8192 L=4096 +-512 {name}: max |dcoef| pos {dp:.3e} neg {dn:.3e}; lag agreement {agree:.6f} "
              f"on {int(keep.sum())} sampled pairs")
        assert dp <= coef_tol and dn <= coef_tol
        sure = keep & (ref["gap_pos"] >= gap)
        assert np.array_equal(lp[idx][sure], ref["lag_pos"][sure])
        keep_n = ref["neg"] < -1e-3
        sure_n = keep_n & (ref["gap_neg"] >= gap)
        assert np.array_equal(ln[idx][sure_n], ref["lag_neg"][sure_n])
        del vp, lp, vn, ln
