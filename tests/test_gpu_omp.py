"""GPU OMP-ABI drop-in (correlation_c.so: correlationCPP with threshold and
histograms) against the oracle's C/OpenMP restatement of
Correlation_in_OMP/correlation_c.cpp.

Bars: coefficients within 1e-5 of the oracle; lags equal except where the
FP64 gap between the top two values is below 1e-5; histograms equal to the
classes of the GPU's own pre-threshold values exactly, and to the oracle's
except for pairs whose value lies within 1e-5 of a class edge."""
import numpy as np
import pytest

from seismic_amd import omp_lib, synth
from tests import np_ref

pytestmark = pytest.mark.gpu


def _classes(v, n):
    d = np.float32(1.0 / n)
    q = (np.abs(v).astype(np.float32) / d).astype(np.int32)
    return np.where(np.abs(v) >= 1, n - 1, np.minimum(q, n - 1))


@pytest.mark.parametrize("n,L,shift,thr,nc", [(256, 1024, 1024, 0.0, 100), (60, 1500, 548, 0.3, 37),
                                               (40, 700, 300, 0.1, 20000)])
def test_omp_abi_vs_oracle(oracle, n, L, shift, thr, nc):
    ev, _ = synth.make_events(n, 1, L, seed=n + shift)
    x = ev[:, 0]
    print(f"case omp-abi n={n} L={L} shift={shift} thr={thr} classes={nc}", flush=True)
    vp, lp, vn, ln, hp, hn = omp_lib.compute_correlation(x, shift, 8, 4, thr, nc)
    padded, rev, Lr, P = synth.pad_like_binding(x)
    ovp, olp, ovn, oln, ohp, ohn = oracle.correlation_cpp(padded, rev, Lr, shift, thr, nc)
    ref = np_ref.reference_f64(padded, rev, Lr, shift)
    npairs = n * (n + 1) // 2
    assert hp.sum() == npairs and hn.sum() == npairs
    dp, dn = np.abs(vp - ovp).max(), np.abs(vn - ovn).max()
    print(f"omp-abi: max |dcoef| vs oracle pos {dp:.2e} neg {dn:.2e}")
    # threshold edges: a value within 1e-5 of the threshold may fall either side
    near_thr = (np.abs(ref["pos"] - thr) < 1e-5) | (np.abs(ref["neg"] + thr) < 1e-5)
    assert np.abs(vp - ovp)[~near_thr].max() <= 1e-5 and np.abs(vn - ovn)[~near_thr].max() <= 1e-5
    sure_p = (ref["gapp"] >= 1e-5) & ~near_thr
    sure_n = (ref["gapn"] >= 1e-5) & ~near_thr
    assert np.array_equal(lp[sure_p], olp[sure_p]) and np.array_equal(ln[sure_n], oln[sure_n])
    if thr == 0.0:
        # the GPU histograms are exactly the classes of its own values
        assert np.array_equal(hp, np.bincount(_classes(vp, nc), minlength=nc))
        assert np.array_equal(hn, np.bincount(_classes(vn, nc), minlength=nc))
    # against the oracle: only pairs within 1e-5 of a class edge may move
    d = 1.0 / nc
    for h, oh, v in ((hp, ohp, ref["pos"]), (hn, ohn, -ref["neg"])):
        frac = np.abs(v) / d
        edge = np.minimum(frac - np.floor(frac), np.ceil(frac) - frac) * d < 1e-5
        moved = np.abs(np.cumsum(h.astype(np.int64) - oh)).max()
        assert moved <= int(edge.sum()), (moved, int(edge.sum()))
