"""Host-buffer calls of at least 1024 events run blocked in event bands:
inputs copied in band by band while earlier bands compute, the last column
band completed row band by row band while finished row bands are copied
out, each launch group resolving its own near ties.  Same kernels per pair
and the same near-tie records, so the outputs must equal the single-launch
device execute's bit for bit: sxc_allpairs_host against sxc_plan_execute,
and the drop-in correlationCUDA (reversed rows) against
sxc_plan_execute_reversed."""
import numpy as np
import pytest

from seismic_amd import correlation_lib, engine, omp_lib, synth
from seismic_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _device_execute(rows, P, lo, hi, trace_len, rev=None, threshold=0.0, demean=False):
    n, C, _ = rows.shape
    plan = engine.AllPairsPlan(n, C, trace_len, P, lo, hi, threshold=threshold, demean=demean)
    tr = DeviceArray.from_numpy(rows)
    out = plan.execute(tr) if rev is None else plan.execute_reversed(tr, DeviceArray.from_numpy(rev))
    synchronize()
    res = [o.numpy() for o in out]
    st = plan.stats()
    plan.close()
    return res, st


@pytest.mark.parametrize("n,C,L,P,lo,hi", [(1024, 1, 1000, 2048, -999, 1000), (1500, 3, 700, 1024, -300, 500)])
def test_banded_host_call_equals_device_execute(n, C, L, P, lo, hi):
    ev, _ = synth.make_events(n, C, L, seed=n + C)
    print(f"case banded host call N={n} C={C} L={L} P={P} lags [{lo}, {hi}]", flush=True)
    got = engine.allpairs_host(ev, P, lo, hi)
    want, st = _device_execute(ev, P, lo, hi, L)
    assert st["near_ties"] > 0  # the groups' resolver launches have records to settle
    for name, g, w in zip(("vals_pos", "lags_pos", "vals_neg", "lags_neg"), got, want):
        assert np.array_equal(g, w), name


def test_banded_dropin_equals_device_execute_reversed():
    n, L = 1200, 1024  # P = 2048, shift = L: lags -1023 .. 1024
    ev, _ = synth.make_events(n, 1, L, seed=1200)
    x = ev[:, 0]
    print(f"case banded drop-in N={n} L={L}", flush=True)
    got = correlation_lib.compute_correlation(x, L, 1, 0.0)
    _, L_, P, padded, rev, _ = correlation_lib.prepare(x)
    assert P == 2048
    want, _ = _device_execute(padded[:, None, :], P, -L_ + 1, L_, L_, rev=rev[:, None, :])
    for name, g, w in zip(("vals_pos", "lags_pos", "vals_neg", "lags_neg"), got, want):
        assert np.array_equal(g, w), name


def test_banded_host_call_demean_threshold():
    """FP64 demeaning in the blocked run (each band's transform writes its
    rows of the demeaned copy the resolver reads) and a threshold."""
    n, C, L, P, lo, hi, thr = 1100, 2, 900, 2048, -600, 400, 0.15
    ev, _ = synth.make_events(n, C, L, seed=1100)
    ev = ev + np.float32(0.3)  # a mean for the demeaning to remove
    print(f"case banded host call with demean N={n} C={C} L={L} P={P} thr={thr}", flush=True)
    got = engine.allpairs_host(ev, P, lo, hi, threshold=thr, demean=True)
    want, st = _device_execute(ev, P, lo, hi, L, threshold=thr, demean=True)
    assert st["near_ties"] > 0
    for name, g, w in zip(("vals_pos", "lags_pos", "vals_neg", "lags_neg"), got, want):
        assert np.array_equal(g, w), name


@pytest.mark.parametrize("thr,nc", [(0.2, 37), (0.0, 20000)])
def test_banded_omp_abi_equals_unblocked(monkeypatch, thr, nc):
    """The OMP ABI (correlationCPP): its histogram and threshold pass runs per
    row band in the blocked run; values, lags and both histograms must equal
    the unblocked run's (SEISMIC_XCORR_BLOCKED=0) bit for bit."""
    n, L = 1100, 1024  # P = 2048, shift = L
    ev, _ = synth.make_events(n, 1, L, seed=1101)
    x = ev[:, 0]
    print(f"case banded OMP ABI N={n} L={L} thr={thr} classes={nc}", flush=True)
    monkeypatch.delenv("SEISMIC_XCORR_BLOCKED", raising=False)
    got = omp_lib.compute_correlation(x, L, 8, 4, thr, nc)
    monkeypatch.setenv("SEISMIC_XCORR_BLOCKED", "0")
    want = omp_lib.compute_correlation(x, L, 8, 4, thr, nc)
    assert int(got[4].sum()) == n * (n + 1) // 2
    for name, g, w in zip(("vals_pos", "lags_pos", "vals_neg", "lags_neg", "hist_pos", "hist_neg"), got, want):
        assert np.array_equal(g, w), name
