"""sxc_plan_execute_noalloc against sxc_plan_execute, through the C ABI.

The no-allocation entry point runs the same kernels with the same launch
shapes on workspace held in the library's static device array, so on the
same inputs its four packed outputs must equal the plan execute's bit for
bit.  Two calls in a row, the second smaller (the array keeps the first
call's contents: the second must not see them), then sizes beyond the
array's capacity must be refused with an error before any launch."""
import ctypes

import numpy as np
import pytest

from seismic_amd import _lib, engine, synth
from seismic_amd.device import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def _noalloc(lib, desc, tr, stride, in_len, outs, stream):
    vp, lp, vn, ln = outs
    return lib.sxc_plan_execute_noalloc(ctypes.byref(desc), ctypes.c_void_p(tr.ptr), stride, in_len,
                                        ctypes.c_void_p(vp.ptr), ctypes.c_void_p(lp.ptr),
                                        ctypes.c_void_p(vn.ptr), ctypes.c_void_p(ln.ptr),
                                        ctypes.c_void_p(stream.handle))


@pytest.mark.parametrize("cases", [
    # (N, C, L, P, lag_lo, lag_hi, threshold, demean): first call, then a smaller one
    ((700, 3, 2048, 4096, -2047, 2048, 0.0, False), (130, 3, 1500, 2048, -300, 301, 0.1, False)),
    ((512, 1, 3000, 6144, -1000, 1200, 0.0, True), (64, 2, 900, 1024, -511, 512, 0.0, False)),
])
def test_noalloc_matches_plan_execute(cases):
    lib = _lib.load_engine()
    stream = Stream()
    for k, (n, c, L, P, lo, hi, thr, dm) in enumerate(cases):
        print(f"case noalloc call {k}: N={n} C={c} L={L} P={P} [{lo},{hi}] thr={thr} demean={dm}", flush=True)
        ev, _ = synth.make_events(n, c, L, seed=n + L + k)
        tr = DeviceArray.from_numpy(ev)
        plan = engine.AllPairsPlan(n, c, L, P, lo, hi, threshold=thr, demean=dm)
        ref = plan.execute(tr, in_len=L, stream=stream)
        stream.synchronize()
        ref = [o.numpy() for o in ref]
        outs = plan.alloc_outputs()
        for o in outs:
            o.zero_()
        rc = _noalloc(lib, plan.desc, tr, L, L, outs, stream)
        assert rc == 0, _lib.last_error(lib)
        stream.synchronize()
        got = [o.numpy() for o in outs]
        for a, b, name in zip(ref, got, ("vals_pos", "lags_pos", "vals_neg", "lags_neg")):
            assert np.array_equal(a, b), f"call {k}: {name} differs from sxc_plan_execute"
        plan.close()


def test_noalloc_rejects_oversize():
    lib = _lib.load_engine()
    stream = Stream()
    n, c, L, P = 20000, 3, 4096, 8192   # spectra alone ~2 GB
    desc = engine._desc(n, c, L, P, -100, 100, 0.0, 0.0, False, engine.FORM_FFT_F32)
    tiny = DeviceArray(16, np.float32)
    rc = lib.sxc_plan_execute_noalloc(ctypes.byref(desc), ctypes.c_void_p(tiny.ptr), L, L,
                                      ctypes.c_void_p(tiny.ptr), ctypes.c_void_p(tiny.ptr),
                                      ctypes.c_void_p(tiny.ptr), ctypes.c_void_p(tiny.ptr),
                                      ctypes.c_void_p(stream.handle))
    assert rc == -1 and "no-allocation entry point holds at most" in _lib.last_error(lib)
    lag = engine._desc(64, 1, 512, 0, -10, 10, 0.0, 0.0, False, engine.FORM_LAG_BF16X3)
    rc = lib.sxc_plan_execute_noalloc(ctypes.byref(lag), ctypes.c_void_p(tiny.ptr), 512, 512,
                                      ctypes.c_void_p(tiny.ptr), ctypes.c_void_p(tiny.ptr),
                                      ctypes.c_void_p(tiny.ptr), ctypes.c_void_p(tiny.ptr),
                                      ctypes.c_void_p(stream.handle))
    assert rc == -1 and "FFT form only" in _lib.last_error(lib)
