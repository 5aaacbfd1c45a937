"""Device-resident all-pairs plans over the extended C ABI
(include/seismic_xcorr.h).  Device memory comes from seismic_amd.device (the
engine's own HIP runtime); all compute runs in libseismic_xcorr.so."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import SxcPlanDesc, SxcStats, SxcTiming, c_float_p, c_int_p, last_error, load_engine

FORM_FFT_F32 = 0
FORM_LAG_BF16X3 = 1
FORM_LAG_FP16 = 2


def _desc(n_events, n_channels, trace_len, padded_size, lag_lo, lag_hi, threshold, tie_tol,
          demean, form):
    return SxcPlanDesc(int(n_events), int(n_channels), int(trace_len), int(padded_size),
                       int(lag_lo), int(lag_hi), float(threshold), float(tie_tol),
                       1 if demean else 0, int(form))


def n_packed(n_events: int) -> int:
    return n_events * (n_events + 1) // 2


class AllPairsPlan:
    """All-pairs NCC of N events x C channels, FFT period P, lag window
    [lag_lo, lag_hi] (circular mod P), packed upper-triangle outputs."""

    def __init__(self, n_events, n_channels, trace_len, padded_size, lag_lo, lag_hi,
                 threshold=0.0, tie_tol=0.0, demean=False, form=FORM_FFT_F32):
        self.lib = load_engine()
        self.n_events = int(n_events)
        self.n_channels = int(n_channels)
        self.padded_size = int(padded_size)
        self.desc = _desc(n_events, n_channels, trace_len, padded_size, lag_lo, lag_hi,
                          threshold, tie_tol, demean, form)
        h = ctypes.c_void_p()
        rc = self.lib.sxc_plan_create(ctypes.byref(self.desc), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"sxc_plan_create failed ({rc}): {last_error(self.lib)}")
        self.handle = h

    @property
    def n_pairs(self) -> int:
        return n_packed(self.n_events)

    def alloc_outputs(self):
        from .device import DeviceArray
        n = self.n_pairs
        return (DeviceArray(n, np.float32), DeviceArray(n, np.int32),
                DeviceArray(n, np.float32), DeviceArray(n, np.int32))

    def _check_io(self, traces, out):
        if traces.dtype != np.float32:
            raise ValueError("traces must be float32")
        stride = traces.shape[-1]
        if int(np.prod(traces.shape)) != self.n_events * self.n_channels * stride:
            raise ValueError("traces shape does not match the plan")
        if out is None:
            out = self.alloc_outputs()
        # the kernels write n_pairs entries into each output: a smaller
        # buffer would be written out of bounds on the device
        for o, dt in zip(out, (np.float32, np.int32, np.float32, np.int32)):
            if o.dtype != dt or o.nbytes < self.n_pairs * 4:
                raise ValueError(f"outputs must be {self.n_pairs}-entry float32/int32/float32/int32 "
                                 f"device arrays (got {o.dtype} of {o.nbytes} bytes)")
        return stride, out

    def execute(self, traces, in_len=None, out=None, stream=None):
        """traces: DeviceArray float32 [N, C, stride] (or [N, stride] for C=1).
        Asynchronous on `stream` (a device.Stream; None = default stream)."""
        stride, out = self._check_io(traces, out)
        in_len = stride if in_len is None else int(in_len)
        vp, lp, vn, ln = out
        sh = stream.handle if stream is not None else 0
        rc = self.lib.sxc_plan_execute(self.handle, traces.ptr, stride, in_len,
                                       vp.ptr, lp.ptr, vn.ptr, ln.ptr, ctypes.c_void_p(sh))
        if rc != 0:
            raise RuntimeError(f"sxc_plan_execute failed ({rc}): {last_error(self.lib)}")
        return out

    def execute_reversed(self, traces, reversed_traces, in_len=None, out=None, stream=None):
        """The reference's X_i * R_j with the caller's reversed rows
        (sxc_plan_execute_reversed); both DeviceArrays of the same shape."""
        stride, out = self._check_io(traces, out)
        if reversed_traces.dtype != np.float32 or reversed_traces.shape != traces.shape:
            raise ValueError("reversed traces must be float32 of the traces' shape")
        in_len = stride if in_len is None else int(in_len)
        vp, lp, vn, ln = out
        sh = stream.handle if stream is not None else 0
        rc = self.lib.sxc_plan_execute_reversed(self.handle, traces.ptr, reversed_traces.ptr, stride, in_len,
                                                vp.ptr, lp.ptr, vn.ptr, ln.ptr, ctypes.c_void_p(sh))
        if rc != 0:
            raise RuntimeError(f"sxc_plan_execute_reversed failed ({rc}): {last_error(self.lib)}")
        return out

    def stats(self):
        s = SxcStats()
        rc = self.lib.sxc_plan_stats(self.handle, ctypes.byref(s))
        if rc != 0:
            raise RuntimeError(f"sxc_plan_stats failed ({rc}): {last_error(self.lib)}")
        return {"pairs": s.pairs, "near_ties": s.near_ties, "tie_overflow": s.tie_overflow}

    def enable_timing(self, on=True):
        rc = self.lib.sxc_plan_enable_timing(self.handle, 1 if on else 0)
        if rc != 0:
            raise RuntimeError(f"sxc_plan_enable_timing failed ({rc}): {last_error(self.lib)}")

    def timing(self):
        t = SxcTiming()
        rc = self.lib.sxc_plan_timing(self.handle, ctypes.byref(t))
        if rc != 0:
            raise RuntimeError(f"sxc_plan_timing failed ({rc}): {last_error(self.lib)}")
        return {"executes": t.executes, "ms_forward": t.ms_forward, "ms_pairs": t.ms_pairs,
                "ms_resolve": t.ms_resolve}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.sxc_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def allpairs_host(traces, padded_size, lag_lo, lag_hi, trace_len=None, threshold=0.0,
                  tie_tol=0.0, demean=False, in_len=None, form=FORM_FFT_F32):
    """Host-memory convenience call.  traces: [N, C, stride] or [N, stride] float32."""
    lib = load_engine()
    tr = np.ascontiguousarray(traces, dtype=np.float32)
    if tr.ndim == 2:
        tr = tr[:, None, :]
    n, c, stride = tr.shape
    L = stride if trace_len is None else trace_len
    d = _desc(n, c, L, padded_size, lag_lo, lag_hi, threshold, tie_tol, demean, form)
    npk = n_packed(n)
    vp = np.zeros(npk, np.float32); lp = np.zeros(npk, np.int32)
    vn = np.zeros(npk, np.float32); ln = np.zeros(npk, np.int32)
    rc = lib.sxc_allpairs_host(ctypes.byref(d), tr.ctypes.data_as(c_float_p), stride,
                               stride if in_len is None else int(in_len),
                               vp.ctypes.data_as(c_float_p), lp.ctypes.data_as(c_int_p),
                               vn.ctypes.data_as(c_float_p), ln.ctypes.data_as(c_int_p))
    if rc != 0:
        raise RuntimeError(f"sxc_allpairs_host failed ({rc}): {last_error(lib)}")
    return vp, lp, vn, ln
