"""TEST INFRASTRUCTURE ONLY: ctypes wrappers of oracle/build/liboracle.so.

  correlation_cpp   restatement of Correlation_in_OMP/correlation_c.cpp:209-298
                    (16-argument correlationCPP ABI, own FP32 FFT)
  pairs_f64         FP64 direct circular/linear NCC for a list of pairs
  rfft_f32/irfft_f32  the oracle's FP32 FFT, exposed so tests pin it to numpy
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    lib.correlationCPP.restype = None
    lib.correlationCPP.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, ctypes.c_int, c_float_p,
                                   c_int_p, c_float_p, c_int_p, c_int_p, c_int_p]
    lib.oracle_pairs_f64.restype = ctypes.c_int
    lib.oracle_pairs_f64.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_longlong, c_int_p, c_int_p, ctypes.c_int,
                                     c_double_p, c_int_p, c_double_p, c_int_p,
                                     c_double_p, c_double_p, c_double_p]
    lib.oracle_allpairs_stacked_f32.restype = None
    lib.oracle_allpairs_stacked_f32.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                                c_float_p, c_int_p, c_float_p, c_int_p]
    lib.oracle_rfft_f32.restype = None
    lib.oracle_rfft_f32.argtypes = [c_float_p, ctypes.c_int, ctypes.c_int, c_float_p, c_float_p]
    lib.oracle_irfft_f32.restype = None
    lib.oracle_irfft_f32.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int, c_float_p]
    _lib = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(c_float_p)


def _ip(a):
    return a.ctypes.data_as(c_int_p)


def _dp(a):
    return a.ctypes.data_as(c_double_p)


def correlation_cpp(padded, padded_rev, event_length, shift, threshold=0.0, n_classes_hist=20,
                    num_threads=0, chunk_size=4):
    """Packed (vals_pos, lags_pos, vals_neg, lags_neg, max_hist, min_hist)."""
    lib = load()
    padded = np.ascontiguousarray(padded, dtype=np.float32)
    padded_rev = np.ascontiguousarray(padded_rev, dtype=np.float32)
    n, p = padded.shape
    ne = n * (n + 1) // 2
    vp = np.zeros(ne, np.float32); lp = np.zeros(ne, np.int32)
    vn = np.zeros(ne, np.float32); ln = np.zeros(ne, np.int32)
    hmax = np.zeros(max(n_classes_hist, 1), np.int32); hmin = np.zeros(max(n_classes_hist, 1), np.int32)
    lib.correlationCPP(_fp(padded), _fp(padded_rev), n, event_length, shift, p,
                       num_threads or (os.cpu_count() or 1), chunk_size, float(threshold),
                       n_classes_hist, _fp(vp), _ip(lp), _fp(vn), _ip(ln), _ip(hmax), _ip(hmin))
    return vp, lp, vn, ln, hmax, hmin


def allpairs_stacked(events, L, S, threshold=0.0, num_threads=0, chunk_size=4):
    """Channel-stacked CPU restatement: events [N, C, P] zero-padded float32."""
    lib = load()
    ev = np.ascontiguousarray(events, dtype=np.float32)
    n, c, p = ev.shape
    rev = np.zeros_like(ev)
    rev[:, :, :L] = ev[:, :, :L][:, :, ::-1]
    ne = n * (n + 1) // 2
    vp = np.zeros(ne, np.float32); lp = np.zeros(ne, np.int32)
    vn = np.zeros(ne, np.float32); ln = np.zeros(ne, np.int32)
    lib.oracle_allpairs_stacked_f32(_fp(ev), _fp(rev), n, c, L, S, p, float(threshold),
                                    num_threads or (os.cpu_count() or 1), chunk_size,
                                    _fp(vp), _ip(lp), _fp(vn), _ip(ln))
    return vp, lp, vn, ln


def pairs_f64(traces, pi, pj, lag_lo, lag_hi, P=0, circular=True, trace_len=None,
              num_threads=0, want_curve=False):
    """traces: [N, C, stride] (or [N, stride]) float32.  Returns dict of arrays."""
    lib = load()
    tr = np.ascontiguousarray(traces, dtype=np.float32)
    if tr.ndim == 2:
        tr = tr[:, None, :]
    n, c, stride = tr.shape
    L = stride if trace_len is None else trace_len
    pi = np.ascontiguousarray(pi, dtype=np.int32)
    pj = np.ascontiguousarray(pj, dtype=np.int32)
    q = pi.shape[0]
    out = {k: np.zeros(q, np.float64) for k in ("pos", "neg", "gap_pos", "gap_neg")}
    out["lag_pos"] = np.zeros(q, np.int32)
    out["lag_neg"] = np.zeros(q, np.int32)
    W = lag_hi - lag_lo + 1
    curve = np.zeros((q, W), np.float64) if want_curve else None
    rc = lib.oracle_pairs_f64(_fp(tr), n, c, L, stride, P, lag_lo, lag_hi, 1 if circular else 0,
                              q, _ip(pi), _ip(pj), num_threads or (os.cpu_count() or 1),
                              _dp(out["pos"]), _ip(out["lag_pos"]), _dp(out["neg"]), _ip(out["lag_neg"]),
                              _dp(out["gap_pos"]), _dp(out["gap_neg"]),
                              _dp(curve) if curve is not None else None)
    if rc != 0:
        raise ValueError("oracle_pairs_f64 rejected its arguments")
    if curve is not None:
        out["curve"] = curve
    return out


def rfft_f32(x, p):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, p)
    n = x.shape[0]
    m = p // 2
    xr = np.zeros((n, m + 1), np.float32); xi = np.zeros((n, m + 1), np.float32)
    lib.oracle_rfft_f32(_fp(x), p, n, _fp(xr), _fp(xi))
    return xr + 1j * xi.astype(np.float64)


def irfft_f32(Y, p):
    lib = load()
    Y = np.asarray(Y).reshape(-1, p // 2 + 1)
    yr = np.ascontiguousarray(Y.real, dtype=np.float32)
    yi = np.ascontiguousarray(Y.imag, dtype=np.float32)
    n = Y.shape[0]
    y = np.zeros((n, p), np.float32)
    lib.oracle_irfft_f32(_fp(yr), _fp(yi), p, n, _fp(y))
    return y


def packed_pairs(n):
    """(i, j) arrays in packed upper-triangle order (j >= i)."""
    iu = np.triu_indices(n)
    return iu[0].astype(np.int32), iu[1].astype(np.int32)
