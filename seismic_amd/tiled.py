"""Tiled N x N runs of the lag-window forms over the extended C ABI
(sxc_tiled_*): upper-triangle blocks streamed to the host as float16
coefficients and int16 lags, plus a helper that assembles the full N x N
matrices (e.g. into numpy memmaps)."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import SxcBlockInfo, SxcStats, last_error, load_engine
from .engine import FORM_LAG_BF16X3, _desc


class TiledRun:
    def __init__(self, n_events, n_channels, trace_len, lag_lo, lag_hi, form=FORM_LAG_BF16X3,
                 block=8192, want_neg=False, threshold=0.0, tie_tol=0.0, demean=False):
        self.lib = load_engine()
        self.n_events = int(n_events)
        self.n_channels = int(n_channels)
        self.want_neg = bool(want_neg)
        self.block = int(block)
        self.edge = min(self.block, self.n_events)
        self.desc = _desc(n_events, n_channels, trace_len, 0, lag_lo, lag_hi, threshold, tie_tol,
                          demean, form)
        h = ctypes.c_void_p()
        rc = self.lib.sxc_tiled_create(ctypes.byref(self.desc), int(block), 1 if want_neg else 0,
                                       ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"sxc_tiled_create failed ({rc}): {last_error(self.lib)}")
        self.handle = h

    def start(self, traces, in_len=None):
        """traces: DeviceArray float32 [N, C, stride] (or [N, stride] for C=1)."""
        stride = traces.shape[-1]
        if int(np.prod(traces.shape)) != self.n_events * self.n_channels * stride:
            raise ValueError("traces shape does not match the run")
        in_len = stride if in_len is None else int(in_len)
        rc = self.lib.sxc_tiled_start(self.handle, traces.ptr, stride, in_len)
        if rc != 0:
            raise RuntimeError(f"sxc_tiled_start failed ({rc}): {last_error(self.lib)}")

    def next_block(self, out=None):
        """Returns (info dict, coef_pos float16, lag_pos int16[, coef_neg,
        lag_neg]) of the next block, or None when the run is complete.
        `out`: optional arrays of at least [edge, edge] to fill (views of
        them are returned), edge = min(block, n_events)."""
        info = SxcBlockInfo()
        k = 4 if self.want_neg else 2
        if out is None:
            e = self.edge
            out = [np.empty((e, e), np.float16 if q % 2 == 0 else np.int16) for q in range(k)]
        # the library writes up to edge rows of row stride ld into each array
        if len(out) < k or any(
                o.ndim != 2 or not o.flags.c_contiguous or o.shape[0] < self.edge
                or o.shape[1] != out[0].shape[1] or o.shape[1] < self.edge
                or o.dtype != (np.float16 if q % 2 == 0 else np.int16)
                for q, o in enumerate(out[:k])):
            raise ValueError(f"out must hold {k} C-contiguous arrays of at least [{self.edge}, {self.edge}], "
                             "float16 coefficients and int16 lags alternating, one row stride")
        out = list(out[:k])
        ld = out[0].shape[1]
        ptrs = [o.ctypes.data_as(ctypes.c_void_p) for o in out]
        if not self.want_neg:
            ptrs += [None, None]
        rc = self.lib.sxc_tiled_next(self.handle, ctypes.byref(info), ptrs[0], ptrs[1], ptrs[2],
                                     ptrs[3], ld)
        if rc != 0:
            raise RuntimeError(f"sxc_tiled_next failed ({rc}): {last_error(self.lib)}")
        if info.index < 0:
            return None
        d = {f: getattr(info, f) for f, _ in SxcBlockInfo._fields_}
        r, c = info.rows, info.cols
        return (d,) + tuple(o[:r, :c] for o in out)

    def blocks(self):
        while True:
            res = self.next_block()
            if res is None:
                return
            yield res

    def stats(self):
        s = SxcStats()
        rc = self.lib.sxc_tiled_stats(self.handle, ctypes.byref(s))
        if rc != 0:
            raise RuntimeError(f"sxc_tiled_stats failed ({rc}): {last_error(self.lib)}")
        return {"pairs": s.pairs, "near_ties": s.near_ties, "tie_overflow": s.tie_overflow}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.sxc_tiled_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble(run: TiledRun, coef, lag):
    """Writes every block of a started run into the N x N arrays coef
    (float16) and lag (int16), the blocks below the diagonal as transposes
    with lags negated.  Returns the number of blocks."""
    n = 0
    for res in run.blocks():
        info, c, l = res[0], res[1], res[2]
        r0, c0, rows, cols = info["row0"], info["col0"], info["rows"], info["cols"]
        coef[r0:r0 + rows, c0:c0 + cols] = c
        lag[r0:r0 + rows, c0:c0 + cols] = l
        if r0 != c0:
            coef[c0:c0 + cols, r0:r0 + rows] = c.T
            lag[c0:c0 + cols, r0:r0 + rows] = -l.T
        n += 1
    return n
