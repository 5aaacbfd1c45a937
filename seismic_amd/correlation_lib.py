"""Mirror of the reference binding Correlation_in_CUDA/correlation_lib.py.

Same functions, argument meaning and behaviour, loading this repository's
``lib/correlation_cuda.so`` (MI355X) instead of the cluster path the reference
hard-codes (correlation_lib.py:24).
"""
from __future__ import annotations

import ctypes
import logging
import sys

import numpy as np
import numpy.ctypeslib as ctl

from ._lib import LIB_DIR

c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)


def initialise_library():
    """Correlation_in_CUDA/correlation_lib.py:17-29: load or exit(1)."""
    try:
        library = ctl.load_library("correlation_cuda.so", LIB_DIR)
    except OSError:
        logging.error("Library has not been properly loaded")
        sys.exit(1)
    library.correlationCUDA.restype = None
    library.correlationCUDA.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_float, c_float_p, c_int_p, c_float_p, c_int_p]
    library.initialiseCUDA.restype = None
    library.initialiseCUDA.argtypes = []
    library.seismic_last_error.restype = ctypes.c_char_p
    library.seismic_last_error.argtypes = []
    return library


def _trace(ev):
    return np.asarray(ev.data if hasattr(ev, "data") else ev)


def prepare(events):
    """Padding exactly as correlation_lib.py:35-72: L from event 0, P = 1 << L.bit_length(),
    zero-padded rows and flipped rows, zeroed packed outputs of N(N+1)/2."""
    n_events = events.shape[0] if hasattr(events, "shape") else len(events)
    max_len = 0
    for _ in range(n_events):
        # the reference reads event 0's length on every iteration (correlation_lib.py:45)
        if _trace(events[0]).shape[0] > max_len:
            max_len = _trace(events[0]).shape[0]
    max_pad = 1 << max_len.bit_length()
    padded = np.zeros((n_events, max_pad), dtype=np.float32)
    padded_rev = np.zeros((n_events, max_pad), dtype=np.float32)
    for i in range(n_events):
        x = _trace(events[i])
        padded[i, :] = np.pad(x, (0, max_pad - x.shape[0]), "constant").astype(float)
        padded_rev[i, :] = np.pad(np.flip(x, 0), (0, max_pad - x.shape[0]), "constant").astype(float)
    n_elements = int(n_events * (n_events + 1) / 2)
    outs = (np.zeros(n_elements, dtype=np.float32), np.zeros(n_elements, dtype=np.int32),
            np.zeros(n_elements, dtype=np.float32), np.zeros(n_elements, dtype=np.int32))
    return n_events, max_len, max_pad, padded, padded_rev, outs


def compute_correlation(events, shift, num_threads, threshold):
    """Correlation_in_CUDA/correlation_lib.py:31-99: returns the packed
    (xcm_pos, xclags_pos, xcm_neg, xclags_neg) upper triangles."""
    library = initialise_library()
    n_events, max_len, max_pad, padded, padded_rev, outs = prepare(events)
    xcm_pos, xclags_pos, xcm_neg, xclags_neg = outs
    library.correlationCUDA(padded.ctypes.data_as(c_float_p),
                            padded_rev.ctypes.data_as(c_float_p),
                            ctypes.c_int(n_events), ctypes.c_int(max_len), ctypes.c_int(shift),
                            ctypes.c_int(max_pad), ctypes.c_int(num_threads),
                            ctypes.c_float(threshold),
                            xcm_pos.ctypes.data_as(c_float_p), xclags_pos.ctypes.data_as(c_int_p),
                            xcm_neg.ctypes.data_as(c_float_p), xclags_neg.ctypes.data_as(c_int_p))
    err = library.seismic_last_error()
    if err:
        raise RuntimeError("correlationCUDA: " + err.decode())
    return xcm_pos, xclags_pos, xcm_neg, xclags_neg


def initialiseRunTime():
    """Correlation_in_CUDA/correlation_lib.py:101-106."""
    library = initialise_library()
    library.initialiseCUDA()
