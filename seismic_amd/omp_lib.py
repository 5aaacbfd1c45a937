"""Mirror of the reference binding Correlation_in_OMP/correlation_lib.py.

Same function, argument meaning and return values (packed results plus the
two histograms), loading this repository's ``lib/correlation_c.so`` (the
OMP ABI computed on MI355X) instead of the cluster path the reference
hard-codes (correlation_lib.py:24).
"""
from __future__ import annotations

import ctypes
import logging
import sys

import numpy as np
import numpy.ctypeslib as ctl

from ._lib import LIB_DIR
from .correlation_lib import c_float_p, c_int_p, prepare


def initialize_library():
    """Correlation_in_OMP/correlation_lib.py:17-29: load or exit(1)."""
    try:
        library = ctl.load_library("correlation_c.so", LIB_DIR)
    except OSError:
        logging.error("Something bad is happening..")
        sys.exit(1)
    library.correlationCPP.restype = None
    library.correlationCPP.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_int, c_float_p,
                                       c_int_p, c_float_p, c_int_p, c_int_p, c_int_p]
    library.seismic_last_error.restype = ctypes.c_char_p
    library.seismic_last_error.argtypes = []
    return library


def compute_correlation(events, shift, num_threads, chunk_size, threshold, n_classes_hist):
    """Correlation_in_OMP/correlation_lib.py:31-113: returns (xcm_pos,
    xclags_pos, xcm_neg, xclags_neg, max_hist, min_hist)."""
    library = initialize_library()
    max_hist = np.zeros(n_classes_hist, dtype=np.int32)
    min_hist = np.zeros(n_classes_hist, dtype=np.int32)
    n_events, max_len, max_pad, padded, padded_rev, outs = prepare(events)
    xcm_pos, xclags_pos, xcm_neg, xclags_neg = outs
    library.correlationCPP(padded.ctypes.data_as(c_float_p), padded_rev.ctypes.data_as(c_float_p),
                           ctypes.c_int(n_events), ctypes.c_int(max_len), ctypes.c_int(shift),
                           ctypes.c_int(max_pad), ctypes.c_int(num_threads),
                           ctypes.c_int(chunk_size), ctypes.c_float(threshold),
                           ctypes.c_int(n_classes_hist),
                           xcm_pos.ctypes.data_as(c_float_p), xclags_pos.ctypes.data_as(c_int_p),
                           xcm_neg.ctypes.data_as(c_float_p), xclags_neg.ctypes.data_as(c_int_p),
                           max_hist.ctypes.data_as(c_int_p), min_hist.ctypes.data_as(c_int_p))
    err = library.seismic_last_error()
    if err:
        raise RuntimeError(err.decode())
    return xcm_pos, xclags_pos, xcm_neg, xclags_neg, max_hist, min_hist
