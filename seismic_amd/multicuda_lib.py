"""Mirror of the reference binding Correlation_in_MultiCUDA/correlation_lib.py,
loading this repository's ``lib/correlation_multicuda.so``.  Every nGPUS runs
on the one MI355X (the reference clamps nGPUS to the devices present,
correlation_multicuda.cu:128-129)."""
from __future__ import annotations

import ctypes
import logging
import sys
import time

import numpy.ctypeslib as ctl

from ._lib import LIB_DIR
from .correlation_lib import c_float_p, c_int_p, prepare


def initialise_library():
    try:
        library = ctl.load_library("correlation_multicuda.so", LIB_DIR)
    except OSError:
        logging.error("Library has not been properly loaded")
        sys.exit(1)
    library.correlationCUDA.restype = None
    library.correlationCUDA.argtypes = [c_float_p, c_float_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                        c_float_p, c_int_p, c_float_p, c_int_p]
    library.seismic_last_error.restype = ctypes.c_char_p
    library.seismic_last_error.argtypes = []
    return library


def compute_correlation(events, shift, nGPUS, threshold):
    """Correlation_in_MultiCUDA/correlation_lib.py:31-101: returns
    (xcm_pos, xclags_pos, xcm_neg, xclags_neg, prep_seconds)."""
    start = time.time()
    library = initialise_library()
    n_events, max_len, max_pad, padded, padded_rev, outs = prepare(events)
    xcm_pos, xclags_pos, xcm_neg, xclags_neg = outs
    diff_time = time.time() - start
    library.correlationCUDA(padded.ctypes.data_as(c_float_p),
                            padded_rev.ctypes.data_as(c_float_p),
                            ctypes.c_int(n_events), ctypes.c_int(max_len), ctypes.c_int(shift),
                            ctypes.c_int(max_pad), ctypes.c_float(threshold), ctypes.c_int(nGPUS),
                            xcm_pos.ctypes.data_as(c_float_p), xclags_pos.ctypes.data_as(c_int_p),
                            xcm_neg.ctypes.data_as(c_float_p), xclags_neg.ctypes.data_as(c_int_p))
    err = library.seismic_last_error()
    if err:
        raise RuntimeError("correlationCUDA: " + err.decode())
    return xcm_pos, xclags_pos, xcm_neg, xclags_neg, diff_time


def initialiseRunTime(nGPUS):
    """Correlation_in_MultiCUDA/correlation_lib.py:103-108 (argtypes=[] plus an
    int argument, which ctypes forwards)."""
    library = initialise_library()
    library.initialiseCUDA.restype = None
    library.initialiseCUDA.argtypes = []
    library.initialiseCUDA(nGPUS)
