"""Input and output formats of the reference's driver scripts.

* Events: ``np.load(path, fix_imports=True, encoding='latin1')`` of an
  ``.npy`` holding either a float array [n_events, L] or an object array of
  traces (ndarrays or objects with a ``.data`` ndarray, e.g. obspy Traces)
  (Correlation_in_CUDA/test_cuda.py:22, Correlation_in_OMP/test_c.py:27).
  Object arrays are pickles: they load only with ``allow_pickle=True``, which
  executes code from the file — pass it for files you trust; the classes the
  pickle names (obspy's) must be importable.
* Results: each packed array (and each histogram) as a 1 x M
  ``scipy.sparse.csr_matrix`` written by ``save_npz(..., compressed=True)``
  under the reference's file names (Correlation_in_OMP/test_c.py:43-55,
  Correlation_in_CUDA/test_cuda.py:39-47).
"""
from __future__ import annotations

import os

import numpy as np

RESULT_FILES = {"xcm_pos": "out_xcm_pos.npz", "xclags_pos": "out_xcl_pos.npz",
                "xcm_neg": "out_xcm_neg.npz", "xclags_neg": "out_xcl_neg.npz",
                "max_hist": "out_xmax_hist.npz", "min_hist": "out_xmin_hist.npz"}
_DTYPES = {"xcm_pos": np.float32, "xclags_pos": np.int32, "xcm_neg": np.float32,
           "xclags_neg": np.int32, "max_hist": np.int32, "min_hist": np.int32}


def load_events(path: str, allow_pickle: bool = False):
    """Events as the reference drivers load them (see module docstring)."""
    try:
        return np.load(path, fix_imports=True, encoding="latin1", allow_pickle=allow_pickle)
    except ValueError as e:
        if not allow_pickle and "allow_pickle" in str(e):
            raise ValueError(f"{path} holds pickled objects (e.g. obspy Traces); load it with "
                             "allow_pickle=True only if you trust the file") from e
        raise


def save_results(directory: str, xcm_pos, xclags_pos, xcm_neg, xclags_neg, max_hist=None,
                 min_hist=None) -> list:
    """Writes the packed results (and histograms, when given) as 1 x M CSR
    matrices with the reference's names; returns the paths written."""
    from scipy.sparse import csr_matrix, save_npz
    arrays = {"xcm_pos": xcm_pos, "xclags_pos": xclags_pos, "xcm_neg": xcm_neg,
              "xclags_neg": xclags_neg, "max_hist": max_hist, "min_hist": min_hist}
    os.makedirs(directory, exist_ok=True)
    paths = []
    for key, arr in arrays.items():
        if arr is None:
            continue
        path = os.path.join(directory, RESULT_FILES[key])
        save_npz(path, csr_matrix(np.asarray(arr), dtype=_DTYPES[key]), compressed=True)
        paths.append(path)
    return paths


def load_results(directory: str) -> dict:
    """Reads back whatever save_results (or the reference drivers) wrote, as
    flat arrays."""
    from scipy.sparse import load_npz
    out = {}
    for key, name in RESULT_FILES.items():
        path = os.path.join(directory, name)
        if os.path.exists(path):
            out[key] = np.asarray(load_npz(path).toarray()).ravel().astype(_DTYPES[key])
    return out


def packed_to_dense(n_events: int, values, lags):
    """Packed upper triangle (index i*N - i(i-1)/2 + (j-i)) -> N x N arrays,
    mirrored below the diagonal with the lag negated.  The mirror is pair
    (j, i) over the mirrored window: exact for the linear lag forms with a
    symmetric window; for the reference window -S+1..S the mirrored lags span
    -S..S-1."""
    v = np.asarray(values)
    g = np.asarray(lags)
    if v.shape != (n_events * (n_events + 1) // 2,):
        raise ValueError("values do not hold n_events*(n_events+1)/2 entries")
    V = np.zeros((n_events, n_events), v.dtype)
    G = np.zeros((n_events, n_events), g.dtype)
    iu = np.triu_indices(n_events)
    V[iu] = v
    G[iu] = g
    il = np.tril_indices(n_events, -1)
    V[il] = V.T[il]
    G[il] = -G.T[il]
    return V, G
