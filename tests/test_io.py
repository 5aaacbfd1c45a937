"""Caller formats (CPU): events loaded like the reference drivers, results
written as 1 x M CSR .npz files under the reference's names, packed ->
dense expansion."""
import os

import numpy as np
import pytest
from scipy.sparse import load_npz

from seismic_amd import io


def test_load_events_plain_and_object(tmp_path):
    ev = np.arange(12, dtype=np.float32).reshape(3, 4)
    p = tmp_path / "ev.npy"
    np.save(p, ev)
    assert np.array_equal(io.load_events(str(p)), ev)
    obj = np.empty(2, dtype=object)
    obj[0] = np.ones(5, np.float32)
    obj[1] = np.zeros(5, np.float32)
    q = tmp_path / "obj.npy"
    np.save(q, obj, allow_pickle=True)
    with pytest.raises(ValueError, match="allow_pickle=True"):
        io.load_events(str(q))
    got = io.load_events(str(q), allow_pickle=True)   # our own file: trusted
    assert got.dtype == object and np.array_equal(got[0], obj[0])


def test_save_results_reference_layout(tmp_path):
    rng = np.random.default_rng(0)
    n = 7
    m = n * (n + 1) // 2
    vp = rng.random(m).astype(np.float32)
    vp[::3] = 0.0                                   # sub-threshold entries stay implicit zeros
    lp = rng.integers(-50, 50, m).astype(np.int32)
    vn = -rng.random(m).astype(np.float32)
    ln = rng.integers(-50, 50, m).astype(np.int32)
    hp = rng.integers(0, 9, 20).astype(np.int32)
    hn = rng.integers(0, 9, 20).astype(np.int32)
    paths = io.save_results(str(tmp_path), vp, lp, vn, ln, hp, hn)
    assert sorted(os.path.basename(p) for p in paths) == sorted(io.RESULT_FILES.values())
    x = load_npz(str(tmp_path / "out_xcm_pos.npz"))
    assert x.shape == (1, m) and x.dtype == np.float32 and x.nnz == int((vp != 0).sum())
    back = io.load_results(str(tmp_path))
    for k, a in (("xcm_pos", vp), ("xclags_pos", lp), ("xcm_neg", vn), ("xclags_neg", ln),
                 ("max_hist", hp), ("min_hist", hn)):
        assert np.array_equal(back[k], a), k


def test_packed_to_dense_mirror():
    n = 5
    m = n * (n + 1) // 2
    v = np.arange(m, dtype=np.float32)
    g = np.arange(m, dtype=np.int32) - 7
    V, G = io.packed_to_dense(n, v, g)
    for i in range(n):
        for j in range(i, n):
            k = i * n - i * (i - 1) // 2 + (j - i)
            assert V[i, j] == v[k] and G[i, j] == g[k]
            assert V[j, i] == v[k] and G[j, i] == -g[k] or i == j
