"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from this
repository's FP64 code): the CPU oracle on the CPU, and the GPU forms through
the C ABI."""
import glob
import os

import numpy as np
import pytest

from seismic_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
CIRC = ["ncc_full_n8_l1024.npz", "ncc_window_n10_l1500_s300.npz", "ncc_known_answers.npz"]


def _load(name):
    return dict(np.load(os.path.join(HERE, "golden", name), allow_pickle=False))


def _check(vp, lp, vn, ln, g, coef_tol=1e-5, gap=1e-5):
    assert np.abs(vp - g["pos"]).max() <= coef_tol
    assert np.abs(vn - g["neg"]).max() <= coef_tol
    assert np.all((lp == g["lag_pos"]) | (g["gap_pos"] < gap))
    assert np.all((ln == g["lag_neg"]) | (g["gap_neg"] < gap))


def test_fixtures_present():
    assert len(glob.glob(os.path.join(HERE, "golden", "*.npz"))) >= 5


@pytest.mark.parametrize("name", CIRC)
def test_oracle_against_golden(oracle, name):
    g = _load(name)
    padded, rev, L, P = synth.pad_like_binding(g["events"])
    vp, lp, vn, ln, _, _ = oracle.correlation_cpp(padded, rev, L, int(g["shift"]), float(g["threshold"]))
    _check(vp, lp, vn, ln, g)


def test_direct_oracle_against_golden(oracle):
    g = _load("ncc_3ch_n6_l512_p1024.npz")
    pi, pj = oracle.packed_pairs(6)
    r = oracle.pairs_f64(g["events"], pi, pj, int(g["lag_lo"]), int(g["lag_hi"]), P=int(g["P"]), circular=True)
    assert np.abs(r["pos"] - g["pos"]).max() < 1e-12
    assert np.array_equal(r["lag_pos"], g["lag_pos"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CIRC)
def test_dropin_against_golden(name):
    from seismic_amd import correlation_lib
    g = _load(name)
    vp, lp, vn, ln = correlation_lib.compute_correlation(g["events"], int(g["shift"]), 1, float(g["threshold"]))
    _check(vp, lp, vn, ln, g)


@pytest.mark.gpu
def test_multichannel_against_golden():
    from seismic_amd import engine
    g = _load("ncc_3ch_n6_l512_p1024.npz")
    vp, lp, vn, ln = engine.allpairs_host(g["events"], int(g["P"]), int(g["lag_lo"]), int(g["lag_hi"]))
    _check(vp, lp, vn, ln, g)


@pytest.mark.gpu
@pytest.mark.parametrize("form,tol,gap", [(1, 1e-5, 4e-5), (2, 2e-3, 1e-3)])
def test_lag_forms_against_golden(form, tol, gap):
    from seismic_amd import engine
    g = _load("lag_2ch_n8_l600_w100.npz")
    vp, lp, vn, ln = engine.allpairs_host(g["events"], 0, int(g["lag_lo"]), int(g["lag_hi"]),
                                          threshold=float(g["threshold"]), form=form)
    assert np.abs(vp - g["pos"]).max() <= tol and np.abs(vn - g["neg"]).max() <= tol
    sure = (g["gap_pos"] >= gap) & (g["pos"] > 0)
    assert np.array_equal(lp[sure], g["lag_pos"][sure])
