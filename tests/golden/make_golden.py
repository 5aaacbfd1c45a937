"""Writes the golden fixtures in tests/golden/*.npz.

The reference ships no data, outputs or assertions for this path, so the
expected values come from this repository's own FP64 code: tests/np_ref.py
(numpy FP64 rendering of Correlation_in_OMP/correlation_c.cpp) for the
circular cases and oracle.pairs_f64 (FP64 direct sums) for the multi-channel
and linear-window cases.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from seismic_amd import synth  # noqa: E402
from tests import np_ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def circular_case(name, events, shift, thr):
    padded, rev, L, P = synth.pad_like_binding(events)
    ref = np_ref.reference_f64(padded, rev, L, shift, threshold=thr)
    np.savez_compressed(os.path.join(OUT, name), events=np.asarray(events, np.float32), shift=shift,
                        threshold=thr, L=L, P=P, pos=ref["pos"], lag_pos=ref["lagp"].astype(np.int32),
                        neg=ref["neg"], lag_neg=ref["lagn"].astype(np.int32), gap_pos=ref["gapp"],
                        gap_neg=ref["gapn"])


def direct_case(name, events, lo, hi, circular, P, thr):
    n = events.shape[0]
    pi, pj = oracle.packed_pairs(n)
    ref = oracle.pairs_f64(events, pi, pj, lo, hi, P=P, circular=circular)
    keep_p = ref["pos"] > thr
    keep_n = ref["neg"] < -thr
    np.savez_compressed(os.path.join(OUT, name), events=events.astype(np.float32), lag_lo=lo, lag_hi=hi,
                        circular=int(circular), P=P, threshold=thr,
                        pos=np.where(keep_p, ref["pos"], 0.0), lag_pos=np.where(keep_p, ref["lag_pos"], 0).astype(np.int32),
                        neg=np.where(keep_n, ref["neg"], 0.0), lag_neg=np.where(keep_n, ref["lag_neg"], 0).astype(np.int32),
                        gap_pos=ref["gap_pos"], gap_neg=ref["gap_neg"])


def main():
    ev, _ = synth.make_events(8, 1, 1024, seed=1001)
    circular_case("ncc_full_n8_l1024.npz", ev[:, 0], 1024, 0.0)
    ev, _ = synth.make_events(10, 1, 1500, seed=1002)
    circular_case("ncc_window_n10_l1500_s300.npz", ev[:, 0], 300, 0.2)
    rng = np.random.default_rng(1003)
    base = np.zeros(1024, np.float32)
    base[250:750] = rng.standard_normal(500).astype(np.float32)
    known = np.stack([base, np.roll(base, 23), -base, np.zeros(1024, np.float32), np.roll(base, -101)])
    circular_case("ncc_known_answers.npz", known, 512, 0.0)
    ev, _ = synth.make_events(6, 3, 512, seed=1004)
    direct_case("ncc_3ch_n6_l512_p1024.npz", ev, -511, 512, True, 1024, 0.0)
    ev, _ = synth.make_events(8, 2, 600, seed=1005)
    direct_case("lag_2ch_n8_l600_w100.npz", ev, -100, 120, False, 0, 0.1)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
