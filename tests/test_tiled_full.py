"""The largest single-GPU configuration end to end (opt in: it runs ~90 s):
N=65536 events, 1 channel, 4096 samples, lags -1024..1024, bf16x3, 10%
zero-filled shifted duplicates, log-uniform SNR [0.3, 30], float16/int16
blocks of 8192 x 8192 streamed to host.  Checks every duplicate's lag and
coefficient and 2000 seeded pairs against the FP64 oracle.

    SEISMIC_AMD_FULL=1 python -m pytest tests/test_tiled_full.py -m gpu -s
"""
import os
import time

import numpy as np
import pytest

from seismic_amd import synth, tiled

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.environ.get("SEISMIC_AMD_FULL"),
                                 reason="full N=65536 run; set SEISMIC_AMD_FULL=1")]


def test_full_n65536(oracle):
    from seismic_amd.device import DeviceArray
    N, L, W, B = 65536, 4096, 1024, 8192
    ev, meta = synth.make_events(N, 1, L, seed=65536, snr=(0.3, 30.0), log_snr=True, dup_frac=0.1)
    dups = {}
    for s, d, sh in meta["dups"]:
        i, j = min(s, d), max(s, d)
        dups.setdefault((i // B, j // B), []).append((i, j, -sh if i == s else sh, s, d))
    rng = np.random.default_rng(5)
    k = 2000
    pi = rng.integers(0, N, k)
    pj = rng.integers(0, N, k)
    pi, pj = np.minimum(pi, pj), np.maximum(pi, pj)
    samp = {}
    for q in range(k):
        samp.setdefault((pi[q] // B, pj[q] // B), []).append(q)
    got_v = np.zeros(k, np.float32)
    got_l = np.zeros(k, np.int32)
    e64 = ev[:, 0].astype(np.float64)
    tr = DeviceArray.from_numpy(ev)
    run = tiled.TiledRun(N, 1, L, -W, W, block=B)
    bufs = [np.empty((B, B), np.float16), np.empty((B, B), np.int16)]
    t0 = time.time()
    run.start(tr)
    nb = ndup = bad = 0
    worst = 0.0
    while (res := run.next_block(bufs)) is not None:
        info, c, lg = res[0], res[1], res[2]
        I, J = info["row0"] // B, info["col0"] // B
        for (i, j, want, s, d) in dups.get((I, J), []):
            r, cc = i - info["row0"], j - info["col0"]
            ndup += 1
            bad += int(lg[r, cc]) != want
            worst = max(worst, abs(float(c[r, cc]) - np.linalg.norm(e64[d]) / np.linalg.norm(e64[s])))
        for q in samp.get((I, J), []):
            got_v[q] = float(c[pi[q] - info["row0"], pj[q] - info["col0"]])
            got_l[q] = int(lg[pi[q] - info["row0"], pj[q] - info["col0"]])
        nb += 1
        print(f"block {nb}/{info['n_blocks']} ({I},{J}) at {time.time() - t0:.1f}s", flush=True)
    wall = time.time() - t0
    st = run.stats()
    run.close()
    tr.free()
    print(f"N={N}: {wall:.1f}s, {N * (N + 1) / 2 / wall / 1e6:.2f} M pairs/s, near ties {st['near_ties']}")
    assert nb == 36 and st["tie_overflow"] == 0 and st["pairs"] == N * (N + 1) // 2
    assert ndup == len(meta["dups"]) and bad == 0
    assert worst <= 5e-4                      # float16 spacing near 1 is 4.9e-4
    ref = oracle.pairs_f64(ev, pi.astype(np.int32), pj.astype(np.int32), -W, W, circular=False)
    err = np.abs(got_v - ref["pos"]).max()
    keep = ref["pos"] > 1e-3
    print(f"FP64 oracle, {k} pairs: max |dcoef| {err:.2e}, lag agreement "
          f"{(got_l[keep] == ref['lag_pos'][keep]).mean():.5f}")
    assert err <= 1e-3
    sure = keep & (ref["gap_pos"] >= 4e-5)
    assert np.array_equal(got_l[sure], ref["lag_pos"][sure])
