"""Tiled N x N runs: argument checks on the CPU; on the GPU, equality with
the packed-triangle execute of the same form (values rounded to float16),
mirrored blocks, ragged edges, and exact lags of time-shifted duplicates
at the largest-run trace shape (L=4096, window +-1024)."""
import numpy as np
import pytest

from seismic_amd import engine, synth, tiled


def _packed_full(n, vp, lp):
    """Packed upper triangle -> N x N (upper) index arrays."""
    iu = np.triu_indices(n)
    V = np.zeros((n, n), np.float32)
    Lg = np.zeros((n, n), np.int32)
    V[iu] = vp
    Lg[iu] = lp
    return V, Lg, iu


def test_tiled_argument_checks():
    with pytest.raises(RuntimeError, match="multiple of 32"):
        tiled.TiledRun(100, 1, 256, -10, 10, block=100)
    with pytest.raises(RuntimeError, match="lag-window forms"):
        tiled.TiledRun(100, 1, 256, -10, 10, form=engine.FORM_FFT_F32, block=64)


def _tiled_all(ev, lo, hi, form, block, want_neg):
    from seismic_amd.device import DeviceArray
    n, C, L = ev.shape
    run = tiled.TiledRun(n, C, L, lo, hi, form=form, block=block, want_neg=want_neg)
    tr = DeviceArray.from_numpy(ev)
    run.start(tr)
    coef = np.full((n, n), np.nan, np.float16)
    lag = np.zeros((n, n), np.int16)
    cn = np.full((n, n), np.nan, np.float16)
    ln = np.zeros((n, n), np.int16)
    nb = 0
    for res in run.blocks():
        info = res[0]
        r0, c0, rows, cols = info["row0"], info["col0"], info["rows"], info["cols"]
        coef[r0:r0 + rows, c0:c0 + cols] = res[1]
        lag[r0:r0 + rows, c0:c0 + cols] = res[2]
        if want_neg:
            cn[r0:r0 + rows, c0:c0 + cols] = res[3]
            ln[r0:r0 + rows, c0:c0 + cols] = res[4]
        nb += 1
    st = run.stats()
    run.close()
    tr.free()
    return coef, lag, cn, ln, nb, st


@pytest.mark.gpu
@pytest.mark.parametrize("n,C,L,lo,hi,form,block,neg", [
    (300, 1, 512, -100, 100, engine.FORM_LAG_BF16X3, 128, False),
    (100, 2, 300, -50, 80, engine.FORM_LAG_FP16, 64, True),
])
def test_tiled_equals_packed(n, C, L, lo, hi, form, block, neg):
    ev, _ = synth.make_events(n, C, L, seed=n + L)
    print(f"case tiled n={n} C={C} L={L} [{lo},{hi}] form={form} block={block}", flush=True)
    coef, lag, cn, ln, nb, st = _tiled_all(ev, lo, hi, form, block, neg)
    nbs = (n + block - 1) // block
    assert nb == nbs * (nbs + 1) // 2
    assert st["pairs"] == n * (n + 1) // 2 and st["tie_overflow"] == 0
    vp, lp, vn, lnp = engine.allpairs_host(ev, 0, lo, hi, form=form)
    V, Lg, iu = _packed_full(n, vp, lp)
    # upper triangle (incl. diagonal blocks' upper halves): float16 of the packed values, same lags
    assert np.array_equal(coef[iu], V[iu].astype(np.float16))
    assert np.array_equal(lag[iu].astype(np.int32), Lg[iu])
    # diagonal blocks hold the mirror below the diagonal
    for b in range(nbs):
        s = slice(b * block, min(n, (b + 1) * block))
        cb, lb = coef[s, s], lag[s, s]
        low = np.tril_indices(cb.shape[0], -1)
        assert np.array_equal(cb[low], cb.T[low])
        assert np.array_equal(lb[low], -lb.T[low])
    if neg:
        Vn, Ln, _ = _packed_full(n, vn, lnp)
        assert np.array_equal(cn[iu], Vn[iu].astype(np.float16))
        assert np.array_equal(ln[iu].astype(np.int32), Ln[iu])


@pytest.mark.gpu
def test_tiled_run_state():
    """next() before a start and after a rejected start reports that no run
    is in flight; a second start over the same run delivers the same blocks."""
    from seismic_amd.device import DeviceArray
    n, L = 96, 256
    ev, _ = synth.make_events(n, 1, L, seed=5)
    print(f"case tiled run state n={n} L={L}", flush=True)
    run = tiled.TiledRun(n, 1, L, -20, 20, form=engine.FORM_LAG_BF16X3, block=64)
    tr = DeviceArray.from_numpy(ev)
    try:
        with pytest.raises(RuntimeError, match="no run in flight"):
            run.next_block()
        with pytest.raises(RuntimeError, match="in_len"):
            run.start(tr, in_len=L + 1)
        with pytest.raises(RuntimeError, match="no run in flight"):
            run.next_block()
        run.start(tr)
        small = [np.empty((32, 64), np.float16), np.empty((32, 64), np.int16)]
        with pytest.raises(ValueError, match="out must hold"):
            run.next_block(small)              # rows < edge: rejected before the call
        with pytest.raises(ValueError, match="out must hold"):
            run.next_block([np.empty((64, 64), np.float32), np.empty((64, 64), np.int16)])
        runs = []
        for _ in range(2):
            run.start(tr)
            runs.append([(r[0]["index"], r[1].copy(), r[2].copy()) for r in run.blocks()])
            assert run.next_block() is None   # complete: stays complete
        assert len(runs[0]) == 3 and len(runs[1]) == 3
        for a, b in zip(runs[0], runs[1]):
            assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    finally:
        run.close()
        tr.free()


@pytest.mark.gpu
def test_tiled_l4096_w1024_duplicates(oracle):
    """L=4096, window +-1024, bf16x3, log-uniform SNR in [0.3, 30] and 10%
    zero-filled time-shifted duplicates, N=640 in blocks of 256 (ragged)."""
    n, L, w = 640, 4096, 1024
    ev, meta = synth.make_events(n, 1, L, seed=65536, snr=(0.3, 30.0), log_snr=True, dup_frac=0.1)
    print(f"case tiled L=4096 +-1024 n={n} dups={len(meta['dups'])}", flush=True)
    coef, lag, _, _, nb, st = _tiled_all(ev, -w, w, engine.FORM_LAG_BF16X3, 256, False)
    assert nb == 6 and st["tie_overflow"] == 0
    iu = np.triu_indices(n)
    full_coef = coef.astype(np.float32)
    full_lag = lag.astype(np.int32)
    # mirror the upper blocks to the full matrix
    low = np.tril_indices(n, -1)
    full_coef[low] = full_coef.T[low]
    full_lag[low] = -full_lag.T[low]
    bad = 0
    for s, d, sh in meta["dups"]:
        i, j = min(s, d), max(s, d)
        want = -sh if i == s else sh   # row i delayed relative to column j by lag
        e64 = ev.astype(np.float64)
        expect = np.linalg.norm(e64[d]) / np.linalg.norm(e64[s])
        if full_lag[i, j] != want:
            bad += 1
        assert abs(full_coef[i, j] - expect) <= 1e-3, (s, d, sh)
        assert full_lag[j, i] == -want
    assert bad == 0, f"{bad} duplicate lags differ"
    # FP64 oracle on seeded sample pairs
    rng = np.random.default_rng(7)
    k = 1500
    pi = rng.integers(0, n, size=k).astype(np.int32)
    pj = rng.integers(0, n, size=k).astype(np.int32)
    pi, pj = np.minimum(pi, pj), np.maximum(pi, pj)
    ref = oracle.pairs_f64(ev, pi, pj, -w, w, circular=False)
    got = full_coef[pi, pj]
    err = np.abs(got - ref["pos"]).max()
    sure = ref["gap_pos"] >= 4e-5
    agree = (full_lag[pi, pj] == ref["lag_pos"]).mean()
    print(f"tiled L=4096 +-1024: max |dcoef| vs FP64 {err:.3e} (float16 output), lag agreement {agree:.5f}")
    assert err <= 1e-3
    assert np.array_equal(full_lag[pi, pj][sure], ref["lag_pos"][sure])
