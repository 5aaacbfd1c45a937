"""CPU checks of the reversed-row column spectra the drop-ins use.

The device stores, for the caller's reversed row r_j, the column spectrum
conj(rfft(r_j)[k] * exp(+2 pi i q / P)) with q = k (L - 1) mod P, taking the
phase from the post-twiddle table post[k] = exp(-2 pi i k / P), k <= P/2:
conj(post[q]) for q <= P/2 and post[P - q] above.  The pair kernel then
inverse-transforms X_i * conj(column) and reads lag l at slot l mod P.

Pinned here, independently of that construction:
  * the table lookup equals exp(+2 pi i q / P) for every q (a swapped branch,
    a missing conjugate or an off-by-one index fails);
  * the whole construction reproduces the reference algorithm -- the literal
    rfft(events) * rfft(reversed), unnormalised irfft and corr_t[L + i] scan
    of tests/np_ref.py -- on events of uneven length, lags and values."""
import math

import numpy as np

from tests import np_ref


def twiddle_f32(k, n, sign):
    """exp(sign * 2 pi i k / n), FP64 with the angle reduced to an octant,
    rounded once to FP32 (the host table builder's construction)."""
    k %= n
    k8 = 8 * k
    octant, rem = divmod(k8, n)
    a = rem / n * (math.pi / 4)
    ca, sa, cb, sb = math.cos(a), math.sin(a), math.cos(math.pi / 4 - a), math.sin(math.pi / 4 - a)
    if rem == 0:
        ca, sa, cb, sb = 1.0, 0.0, math.sqrt(0.5), math.sqrt(0.5)
    c, s = [(ca, sa), (sb, cb), (-sa, ca), (-cb, sb), (-ca, -sa), (-sb, -cb), (sa, -ca), (cb, -sb)][octant]
    return complex(np.float32(c), np.float32(sign * s))


def phase_from_table(k, L, P):
    M = P // 2
    post = [twiddle_f32(x, P, -1) for x in range(M + 1)]
    q = (k * (L - 1)) % P
    return post[q].conjugate() if q <= M else post[P - q]


def test_phase_lookup_is_exp_plus():
    for P in (1024, 3072, 4096):
        for L in (1, 2, 700, P // 2, P - 1):
            for k in (0, 1, 5, P // 4, P // 2 - 1, P // 2):
                ph = phase_from_table(k, L, P)
                q = (k * (L - 1)) % P
                assert abs(ph - np.exp(2j * np.pi * q / P)) < 2e-7, (P, L, k)


def test_rotated_reversed_spectra_reproduce_reference_on_ragged_events():
    rng = np.random.default_rng(3)
    lens = [1000, 700, 1023, 512, 1024, 990, 1000]
    L = lens[0]
    P = 1 << L.bit_length()
    M = P // 2
    pad = np.zeros((len(lens), P), np.float32)
    rev = np.zeros((len(lens), P), np.float32)
    for i, n in enumerate(lens):
        x = rng.standard_normal(n).astype(np.float32)
        pad[i, :n] = x
        rev[i, :n] = x[::-1]
    S = 24  # L + S <= P
    ref = np_ref.reference_f64(pad, rev, L, S)
    X = np.fft.rfft(pad.astype(np.float64), axis=1)
    post = np.array([twiddle_f32(x, P, -1) for x in range(M + 1)])
    q = (np.arange(M + 1) * (L - 1)) % P
    ph = np.where(q <= M, np.conj(post[np.minimum(q, M)]), post[np.clip(P - q, 0, M)])
    col = np.conj(np.fft.rfft(rev.astype(np.float64), axis=1) * ph)
    col[:, 0] = col[:, 0].real
    col[:, M] = col[:, M].real
    norms = np.sqrt((pad[:, :L].astype(np.float64) ** 2).sum(1))
    lags = np.arange(-S + 1, S + 1)
    k = 0
    for i in range(len(lens)):
        for j in range(i, len(lens)):
            c = np.fft.irfft(X[i] * np.conj(col[j]), n=P) / (norms[i] * norms[j])
            w = c[lags % P]
            a, b = int(np.argmax(w)), int(np.argmin(w))
            assert abs(max(w[a], 0.0) - ref["pos"][k]) < 1e-6
            assert abs(min(w[b], 0.0) - ref["neg"][k]) < 1e-6
            if ref["gapp"][k] > 1e-6 and w[a] > 0:
                assert lags[a] == ref["lagp"][k]
            if ref["gapn"][k] > 1e-6 and w[b] < 0:
                assert lags[b] == ref["lagn"][k]
            k += 1
