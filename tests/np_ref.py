"""Independent FP64 numpy rendering of the reference algorithm, used to pin
the C oracle (not a product path).

It follows Correlation_in_OMP/correlation_c.cpp literally: X_i = rfft of the
padded event, R_j = rfft of the padded reversed event (:32-49), product
X_i * R_j (:188-193), unnormalised c2r (:63), scan corr_t[L+i] for
i in [-S, S) with lag i+1, strict > / < from 0 (first occurrence) (:102-111),
values / (P * n_i * n_j) (:115, :127), all in FP64 with numpy's FFT.
"""
import numpy as np


def norms_f64(padded, L):
    x = padded[:, :L].astype(np.float64)
    return np.sqrt(np.sum(x * x, axis=1))


def scan_window(w, S):
    """w: [..., 2S] window values for i = -S..S-1; returns pos, lagp, neg, lagn
    with the reference's strict rule starting from (0, lag 0)."""
    ap = np.argmax(w, axis=-1)
    mp = np.take_along_axis(w, ap[..., None], -1)[..., 0]
    an = np.argmin(w, axis=-1)
    mn = np.take_along_axis(w, an[..., None], -1)[..., 0]
    pos = np.where(mp > 0, mp, 0.0)
    lagp = np.where(mp > 0, ap - S + 1, 0)
    neg = np.where(mn < 0, mn, 0.0)
    lagn = np.where(mn < 0, an - S + 1, 0)
    return pos, lagp, neg, lagn


def second_gap(w, S):
    """FP64 gap between the winning value and the best other window value
    (pos side and neg side), to classify near ties."""
    ap = np.argmax(w, axis=-1)
    an = np.argmin(w, axis=-1)
    wp = w.copy()
    np.put_along_axis(wp, ap[..., None], -np.inf, -1)
    wn = w.copy()
    np.put_along_axis(wn, an[..., None], np.inf, -1)
    mp = np.take_along_axis(w, ap[..., None], -1)[..., 0]
    mn = np.take_along_axis(w, an[..., None], -1)[..., 0]
    return mp - wp.max(axis=-1), wn.min(axis=-1) - mn


def reference_f64(padded, padded_rev, L, S, threshold=0.0):
    """Packed FP64 results of the reference algorithm plus near-tie gaps."""
    n, P = padded.shape
    F = np.fft.rfft(padded.astype(np.float64), axis=1)
    Rv = np.fft.rfft(padded_rev.astype(np.float64), axis=1)
    nr = norms_f64(padded, L)
    out = {k: [] for k in ("pos", "lagp", "neg", "lagn", "gapp", "gapn", "rawpos", "rawneg")}
    for i in range(n):
        ct = np.fft.irfft(F[i][None, :] * Rv[i:], n=P, axis=1) * P
        with np.errstate(invalid="ignore", divide="ignore"):
            # a zero-energy event gives NaN here; the scan then yields (0, 0)
            w = ct[:, L - S:L + S] / (P * nr[i] * nr[i:, None])
        pos, lagp, neg, lagn = scan_window(w, S)
        gp, gn = second_gap(w, S)
        keep_p = pos > threshold
        keep_n = neg < -threshold
        out["rawpos"].append(pos)
        out["rawneg"].append(neg)
        out["pos"].append(np.where(keep_p, pos, 0.0))
        out["lagp"].append(np.where(keep_p, lagp, 0))
        out["neg"].append(np.where(keep_n, neg, 0.0))
        out["lagn"].append(np.where(keep_n, lagn, 0))
        out["gapp"].append(gp)
        out["gapn"].append(gn)
    return {k: np.concatenate(v) for k, v in out.items()}


def direct_circular(xi, xj, P, lags):
    """Brute force sum_n xi[n] xj[(n - l) mod P] (FP64), xi/xj zero-padded to P."""
    a = np.zeros(P); b = np.zeros(P)
    a[:len(xi)] = xi; b[:len(xj)] = xj
    return np.array([sum(a[n] * b[(n - l) % P] for n in range(P)) for l in lags])
