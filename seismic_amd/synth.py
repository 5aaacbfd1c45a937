"""Seeded synthetic seismic events: band-passed noise plus a Ricker wavelet.

The reference ships no data (its driver loads ``events_eruption_v8.npy``,
Correlation_in_CUDA/test_cuda.py:22, which is git-ignored upstream), so every
workload here is generated.  Each trace is

    x_c[n] = noise_c[n] + A_c * ricker(n - offset)

with ``noise_c`` white Gaussian noise band-passed to 1-20 Hz at 100 Hz (zero
phase, cosine-tapered edges in the frequency domain), a 5 Hz Ricker wavelet
``(1 - 2 pi^2 f^2 t^2) exp(-pi^2 f^2 t^2)`` placed at a per-event uniform
offset shared by all channels, and ``A_c = SNR * rms(noise_c) / max|ricker|``.
Traces are demeaned in FP64 and stored as float32.
"""
from __future__ import annotations

import numpy as np

FS = 100.0
RICKER_F = 5.0
BAND = (1.0, 20.0)


def ricker(fs: float = FS, f: float = RICKER_F, half_width_s: float = 0.4) -> np.ndarray:
    """Sampled Ricker wavelet centred on its middle sample (peak 1.0)."""
    n = int(round(half_width_s * fs))
    t = np.arange(-n, n + 1, dtype=np.float64) / fs
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def _bandpass_gain(n: int, fs: float, band=BAND, taper_lo: float = 0.5, taper_hi: float = 2.0) -> np.ndarray:
    f = np.fft.rfftfreq(n, d=1.0 / fs)
    lo, hi = band
    g = np.zeros_like(f)
    g[(f >= lo) & (f <= hi)] = 1.0
    m = (f >= lo - taper_lo) & (f < lo)
    g[m] = 0.5 - 0.5 * np.cos(np.pi * (f[m] - (lo - taper_lo)) / taper_lo)
    m = (f > hi) & (f <= hi + taper_hi)
    g[m] = 0.5 + 0.5 * np.cos(np.pi * (f[m] - hi) / taper_hi)
    return g


def make_events(n_events: int, n_channels: int, n_samples: int, seed: int,
                snr=(0.5, 10.0), log_snr: bool = False, dup_frac: float = 0.0,
                fs: float = FS, chunk: int = 4096):
    """Return (events[N, C, L] float32, meta dict).

    ``dup_frac`` > 0 turns that fraction of events into exact copies of an
    earlier event shifted by a random lag, with zero fill (``meta['dups']``
    lists (src, dst, shift); dst's samples are src's delayed by ``shift``).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    L = int(n_samples)
    w = ricker(fs)
    half = len(w) // 2
    if L < len(w) + 2:
        raise ValueError("trace too short for the wavelet")
    gain = _bandpass_gain(L, fs)
    out = np.empty((n_events, n_channels, L), dtype=np.float32)
    offsets = rng.integers(half, L - half, size=n_events)
    if log_snr:
        snrs = np.exp(rng.uniform(np.log(snr[0]), np.log(snr[1]), size=(n_events, n_channels)))
    else:
        snrs = rng.uniform(snr[0], snr[1], size=(n_events, n_channels))
    for e0 in range(0, n_events, chunk):
        e1 = min(n_events, e0 + chunk)
        noise = rng.standard_normal((e1 - e0, n_channels, L))
        noise = np.fft.irfft(np.fft.rfft(noise, axis=-1) * gain, n=L, axis=-1)
        rms = np.sqrt(np.mean(noise ** 2, axis=-1))
        sig = noise
        for e in range(e0, e1):
            o = int(offsets[e])
            amp = snrs[e] * rms[e - e0]  # max|ricker| == 1
            sig[e - e0, :, o - half:o + half + 1] += amp[:, None] * w[None, :]
        sig -= sig.mean(axis=-1, keepdims=True)
        out[e0:e1] = sig.astype(np.float32)
    dups = []
    if dup_frac > 0.0:
        n_dup = int(round(dup_frac * n_events))
        dst = rng.choice(np.arange(1, n_events), size=n_dup, replace=False) if n_dup else []
        for d in sorted(int(x) for x in dst):
            s = int(rng.integers(0, d))
            sh = int(rng.integers(-L // 8, L // 8 + 1))
            src = out[s]
            y = np.zeros_like(src)
            if sh >= 0:
                y[:, sh:] = src[:, :L - sh]
            else:
                y[:, :L + sh] = src[:, -sh:]
            out[d] = y
            dups.append((s, d, sh))
    meta = {"seed": seed, "offsets": offsets, "snr": snrs, "dups": dups, "fs": fs}
    return out, meta


def pad_like_binding(events: np.ndarray):
    """Padded and reversed copies exactly as the reference binding builds them
    (Correlation_in_CUDA/correlation_lib.py:38-62): L taken from event 0,
    P = 1 << L.bit_length(), rows zero-padded to P, reversed rows are the
    flipped trace zero-padded to P."""
    ev = [np.asarray(e.data if hasattr(e, "data") else e, dtype=np.float32) for e in events]
    L = int(ev[0].shape[0])
    P = 1 << L.bit_length()
    n = len(ev)
    padded = np.zeros((n, P), dtype=np.float32)
    rev = np.zeros((n, P), dtype=np.float32)
    for i, x in enumerate(ev):
        padded[i, :x.shape[0]] = x
        rev[i, :x.shape[0]] = x[::-1]
    return padded, rev, L, P
