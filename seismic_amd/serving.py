"""Serving loop over the engine's C ABI (sxc_server_*): a catalog resident in
HBM matched against requests of new events with continuous batching, and a
seeded Poisson load generator that reports p50/p99 latency and throughput."""
from __future__ import annotations

import ctypes
import threading
import time

import numpy as np

from ._lib import (SxcMatch, SxcServerDesc, SxcServerStats, c_float_p, c_int_p, last_error,
                   load_engine)


SXC_ERR_OVERLOAD = -5


class Overloaded(RuntimeError):
    """A submit shed by the server's admission (latency budget or queue bound);
    the request was not queued and may be submitted again later."""


class Server:
    def __init__(self, catalog: np.ndarray, max_query_len: int, top_k: int = 10,
                 max_batch_events: int = 256, n_slots: int = 4, mem_budget: int = 0,
                 demean: bool = False, latency_budget_ms: float = 0.0):
        """catalog: [n_catalog, n_channels, catalog_len] float32.
        latency_budget_ms > 0 turns on deadline admission (see
        set_latency_budget)."""
        self.lib = load_engine()
        cat = np.ascontiguousarray(catalog, dtype=np.float32)
        n, c, L = cat.shape
        self.desc = SxcServerDesc(n, c, L, int(max_query_len), int(top_k), int(max_batch_events),
                                  int(n_slots), int(mem_budget), 1 if demean else 0)
        self.top_k = int(top_k)
        self.n_channels = c
        h = ctypes.c_void_p()
        rc = self.lib.sxc_server_create(ctypes.byref(self.desc), cat.ctypes.data_as(c_float_p),
                                        ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"sxc_server_create failed ({rc}): {last_error(self.lib)}")
        self.handle = h
        self._pending = {}
        if latency_budget_ms:
            self.set_latency_budget(latency_budget_ms)

    def set_latency_budget(self, budget_ms: float) -> None:
        """Submits whose predicted latency exceeds budget_ms raise Overloaded
        (unless the device is idle); 0 turns admission by deadline off."""
        rc = self.lib.sxc_server_set_latency_budget(self.handle, float(budget_ms))
        if rc != 0:
            raise RuntimeError(f"sxc_server_set_latency_budget failed ({rc}): {last_error(self.lib)}")

    def submit(self, events: np.ndarray, lengths) -> int:
        """events: [n, n_channels, stride] float32 ([n, stride] with one
        channel); lengths: n sample counts."""
        ev = np.ascontiguousarray(events, dtype=np.float32)
        if ev.ndim == 2 and self.n_channels == 1:
            ev = ev[:, None, :]
        if ev.ndim != 3 or ev.shape[1] != self.n_channels:
            raise ValueError(f"events must be [n, {self.n_channels}, stride] (got {events.shape})")
        ln = np.ascontiguousarray(lengths, dtype=np.int32).reshape(-1)
        if ln.shape[0] != ev.shape[0]:
            raise ValueError(f"{ln.shape[0]} lengths for {ev.shape[0]} events")
        rid = ctypes.c_longlong()
        rc = self.lib.sxc_server_submit(self.handle, ev.ctypes.data_as(c_float_p), ev.shape[-1],
                                        ln.ctypes.data_as(c_int_p), ev.shape[0], ctypes.byref(rid))
        if rc == SXC_ERR_OVERLOAD:
            raise Overloaded(last_error(self.lib))
        if rc != 0:
            raise RuntimeError(f"sxc_server_submit failed ({rc}): {last_error(self.lib)}")
        self._pending[rid.value] = ev.shape[0]
        return rid.value

    def wait(self, request_id: int):
        """Returns (values [n, k], lags [n, k], indices [n, k], latency_ms)."""
        n = self._pending.pop(request_id, None)
        if n is None:
            raise RuntimeError(f"unknown or already collected request {request_id}")
        out = (SxcMatch * (n * self.top_k))()
        lat = ctypes.c_double()
        rc = self.lib.sxc_server_wait(self.handle, request_id, out, ctypes.byref(lat))
        if rc != 0:
            raise RuntimeError(f"sxc_server_wait failed ({rc}): {last_error(self.lib)}")
        arr = np.frombuffer(out, dtype=np.dtype([("value", "<f4"), ("lag", "<i4"), ("index", "<i4")]))
        arr = arr.reshape(n, self.top_k)
        return arr["value"].copy(), arr["lag"].copy(), arr["index"].copy(), lat.value

    def stats(self) -> dict:
        st = SxcServerStats()
        rc = self.lib.sxc_server_stats(self.handle, ctypes.byref(st))
        if rc != 0:
            raise RuntimeError(f"sxc_server_stats failed ({rc}): {last_error(self.lib)}")
        return {f: getattr(st, f) for f, _ in SxcServerStats._fields_}

    def reset_stats(self):
        self.lib.sxc_server_reset_stats(self.handle)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.sxc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_requests(n_requests, n_channels, len_range, size_range, seed, pool_events=512):
    """Seeded requests: sizes uniform in size_range, lengths uniform in
    len_range, waveforms cut from a pool of synthetic events."""
    from . import synth
    rng = np.random.Generator(np.random.PCG64(seed))
    lmax = len_range[1]
    pool, _ = synth.make_events(pool_events, n_channels, lmax, seed=seed + 1)
    reqs = []
    for _ in range(n_requests):
        n = int(rng.integers(size_range[0], size_range[1] + 1))
        idx = rng.integers(0, pool_events, size=n)
        lens = rng.integers(len_range[0], len_range[1] + 1, size=n).astype(np.int32)
        ev = pool[idx].copy()
        for e in range(n):
            ev[e, :, lens[e]:] = 0.0
        reqs.append((ev, lens))
    return reqs


def run_poisson(server: Server, requests, rate_hz: float, seed: int, collect=False):
    """Submits `requests` at Poisson arrival times (rate_hz) from one thread
    and waits for the admitted ones.  Returns per-request latencies (NaN for
    a request shed by admission), the admitted mask, sizes and timing;
    collect=True also returns every admitted request's results."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate_hz, size=len(requests))
    arrivals = np.cumsum(gaps) - gaps[0]
    ids = [None] * len(requests)
    t0 = time.perf_counter()

    def submitter():
        for k, (ev, lens) in enumerate(requests):
            delay = t0 + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            try:
                ids[k] = server.submit(ev, lens)
            except Overloaded:
                ids[k] = -1

    th = threading.Thread(target=submitter)
    th.start()
    lat = np.full(len(requests), np.nan)
    sizes = np.array([r[0].shape[0] for r in requests])
    results = [None] * len(requests)
    for k in range(len(requests)):
        while ids[k] is None:
            time.sleep(0.0005)
        if ids[k] < 0:
            continue
        r = server.wait(ids[k])
        lat[k] = r[3]
        if collect:
            results[k] = r[:3]
    th.join()
    wall = time.perf_counter() - t0
    out = {"latency_ms": lat, "admitted": ~np.isnan(lat), "sizes": sizes, "wall_s": wall,
           "offered_s": float(arrivals[-1])}
    if collect:
        out["results"] = results
    return out
