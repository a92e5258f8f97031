"""Draw-request serving on one MI355X: ``Server`` (the C-ABI server of
libftl_hip.so) and ``serve_stream``, which replays a stream of timed
requests against it and measures per-request latency.

Requests are KRP leverage-sample draws over factor sets of uneven shape
(N, I, R, J).  The server admits them FIFO by HBM footprint into a fixed
device arena, runs up to ``max_active`` concurrently on separate streams and
retires them as they finish; a request admitted while others run overlaps
with them on the device.  Factors and uniforms come from the request's seed,
so a result does not depend on what else is being served.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .sampler import PRECISIONS


class Server:
    def __init__(self, hbm_budget_bytes: int, max_active: int = 8):
        _lib.require_device()
        h = ctypes.c_void_p()
        _lib.check(_lib.load().ftl_server_create(int(hbm_budget_bytes), int(max_active), ctypes.byref(h)),
                   "Server")
        self._h = h
        self.max_active = int(max_active)
        self._shapes = {}
        self._waiting_hint = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().ftl_server_destroy(h)
            self._h = None

    def submit(self, dims, R: int, J: int, j=None, seed: int = 0, precision: str = "fp64") -> int:
        """Queue a draw of J samples over the KRP of N factors (I_k = dims[k]),
        excluding mode j (j None or >= N: every mode)."""
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        I = np.ascontiguousarray(dims, dtype=np.uint64)
        N = I.size
        jj = N + 1 if j is None else int(j)
        rid = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_server_submit(self._h, N, I.ctypes.data, int(R), int(J), jj, int(seed),
                                                 PRECISIONS[precision], ctypes.byref(rid)), "submit")
        self._shapes[rid.value] = (N, int(J), [int(d) for d in I], int(R))
        self._waiting_hint += 1
        return int(rid.value)

    def progress(self):
        """Ids of the requests finished since the last call."""
        cap = self.max_active + self._waiting_hint + 1
        buf = np.zeros(cap, dtype=np.uint64)
        n = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_server_progress(self._h, buf.ctypes.data, cap, ctypes.byref(n)), "progress")
        done = [int(x) for x in buf[: int(n.value)]]
        self._waiting_hint = max(0, self._waiting_hint - len(done))
        return done

    def result(self, rid: int) -> np.ndarray:
        """Samples (N x J uint64) of a finished request; releases it."""
        if rid not in self._shapes:
            raise ValueError(f"unknown request {rid}")
        N, J, _, _ = self._shapes.pop(rid)
        out = np.zeros((N, J), dtype=np.uint64)
        _lib.check(_lib.load().ftl_server_result(self._h, rid, out.ctypes.data), "result")
        return out

    def factor(self, rid: int, k: int) -> np.ndarray:
        N, J, dims, R = self._shapes[rid]
        out = np.zeros((dims[k], R))
        _lib.check(_lib.load().ftl_server_factor(self._h, rid, k, out.ctypes.data), "factor")
        return out

    def cancel(self, rid: int):
        _lib.check(_lib.load().ftl_server_cancel(self._h, rid), "cancel")
        self._shapes.pop(rid, None)

    def stats(self) -> dict:
        out = np.zeros(8)
        _lib.check(_lib.load().ftl_server_stats(self._h, out.ctypes.data))
        keys = ["active", "waiting", "finished", "bytes_in_use", "peak_bytes", "admitted", "completed", "cancelled"]
        return {k: int(v) for k, v in zip(keys, out)}

    def wait(self, rid: int, timeout_s: float = 60.0) -> np.ndarray:
        """Drive progress until request rid finishes; return its samples."""
        t0 = time.perf_counter()
        while True:
            if rid in self.progress():
                return self.result(rid)
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError(f"request {rid} did not finish in {timeout_s} s")


def random_requests(n: int, seed: int = 0):
    """Request shapes of the serving stream: N in {3,4,5}, I log-uniform in
    [2^10, 2^22], R in {8,16,32,64}, J log-uniform in [2^10, 2^16]."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        N = int(rng.integers(3, 6))
        dims = [int(2 ** rng.uniform(10, 22)) for _ in range(N)]
        R = int(rng.choice([8, 16, 32, 64]))
        J = int(2 ** rng.uniform(10, 16))
        reqs.append({"dims": dims, "R": R, "J": J, "seed": 100_000 + i})
    return reqs


def serve_stream(server: Server, requests, rate: float, seed: int = 0, precision: str = "fp32",
                 keep_results: bool = False):
    """Replay requests with Poisson arrivals at `rate` per second: each is
    submitted at its arrival time and its latency is completion - arrival,
    completion being the first progress() call that reports it.  Returns
    latencies (s), the makespan (s) and the results when asked."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(requests)))
    arrivals -= arrivals[0]
    pending = {}
    lat = np.full(len(requests), np.nan)
    results = {}
    t0 = time.perf_counter()
    nxt = 0
    remaining = len(requests)
    while remaining:
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            q = requests[nxt]
            rid = server.submit(q["dims"], q["R"], q["J"], None, q["seed"], precision)
            pending[rid] = nxt
            nxt += 1
        for rid in server.progress():
            i = pending.pop(rid)
            lat[i] = (time.perf_counter() - t0) - arrivals[i]
            out = server.result(rid)
            if keep_results:
                results[i] = out
            remaining -= 1
    makespan = time.perf_counter() - t0
    return {"latency_s": lat, "makespan_s": makespan, "arrivals": arrivals, "results": results}
