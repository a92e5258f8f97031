"""Draw-request serving on one MI355X: ``Server`` (the C-ABI server of
libftl_hip.so) and ``serve_stream``, which replays a stream of timed
requests against it and measures per-request latency.

Requests are KRP leverage-sample draws over factor sets of uneven shape
(N, I, R, J): either over factors the server generates on the device from
the request's seed, or over a caller's factor set registered once
(``add_factors`` / ``add_sampler``) whose Gram trees stay resident, with
``update_factor`` rebuilding one mode's tree.  The server admits requests
FIFO by HBM footprint into a fixed device arena; requests whose trees are
complete are drawn together in packed launches (one launch of each computeM
kernel and of the draw kernel per rank and precision).  A result depends only
on the request's factors and seed, so it is the same served alone or in a mix.
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
        # held by the object: module globals may be gone when it is collected at exit
        self._destroy = _lib.load().ftl_server_destroy
        self.max_active = int(max_active)
        self._shapes = {}
        self._sets = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
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
        return int(rid.value)

    def add_factors(self, factors, J_max: int, precision: str = "fp64") -> int:
        """Register a factor set (NumPy float64 I_k x R arrays, or torch device
        tensors); returns its id for ``submit_on``."""
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        on_device = hasattr(factors[0], "data_ptr")
        N = len(factors)
        R = int(factors[0].shape[1])
        dims = np.ascontiguousarray([int(f.shape[0]) for f in factors], dtype=np.uint64)
        if on_device:
            from .sampler import _dev_ptr
            for k, f in enumerate(factors):
                if str(f.dtype) != "torch.float64" or f.dim() != 2 or int(f.shape[1]) != R:
                    raise TypeError(f"factors[{k}] must be a float64 I_k x R device tensor")
            ptrs = (ctypes.c_void_p * N)(*[_dev_ptr(f, f"factors[{k}]") for k, f in enumerate(factors)])
        else:
            from .sampler import _require_array
            for k, f in enumerate(factors):
                _require_array(f, np.float64, 2, f"factors[{k}]")
                if f.shape[1] != R:
                    raise ValueError("factors must have R columns")
            ptrs = (ctypes.c_void_p * N)(*[f.ctypes.data for f in factors])
        fs = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_server_add_factors(self._h, N, dims.ctypes.data, R, ptrs, int(on_device),
                                                      PRECISIONS[precision], int(J_max), ctypes.byref(fs)),
                   "add_factors")
        self._sets[int(fs.value)] = (N, [int(d) for d in dims], R)
        return int(fs.value)

    def add_sampler(self, sampler) -> int:
        """Register the factors of an existing ``Sampler`` (J_max = its J)."""
        fs = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_server_add_sampler(self._h, sampler.handle, ctypes.byref(fs)), "add_sampler")
        self._sets[int(fs.value)] = (sampler.N, list(sampler.dims), sampler.R)
        return int(fs.value)

    def update_factor(self, fs: int, k: int, Uk):
        """Replace factor k of a registered set and rebuild its tree only."""
        on_device = hasattr(Uk, "data_ptr")
        if on_device:
            from .sampler import _dev_ptr
            ptr = _dev_ptr(Uk, "Uk")
        else:
            from .sampler import _require_array
            _require_array(Uk, np.float64, 2, "Uk")
            ptr = Uk.ctypes.data
        _lib.check(_lib.load().ftl_server_update_factor(self._h, int(fs), int(k), ptr, int(on_device)),
                   "update_factor")

    def remove_factors(self, fs: int):
        _lib.check(_lib.load().ftl_server_remove_factors(self._h, int(fs)), "remove_factors")
        self._sets.pop(int(fs), None)

    def submit_on(self, fs: int, J: int, j=None, seed: int = 0) -> int:
        """Queue a draw of J samples on registered factor set fs."""
        if int(fs) not in self._sets:
            raise ValueError(f"unknown factor set {fs}")
        N, dims, R = self._sets[int(fs)]
        jj = N + 1 if j is None else int(j)
        rid = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_server_submit_on(self._h, int(fs), int(J), jj, int(seed), ctypes.byref(rid)),
                   "submit_on")
        self._shapes[rid.value] = (N, int(J), dims, R)
        return int(rid.value)

    def progress(self):
        """Ids of the requests finished and not reported before."""
        cap = max(64, len(self._shapes))
        buf = np.zeros(cap, dtype=np.uint64)
        n = ctypes.c_uint64()
        _lib.check(_lib.load().ftl_server_progress(self._h, buf.ctypes.data, cap, ctypes.byref(n)), "progress")
        return [int(x) for x in buf[: int(n.value)]]

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

    def wait_for(self, rid: int = 0, timeout_s: float = 60.0) -> bool:
        """Drive the server (sleeping between polls of the device's completion
        events) until request rid has finished -- rid 0: until some request
        finishes that progress() has not reported -- or timeout_s passed."""
        done = ctypes.c_int()
        _lib.check(_lib.load().ftl_server_wait(self._h, int(rid), float(timeout_s), ctypes.byref(done)), "wait")
        return bool(done.value)

    def wait(self, rid: int, timeout_s: float = 60.0) -> np.ndarray:
        """Wait until request rid finishes; return its samples."""
        if not self.wait_for(rid, timeout_s):
            raise TimeoutError(f"request {rid} did not finish in {timeout_s} s")
        self.progress()  # report it (and anything else finished) as done
        return self.result(rid)


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
    completion being the first progress() call that reports it.  A request
    with key "fs" draws J samples on that registered factor set
    (``submit_on``); the others bring their own shapes ("dims", "R").  Returns
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
            if "fs" in q:
                rid = server.submit_on(q["fs"], q["J"], None, q["seed"])
            else:
                rid = server.submit(q["dims"], q["R"], q["J"], None, q["seed"], precision)
            pending[rid] = nxt
            nxt += 1
        # sleep in the server until a request finishes or the next one arrives
        until_next = (arrivals[nxt] - (time.perf_counter() - t0)) if nxt < len(requests) else 0.05
        server.wait_for(0, max(0.0, min(until_next, 0.05)))
        for rid in server.progress():
            i = pending.pop(rid)
            lat[i] = (time.perf_counter() - t0) - arrivals[i]
            out = server.result(rid)
            if keep_results:
                results[i] = out
            remaining -= 1
    makespan = time.perf_counter() - t0
    return {"latency_s": lat, "makespan_s": makespan, "arrivals": arrivals, "results": results}
