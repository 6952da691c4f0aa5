"""Serving a stream of instances on one GPU.

`Server` wraps pmdd_server_* (include/pmdd_engine.h): `slots` worker contexts,
each with its own HIP stream, FIFO admission under a device-memory budget,
so that requests arriving while others run overlap with them on the device.
`slots=1` is the plain one-instance-at-a-time path.

`mixed_stream` draws the request mix of the serving configuration: 60%
knapsack (n uniform in [20,70]), 25% set packing (n in [60,150], m = n/5),
15% TSP (n in [8,11]); objective count uniform over {3,4,5}; every instance
seeded (SplitMix64).  `run_stream` replays a stream with Poisson arrivals and
reports instances/s, Pareto points/s and p50/p99 latency measured from
arrival to result (host diagram construction included).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import queue
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import abi
from .abi import PmddError
from .dd import DecisionDiagram
from .engine import EnumerationResult, _stats_dict, _take_frontier
from . import instances as gen


class Server:
    """pack=True turns on packed serving: method-1 requests without state
    dominance on small diagrams go to a packer of their objective count,
    which expands the current layers of up to pack_inflight in-flight
    requests as one packed layer per step; the rest use the slots."""

    def __init__(self, device: int = 0, slots: int = 8, memory_budget: int = 0,
                 pack: bool = False, pack_inflight: int = 0):
        lib = abi.engine_lib()
        self._lib = lib
        cfg = abi.ServerConfig(device, slots, memory_budget, 1 if pack else 0, pack_inflight)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib.pmdd_server_create(C.byref(cfg), C.byref(h), err, len(err))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        self._h = h
        self.slots = slots

    def submit(self, dd, method: int = 1, state_dominance: int = 0,
               max_cand: int = abi.DEFAULT_MAX_CAND,
               max_prod: int = abi.MAX_PROD_ADAPTIVE) -> int:
        o = abi.Opts(int(method), int(dd.problem_type), int(state_dominance), 0, int(max_cand),
                     int(max_prod))
        t = C.c_int64()
        err = C.create_string_buffer(1024)
        rc = self._lib.pmdd_server_submit(self._h, dd.view_ptr, C.byref(o), C.byref(t), err,
                                          len(err))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        return int(t.value)

    def _collect(self, fn, ticket):
        fr = abi.Frontier()
        st = abi.Stats()
        err = C.create_string_buffer(1024)
        rc = fn(self._h, C.c_int64(ticket), C.byref(fr), C.byref(st), err, len(err))
        if rc == abi.PMDD_PENDING:
            return None
        if rc != abi.PMDD_OK:
            raise PmddError(rc, err.value.decode())
        pts = _take_frontier(self._lib, fr, True)
        return EnumerationResult(pts, _stats_dict(st, 0, None, None))

    def poll(self, ticket: int):
        return self._collect(self._lib.pmdd_server_poll, ticket)

    def wait(self, ticket: int) -> EnumerationResult:
        return self._collect(self._lib.pmdd_server_wait, ticket)

    def wait_any(self):
        """Blocks until some request has finished; returns (ticket, result),
        or (ticket, PmddError) when that request failed."""
        fr = abi.Frontier()
        st = abi.Stats()
        t = C.c_int64(-1)
        err = C.create_string_buffer(1024)
        rc = self._lib.pmdd_server_wait_any(self._h, C.byref(t), C.byref(fr), C.byref(st), err,
                                            len(err))
        if t.value < 0:
            raise PmddError(rc, err.value.decode())
        if rc != abi.PMDD_OK:
            return int(t.value), PmddError(rc, err.value.decode())
        pts = _take_frontier(self._lib, fr, True)
        return int(t.value), EnumerationResult(pts, _stats_dict(st, 0, None, None))

    def info(self) -> dict:
        """The server's admission record (pmdd_server_get_info)."""
        inf = abi.ServerInfo()
        rc = self._lib.pmdd_server_get_info(self._h, C.byref(inf))
        if rc != abi.PMDD_OK:
            raise PmddError(rc, "pmdd_server_get_info failed")
        return {name: getattr(inf, name) for name, _ in abi.ServerInfo._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            self._lib.pmdd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def mixed_stream(count: int, seed: int = 2000) -> list[dict]:
    """Request specs of the serving configuration (see module docstring)."""
    rng = gen.SplitMix64(seed)
    specs = []
    for i in range(count):
        r = rng.uniform(0, 99)
        d = rng.uniform(3, 5)
        inst_seed = rng.next() & 0xFFFFFFFF
        if r < 60:
            specs.append({"kind": 1, "n": rng.uniform(20, 70), "d": d, "seed": inst_seed})
        elif r < 85:
            n = rng.uniform(60, 150)
            specs.append({"kind": 2, "n": n, "d": d, "seed": inst_seed, "m": n // 5})
        else:
            specs.append({"kind": 3, "n": rng.uniform(8, 11), "d": d, "seed": inst_seed})
    return specs


def poisson_arrivals(count: int, rate: float, seed: int = 7) -> list[float]:
    """Arrival times (s) of a Poisson process of `rate` requests per second."""
    rng = gen.SplitMix64(seed)
    t = 0.0
    out = []
    for _ in range(count):
        u = ((rng.next() >> 11) + 0.5) / float(1 << 53)
        t += -math.log(u) / rate
        out.append(t)
    return out


def write_stream(specs: list[dict], directory: str) -> list[str]:
    os.makedirs(directory, exist_ok=True)
    paths = []
    for i, sp in enumerate(specs):
        extra = {"m": sp["m"]} if "m" in sp else {}
        inst = gen.generate(sp["kind"], sp["seed"], sp["n"], sp["d"], **extra)
        paths.append(gen.write(inst, os.path.join(directory, f"req{i:05d}.dat")))
    return paths


def run_stream(paths: list[str], specs: list[dict], arrivals: list[float], slots: int,
               method: int = 1, device: int = 0, builders: int = 16,
               results: dict | None = None, progress: int = 0,
               stats: dict | None = None, memory_budget: int = 0, pack: bool = False,
               pack_inflight: int = 0) -> dict:
    """Replay the stream: at each arrival the request's diagram is built on a
    host thread and submitted; one collector thread takes results as they
    finish (pmdd_server_wait_any); latency runs from arrival to result.
    With progress=k, a line goes to stderr after every k results.  `results`
    and `stats` receive each request's frontier and statistics by index; the
    returned summary holds the server's admission record and the largest
    ratio of a request's measured device memory to its estimate."""
    server = Server(device, slots=slots, memory_budget=memory_budget, pack=pack,
                    pack_inflight=pack_inflight)
    n = len(paths)
    done_at = [0.0] * n
    points = [0] * n
    errors = {}
    owner = {}  # ticket -> request
    mem = {}  # request -> (measured peak device bytes, admission estimate)
    lock = threading.Lock()
    events = queue.Queue()  # one item per request: True once submitted, False if not
    t0 = time.perf_counter()

    def submit(i):
        delay = t0 + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        try:
            dd = DecisionDiagram.from_file(paths[i], specs[i]["kind"])
            with lock:
                ticket = server.submit(dd, method=method)
                owner[ticket] = i
            dd.close()
            events.put(True)
        except PmddError as e:
            with lock:
                errors[i] = (e.code, e.message)
            done_at[i] = time.perf_counter()
            events.put(False)

    last_print = [t0]

    def collect():
        for k in range(n):
            now = time.perf_counter()
            # a line after every `progress` results, and after a result that
            # comes 10 s or more after the last line
            if progress and k > 0 and (k % progress == 0 or now - last_print[0] >= 10.0):
                print(f"served {k}/{n} after {now - t0:.1f} s", file=sys.stderr, flush=True)
                last_print[0] = now
            if not events.get():
                continue
            ticket, res = server.wait_any()
            now = time.perf_counter()
            with lock:
                i = owner[ticket]
                if isinstance(res, PmddError):
                    errors[i] = (res.code, res.message)
                else:
                    points[i] = len(res.points)
                    mem[i] = (res.stats["mem_need_peak_bytes"], res.stats["mem_estimate_bytes"])
                    if results is not None:
                        results[i] = res.points
                    if stats is not None:
                        stats[i] = res.stats
            done_at[i] = now

    collector = threading.Thread(target=collect)
    collector.start()
    with ThreadPoolExecutor(max_workers=builders) as pool:
        list(pool.map(submit, range(n)))
    collector.join()
    info = server.info()
    server.close()
    lat = np.array([done_at[i] - (t0 + arrivals[i]) for i in range(n) if i not in errors])
    span = max(done_at) - t0
    ratios = [peak / est for peak, est in mem.values() if est > 0]
    over = sum(1 for peak, est in mem.values() if peak > est)
    return {
        "pack": pack,
        "admission": info,
        "mem_peak_over_estimate_max": max(ratios) if ratios else None,
        "mem_peak_above_estimate": over,
        "slots": slots,
        "requests": n,
        "failed": len(errors),
        "errors": {str(k): v for k, v in list(errors.items())[:5]},
        "span_s": span,
        "inst_per_s": (n - len(errors)) / span,
        "pareto_pts_per_s": sum(points) / span,
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "mean_latency_s": float(lat.mean()) if len(lat) else None,
    }
