"""Serving loop for generate / generate_inv requests (no reference counterpart).

The reference is a library of pure functions; serving many independent requests is this
package's own layer on top of the same kernels.  Its parts:

* ``GraphCache``: resident graphs (points, neighbours, indices, offsets in HBM) keyed by
  the caller's graph key, least-recently-used eviction, byte accounting.
* ``Admission``: an HBM budget that covers the resident graphs plus, for every request in
  flight, its workspace and outputs.  A request is admitted only if its bytes fit after
  evicting idle graphs; otherwise it waits for in-flight requests to finish.
* ``Server``: continuous batching.  Each pass of the scheduler dispatches every request
  that has arrived and fits the budget, on three HIP streams (default priority: small
  work, small builds, the large class), so small requests overlap large ones on the GPU
  instead of queueing behind them; a request whose graph is not resident is built first
  (``build_graph``, on a builder thread) and then cached.
  Completion is observed with HIP events; latency = completion - arrival on the GPU clock.

A request's result does not depend on what else is in flight: every kernel reads only its
own request's graph, table and inputs, and the refinement's values do not depend on the
order in which the dataflow publishes them.  ``tests/test_serving.py`` checks that a
request's output in a mix equals its output when served alone, bit for bit.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Any, Callable, Dict, Hashable, List, Optional

import numpy as np

GiB = 1 << 30


# ---------------------------------------------------------------- byte accounting
ALLOC_SLACK = 2 << 20  # the caching allocator rounds each block up to 2 MiB


def scheduled(n: int, n0: int) -> bool:
    """Whether build_graph gives a graph of this size a dataflow schedule."""
    from .graph import SCHEDULE_MIN_REFINED

    return n - n0 >= SCHEDULE_MIN_REFINED


def graph_bytes(n: int, d: int, k: int, n0: int, point_bytes: int = 4, schedule: bool = False) -> int:
    """Device bytes of a resident Graph: points, neighbours, indices and, with
    ``schedule`` for a graph large enough to get one, its dataflow schedule (slot order,
    neighbour slots, points in slot order)."""
    b = n * d * point_bytes + (n - n0) * k * 4 + n * 4
    if schedule and scheduled(n, n0):
        b += (n - n0) * (k + 1) * 4 + n * d * point_bytes
    return b


def request_bytes(n: int, d: int, k: int, n0: int, n_bins: int, precision: str) -> int:
    """Device bytes a generate / generate_inv call allocates while in flight: the
    library's own workspace for it (tables, graph-order values, seed block, counters,
    FP64 recomputation lists) and the output vector (the input belongs to the caller)."""
    from . import _lib

    vb = 8 if precision == "fp64" else 4
    ws = _lib.lib().ggp_refine_workspace(n, d, n0, k, n_bins, _lib.PREC[precision], 1)
    return ws + n * vb + 2 * ALLOC_SLACK


def build_bytes(n: int, d: int, k: int, n0: int, schedule: bool = False) -> int:
    """Transient device bytes of build_graph beyond the resulting graph: the library's
    build workspace (tree, kNN, depths, ordering), the depth vector and, with
    ``schedule`` for a graph large enough, the schedule's workspace."""
    from . import _lib

    L = _lib.lib()
    b = L.ggp_build_graph_workspace(n, d, 0, n0, k) + n * 4 + 2 * ALLOC_SLACK
    if schedule and scheduled(n, n0):
        b += L.ggp_graph_schedule_workspace(n, n0) + ALLOC_SLACK
    return b


# ---------------------------------------------------------------- resident graphs
@dataclasses.dataclass
class _Entry:
    graph: Any
    nbytes: int
    users: int = 0


class GraphCache:
    """Resident graphs with LRU order; entries in use by an in-flight request are pinned."""

    def __init__(self):
        self._d: "collections.OrderedDict[Hashable, _Entry]" = collections.OrderedDict()
        self.bytes = 0
        self.hits = 0
        self.misses = 0
        self.evictions = 0

    def __contains__(self, key):
        return key in self._d

    def __len__(self):
        return len(self._d)

    def get(self, key):
        e = self._d.get(key)
        if e is None:
            self.misses += 1
            return None
        self.hits += 1
        self._d.move_to_end(key)
        return e.graph

    def put(self, key, graph, nbytes: int):
        if key in self._d:
            self.bytes -= self._d[key].nbytes
        self._d[key] = _Entry(graph, int(nbytes), self._d[key].users if key in self._d else 0)
        self._d.move_to_end(key)
        self.bytes += int(nbytes)

    def pin(self, key):
        self._d[key].users += 1

    def unpin(self, key):
        self._d[key].users -= 1

    def evictable_bytes(self) -> int:
        return sum(e.nbytes for e in self._d.values() if e.users == 0)

    def evict_until(self, need_free: int) -> int:
        """Evict idle graphs, least recently used first, until ``need_free`` bytes are
        released (or nothing idle is left).  Returns the bytes released."""
        freed = 0
        for key in list(self._d.keys()):
            if freed >= need_free:
                break
            e = self._d[key]
            if e.users:
                continue
            del self._d[key]
            self.bytes -= e.nbytes
            freed += e.nbytes
            self.evictions += 1
        return freed


class Admission:
    """HBM budget over resident graphs plus in-flight request bytes."""

    def __init__(self, budget_bytes: int, cache: GraphCache):
        self.budget = int(budget_bytes)
        self.cache = cache
        self.inflight = 0

    def used(self) -> int:
        return self.cache.bytes + self.inflight

    def try_admit(self, need: int) -> bool:
        """Reserve ``need`` bytes, evicting idle graphs if that makes them fit."""
        if need > self.budget:
            raise ValueError(f"request needs {need} bytes, more than the whole budget {self.budget}")
        over = self.used() + need - self.budget
        if over > 0:
            if self.cache.evictable_bytes() < over:
                return False
            self.cache.evict_until(over)
        self.inflight += need
        return True

    def release(self, need: int):
        self.inflight -= need


# ---------------------------------------------------------------- requests
@dataclasses.dataclass
class Request:
    """One call: ``op`` is "generate" (data = xi) or "generate_inv" (data = values).

    ``graph_key`` names the graph; ``points`` (device tensor) is needed only when that
    graph is not resident.  ``arrival`` is seconds after the server's start."""

    rid: int
    op: str
    graph_key: Hashable
    n: int
    d: int
    n0: int
    k: int
    covariance: Any
    data: Any
    arrival: float = 0.0
    precision: Optional[str] = None
    points: Any = None
    n_bins: int = 1000
    # filled in by the server (seconds after the server's start)
    result: Any = None
    done_s: float = math.nan
    dispatch_s: float = math.nan  # host time the work was issued
    start_s: float = math.nan  # GPU time the request's first kernel could start
    built_graph: bool = False

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return math.nan
    i = min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))
    return xs[i]


class Server:
    """Continuous-batching server over HIP streams (see the module docstring).

    Two size classes (up to ``small_n`` points, and larger), each with a builder thread
    for graphs that are not resident.  Three streams: small work, small builds, and the
    large class (its builds and its work).  A process gets four hardware queues (HIP's
    default), one of them taken by the default stream, and streams beyond that share a
    queue, whose packets run in order: a fourth stream would put some small work behind
    large kernels.  So small requests never queue behind large ones on the GPU, nor behind
    a graph build on the host.  The builds synchronise with the host inside
    ``build_graph``; the ctypes calls release the interpreter lock, so dispatch continues
    meanwhile.  Requests for a graph under construction wait on it; every other arrived
    request that fits the HBM budget is dispatched at once, in arrival order, skipping
    those that do not fit yet.
    """

    def __init__(self, hbm_budget_bytes: int = 200 * 10**9, max_inflight: int = 64, small_n: int = 1_000_000,
                 large_cu_fraction: float = 0.5, build_fn: Optional[Callable] = None, schedule: bool = False):
        import concurrent.futures as cf

        import torch

        import graphgp_amd as gp

        self.torch = torch
        self.gp = gp
        self.cache = GraphCache()
        self.adm = Admission(hbm_budget_bytes, self.cache)
        self.small_n = small_n
        self.max_inflight = max_inflight
        self.build_depth = 2  # builds per class handed to its builder: the running one and the next
        # ``schedule``: build large graphs (scheduled(n, n0)) with their dataflow schedule,
        # whose bytes graph_bytes and build_bytes then count.  Off by default: on the
        # synthetic mix, where builds dominate and a graph serves about three requests,
        # computing schedules lowers the batch capacity more than they speed up generate.
        self.schedule = bool(schedule)
        self.build_fn = build_fn or (lambda pts, n0, k: gp.build_graph(pts, n0=n0, k=k, schedule=self.schedule))
        small_work, small_build, large = torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Stream()
        self.work = [small_work, large]  # [small, large]
        self.build_streams = [small_build, large]
        self.builders = [cf.ThreadPoolExecutor(max_workers=1), cf.ThreadPoolExecutor(max_workers=1)]
        cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        # While small requests are waiting or running, the long launches of large requests
        # (persistent refinement phases and the kNN, which yield: their extra workgroups
        # stop taking work; grid-stride passes of a build, sized at launch) keep this share
        # of the CUs and leave the rest to the small requests; otherwise they use the whole
        # device.  The flag the kernels read is a word of pinned host memory that the
        # scheduler sets.  A smaller share lowers the small requests' tail latency and the
        # served throughput (1.0: no reservation).  On the 2000-request mix at 90% load, the
        # share against p99 for N <= 1e6 and the sustained fraction of batch throughput:
        # 0.35: 25 ms, 0.76; 0.5: 40 ms, 0.85; 0.65: 57 ms, 0.88; 0.8: 69 ms, 0.90.
        self.large_cus = max(1, int(cus * large_cu_fraction)) if large_cu_fraction < 1 else 0
        self._yield = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self._yield_np = self._yield.numpy()
        self._lib = gp._lib.lib()

    def close(self):
        """Wait for all work and stop the builder threads."""
        self.torch.cuda.synchronize()
        for b in self.builders:
            b.shutdown(wait=True)

    def _cls(self, req: Request) -> int:
        return 0 if req.n <= self.small_n else 1

    def _reserve(self, req: Request):
        """Launches of a large request on this thread yield to small requests."""
        if self._cls(req) and self.large_cus:
            self.gp._lib.check(self._lib.ggp_set_yield(self._yield.data_ptr(), self.large_cus), "set_yield")

    def _build(self, req: Request):
        torch = self.torch
        stream = self.build_streams[self._cls(req)]
        with torch.cuda.stream(stream):
            self._reserve(req)
            try:
                g = self.build_fn(req.points, req.n0, req.k)
            finally:
                self._lib.ggp_set_yield(None, 0)
            ev = torch.cuda.Event()
            ev.record(stream)
        return g, ev

    def _execute(self, req: Request, g, after=None):
        torch, gp = self.torch, self.gp
        stream = self.work[self._cls(req)]
        with torch.cuda.stream(stream):
            if after is not None:
                stream.wait_event(after)
            req._start_ev = torch.cuda.Event(enable_timing=True)
            req._start_ev.record(stream)
            self._reserve(req)
            try:
                if req.op == "generate":
                    out = gp.generate(g, req.covariance, req.data, precision=req.precision)
                elif req.op == "generate_inv":
                    out = gp.generate_inv(g, req.covariance, req.data, precision=req.precision)
                else:
                    raise ValueError(f"unknown op {req.op!r}")
            finally:
                self._lib.ggp_set_yield(None, 0)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(stream)
        return out, ev

    def run(self, requests: List[Request], realtime: bool = True) -> Dict[str, Any]:
        """Serve ``requests``.  With ``realtime`` a request is not started before its
        arrival time; otherwise all are available at once (batch mode).  Returns latency
        percentiles and throughput; each request gets ``result``, ``dispatch_s``,
        ``start_s`` and ``done_s``."""
        torch = self.torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: List[Request] = []
        building: Dict[Hashable, Any] = {}  # key -> (future, [parked requests], reserved bytes)
        inflight = []  # (req, event, bytes)
        torch.cuda.synchronize()
        start_ev = torch.cuda.Event(enable_timing=True)
        start_ev.record(self.work[0])
        t0 = time.perf_counter()
        done: List[Request] = []

        def need_of(req):
            return request_bytes(req.n, req.d, req.k, req.n0, req.n_bins, req.precision or "fp32_increment")

        def launch(req, g, after=None):
            req.dispatch_s = time.perf_counter() - t0
            self.cache.pin(req.graph_key)
            out, ev = self._execute(req, g, after)
            req.result = out
            inflight.append((req, ev, need_of(req)))

        def retire(block: bool):
            keep = []
            for req, ev, nb in inflight:
                if block or ev.query():
                    ev.synchronize()
                    req.done_s = start_ev.elapsed_time(ev) / 1e3
                    req.start_s = start_ev.elapsed_time(req._start_ev) / 1e3
                    self.adm.release(nb)
                    self.cache.unpin(req.graph_key)
                    done.append(req)
                else:
                    keep.append((req, ev, nb))
            inflight[:] = keep

        def finish_builds():
            progressed = False
            for key in list(building.keys()):
                fut, parked, reserved = building[key]
                if not fut.done():
                    continue
                g, ev = fut.result()
                del building[key]
                r0 = parked[0]
                self.adm.release(reserved)
                self.cache.put(key, g, graph_bytes(r0.n, r0.d, r0.k, r0.n0, schedule=self.schedule))
                for req in parked:
                    launch(req, g, after=ev)
                progressed = True
            return progressed

        small = self.small_n

        def builds_of(cls):
            return sum(1 for b in building.values() if self._cls(b[1][0]) == cls)

        def small_active():
            return (any(r.n <= small for r in waiting) or any(r.n <= small for r, _, _ in inflight)
                    or any(b[1][0].n <= small for b in building.values()))

        while pending or waiting or inflight or building:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                waiting.append(pending.popleft())
            self._yield_np[0] = 1 if small_active() else 0
            progressed = finish_builds()
            still = []
            for req in waiting:
                key = req.graph_key
                need = need_of(req)
                if key in building:
                    if self.adm.try_admit(need):
                        building[key][1].append(req)
                        progressed = True
                        continue
                elif key in self.cache:
                    if len(inflight) < self.max_inflight and self.adm.try_admit(need):
                        launch(req, self.cache.get(key))
                        progressed = True
                        continue
                elif builds_of(self._cls(req)) < self.build_depth:
                    # a build is handed to its class's builder thread (one build at a time)
                    # only while fewer than build_depth are there, and reserves its bytes
                    # only then: queued builds would hold memory that dispatchable work needs
                    if req.points is None:
                        raise ValueError(f"request {req.rid}: graph {key!r} is not resident and no points given")
                    extra = (graph_bytes(req.n, req.d, req.k, req.n0, schedule=self.schedule)
                             + build_bytes(req.n, req.d, req.k, req.n0, schedule=self.schedule))
                    if self.adm.try_admit(need + extra):  # extra is released when the build ends
                        req.built_graph = True
                        building[key] = (self.builders[self._cls(req)].submit(self._build, req), [req], extra)
                        progressed = True
                        continue
                still.append(req)
            waiting = still
            retire(block=False)
            self._yield_np[0] = 1 if small_active() else 0
            if not progressed:
                if inflight or building:
                    time.sleep(1e-4)  # poll completions, builds and arrivals
                elif pending and realtime:
                    time.sleep(max(0.0, min(0.002, pending[0].arrival - (time.perf_counter() - t0))))
        retire(block=True)
        self._yield_np[0] = 0
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = [r.latency_s for r in done]
        pts = sum(r.n - r.n0 for r in done)
        # steady state: refined points completed, and offered, per second over the middle 80%
        # of the arrival window (the first requests meet an empty device and the last ones
        # leave a drain tail; neither is the rate the server sustains)
        t_last = max((r.arrival for r in done), default=0.0)
        w1, w2 = 0.1 * t_last, 0.9 * t_last
        if realtime and w2 > w1:
            sustained = sum(r.n - r.n0 for r in done if w1 <= r.done_s < w2) / (w2 - w1)
            offered = sum(r.n - r.n0 for r in done if w1 <= r.arrival < w2) / (w2 - w1)
        else:
            sustained = offered = math.nan
        small = [r.latency_s for r in done if r.n <= self.small_n]
        warm_small = [r.latency_s for r in done if r.n <= self.small_n and not r.built_graph]
        return {
            "requests": len(done),
            "wall_s": wall,
            "refined_points": pts,
            "points_per_s": pts / wall if wall > 0 else math.nan,
            "sustained_points_per_s": sustained,
            "offered_points_per_s": offered,
            "p50_ms": percentile(lat, 50) * 1e3,
            "p99_ms": percentile(lat, 99) * 1e3,
            "p50_ms_small": percentile(small, 50) * 1e3,
            "p99_ms_small": percentile(small, 99) * 1e3,
            "p99_ms_small_resident_graph": percentile(warm_small, 99) * 1e3,
            "graphs_built": sum(r.built_graph for r in done),
            "cache_hits": self.cache.hits,
            "cache_evictions": self.cache.evictions,
            "resident_bytes": self.cache.bytes,
        }


# ---------------------------------------------------------------- synthetic mix
def synthetic_mix(n_requests: int, rate_per_s: float, seed: int = 0, n_min: float = 1e3, n_max: float = 1e7,
                  reuse: float = 0.7, n_bins: int = 1000, precision: str = "fp32_increment"):
    """Request descriptors of the serving workload: N log-uniform in [n_min, n_max], d in
    {2, 3}, uniform or 8-cluster points, k in {8, 16, 32}, n0 = max(k, min(1000, N // 100)),
    RBF or Matern-3/2 tables with length scale log-uniform in [0.01, 0.3], Poisson arrivals,
    a fraction ``reuse`` of requests reusing an earlier graph with fresh xi.  Returns plain
    dicts; ``materialize`` turns them into Requests with device data."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs, graphs = [], []
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate_per_s)
        if graphs and rng.uniform() < reuse:
            gs = graphs[rng.integers(len(graphs))]
        else:
            n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
            k = int(rng.choice([8, 16, 32]))
            gs = dict(key=f"g{len(graphs)}", n=n, d=int(rng.choice([2, 3])), k=k,
                      n0=max(k, min(1000, n // 100)), clustered=bool(rng.uniform() < 0.5),
                      seed=int(rng.integers(1 << 31)))
            graphs.append(gs)
        cov = dict(kind=str(rng.choice(["rbf", "matern32"])),
                   scale=float(math.exp(rng.uniform(math.log(0.01), math.log(0.3)))), n_bins=n_bins)
        op = "generate" if rng.uniform() < 0.8 else "generate_inv"
        specs.append(dict(rid=rid, arrival=t, graph=gs, cov=cov, op=op, xi_seed=int(rng.integers(1 << 31)),
                          precision=precision))
    return specs


def make_points(n, d, clustered, seed, device="cuda"):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    if not clustered:
        return torch.rand((n, d), generator=g, device=device, dtype=torch.float32)
    centres = torch.rand((8, d), generator=g, device=device, dtype=torch.float64)
    widths = torch.exp(torch.empty(8, device=device, dtype=torch.float64).uniform_(math.log(0.01), math.log(0.1),
                                                                                     generator=g))
    lab = torch.randint(0, 8, (n,), generator=g, device=device)
    p = centres[lab] + widths[lab, None] * torch.randn((n, d), generator=g, device=device, dtype=torch.float64)
    return p.float().contiguous()


def make_table(kind, scale, n_bins):
    """Covariance table of the serving mix.  Jitter 1e-3: with 1000 log bins a piecewise-
    linear RBF table at length scale 0.3 is not positive definite on dense neighbourhoods
    at jitter 1e-4 (every value would then be NaN by the reference's rule)."""
    import graphgp_amd as gp

    if kind == "rbf":
        return gp.extras.rbf_kernel(variance=1.0, scale=scale, r_min=1e-5, r_max=2.0, n_bins=n_bins, jitter=1e-3)
    return gp.extras.matern_kernel(p=1, variance=1.0, cutoff=scale, r_min=1e-5, r_max=2.0, n_bins=n_bins,
                                   jitter=1e-3)


def materialize(specs, device="cuda"):
    """Requests with device inputs from ``synthetic_mix`` descriptors.  generate_inv
    requests invert values drawn as N(0,1) (the inverse is defined for any values)."""
    import torch

    reqs, points = [], {}
    for s in specs:
        gs = s["graph"]
        if gs["key"] not in points:  # the caller keeps its points so an evicted graph can be rebuilt
            points[gs["key"]] = make_points(gs["n"], gs["d"], gs["clustered"], gs["seed"], device)
        gen = torch.Generator(device=device)
        gen.manual_seed(s["xi_seed"])
        data = torch.randn(gs["n"], generator=gen, device=device, dtype=torch.float32)
        reqs.append(Request(rid=s["rid"], op=s["op"], graph_key=gs["key"], n=gs["n"], d=gs["d"], n0=gs["n0"],
                            k=gs["k"], covariance=make_table(s["cov"]["kind"], s["cov"]["scale"], s["cov"]["n_bins"]),
                            data=data, arrival=s["arrival"], precision=s["precision"],
                            points=points[gs["key"]],
                            n_bins=s["cov"]["n_bins"]))
    return reqs


def serve_benchmark(n_requests: int = 2000, seed: int = 5, load: float = 0.9,
                    large_cu_fraction: float = 0.5, schedule: bool = False) -> Dict[str, Any]:
    """The serving measurement: the synthetic mix served all at once with no CU
    reservation (batch mode: the device's capacity on this mix), then with Poisson
    arrivals at ``load`` x that request rate under the serving policy (small requests
    reserve CUs while active); fresh servers (no resident graphs) for both runs.
    ``serve_vs_batch`` is all served points over the whole served run (arrivals, plus the
    drain of the last requests) against the batch capacity; ``sustained_vs_batch`` is the
    completion rate over the middle 80% of the arrival window against the same capacity,
    with the offered rate of that window beside it."""
    import torch

    def fresh():  # every run starts from an empty device cache
        import gc

        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()

    reqs = materialize(synthetic_mix(n_requests, rate_per_s=1e9, seed=seed))
    torch.cuda.synchronize()
    srv = Server(large_cu_fraction=1.0, schedule=schedule)
    batch = srv.run(reqs, realtime=False)
    srv.close()
    del srv, reqs
    fresh()
    rate = load * batch["requests"] / batch["wall_s"]
    reqs = materialize(synthetic_mix(n_requests, rate_per_s=rate, seed=seed))
    torch.cuda.synchronize()
    srv = Server(large_cu_fraction=large_cu_fraction, schedule=schedule)
    serve = srv.run(reqs, realtime=True)
    srv.close()
    del srv, reqs
    fresh()
    return {
        "requests": n_requests,
        "arrival_rate_per_s": round(rate, 2),
        "load": load,
        "large_cu_fraction": large_cu_fraction,
        "schedules": schedule,
        "batch_points_per_s": round(batch["points_per_s"], 1),
        "serve_points_per_s": round(serve["points_per_s"], 1),
        "serve_vs_batch": round(serve["points_per_s"] / batch["points_per_s"], 4),
        # completions over the middle 80% of the arrival window, against the batch capacity
        # and beside the points offered in that window
        "sustained_points_per_s": round(serve["sustained_points_per_s"], 1),
        "offered_points_per_s": round(serve["offered_points_per_s"], 1),
        "sustained_vs_batch": round(serve["sustained_points_per_s"] / batch["points_per_s"], 4),
        "p50_ms": round(serve["p50_ms"], 3),
        "p99_ms": round(serve["p99_ms"], 3),
        "p50_ms_n_le_1e6": round(serve["p50_ms_small"], 3),
        "p99_ms_n_le_1e6": round(serve["p99_ms_small"], 3),
        "p99_ms_n_le_1e6_resident_graph": round(serve["p99_ms_small_resident_graph"], 3),
        "graphs_built": serve["graphs_built"],
    }


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="serving benchmark over the synthetic request mix")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--load", type=float, default=0.9)
    ap.add_argument("--large-cu-fraction", type=float, default=0.5)
    ap.add_argument("--schedule", action="store_true", help="build large graphs with their dataflow schedule")
    a = ap.parse_args()
    print(json.dumps(serve_benchmark(a.requests, a.seed, a.load, a.large_cu_fraction, a.schedule)), flush=True)
