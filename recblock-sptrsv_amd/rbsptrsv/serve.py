"""Serving loop: resident factors, memory-based admission, continuous
cross-request batching, latency percentiles.

Factors are preprocessed once into HBM (one ``Plan`` each).  Requests
(factor id, right-hand side, arrival time) enter a FIFO admission queue; a
request is admitted when the device bytes it needs (its b and x, 16 bytes
per row) fit next to the bytes of requests in flight.  Admitted requests
wait per factor; whenever a factor's stream has finished its previous batch,
all its waiting requests (up to ``max_batch``) go out as one multi-column
solve.  Each factor's solves are ordered on one stream (a plan runs one
solve at a time); different factors run on different streams and overlap on
the device.  Requests that arrive while batches run join the next batch of
their factor: batching is continuous, not a fixed window.

Completion times come from GPU events recorded after each batch on its
stream, measured against an event recorded when the run starts, so latency =
completion - arrival includes queueing, admission and the solve.

No counterpart in the reference, which solves one right-hand side of one
factor per call.
"""
import collections
import time

import numpy as np

from .api import Plan


class Request:
    __slots__ = ("rid", "factor", "b", "arrival", "x", "done_ms", "batch")

    def __init__(self, rid, factor, b, arrival):
        self.rid, self.factor, self.b, self.arrival = rid, factor, b, arrival
        self.x = None
        self.done_ms = None
        self.batch = 0


class Server:
    """``factors``: objects with colptr/rowidx/val/substitution (as for
    ``Plan.from_factor``).  ``mode``: "fp64" (plain solves) or "f32ir" (FP32
    solves refined in FP64 to ``tol``).  ``mem_budget``: device bytes for
    request buffers in flight.  ``streams``: streams the factors are spread
    over (default priority).  ``max_iter``: corrections queued per f32ir
    batch (the loop runs on the device without a host check; corrections
    after convergence replay empty walks, so the count is kept near the
    usual two; a column that needs more than max_iter makes check()
    raise).  ``log``: called with a progress line after each plan is built
    and warmed up."""

    def __init__(self, factors, mode="fp64", tol=1e-12, max_batch=64,
                 mem_budget=1 << 30, streams=4, lv=-1, tune_rhs=8, log=None,
                 max_iter=4):
        import torch
        if mode not in ("fp64", "f32ir"):
            raise ValueError("mode is fp64 or f32ir")
        self.torch = torch
        self.mode, self.tol, self.max_iter = mode, tol, max_iter
        self.max_batch, self.mem_budget = max_batch, mem_budget
        self.log = log
        self.plans = []
        for j, f in enumerate(factors):
            t0 = time.perf_counter()
            self.plans.append(Plan.from_factor(f, lv=lv, tune_rhs=tune_rhs,
                                               mixed=(mode == "f32ir")))
            if log:
                log("plan %d/%d built (n %d, %.2f s)" %
                    (j + 1, len(factors), self.plans[-1].m,
                     time.perf_counter() - t0))
        self.n = [p.m for p in self.plans]
        self.streams = [torch.cuda.Stream() for _ in range(max(1, streams))]
        self.warmup()

    def warmup(self):
        """Batches of every width up to max_batch (powers of two) per
        factor, so that the first requests do not pay for one-time work
        (multi-column buffers, graph capture)."""
        torch = self.torch
        for f in range(len(self.plans)):
            t0 = time.perf_counter()
            z = [torch.zeros(self.n[f], dtype=torch.float64, device="cuda")
                 for _ in range(self.max_batch)]
            reqs = [Request(-1, f, b, 0.0) for b in z]
            k = self.max_batch
            while k >= 1:
                self._launch(f, reqs[:k])
                k //= 2
            torch.cuda.synchronize()
            if self.log:
                self.log("factor %d warmed up (%.2f s)" %
                         (f, time.perf_counter() - t0))

    def close(self):
        for p in self.plans:
            p.close()

    def _stream(self, f):
        return self.streams[f % len(self.streams)]

    def _launch(self, f, reqs):
        """One multi-column solve of factor f's requests on its stream."""
        torch = self.torch
        st = self._stream(f)
        with torch.cuda.stream(st):
            B = torch.stack([r.b for r in reqs], 0).t()   # column-major
            X = torch.empty((len(reqs), self.n[f]), dtype=torch.float64,
                            device="cuda").t()
            if self.mode == "fp64":
                self.plans[f].solve(B, X, stream=st)
            else:
                # no report: the loop runs on the device without host
                # synchronisation; a column short of tol shows in check()
                self.plans[f].solve_f32ir(B, X, tol=self.tol,
                                          max_iter=self.max_iter, stream=st,
                                          report=False)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(st)
        for j, r in enumerate(reqs):
            r.x = X[:, j]
            r.batch = len(reqs)
        return ev, B, X

    def run(self, requests, spin_s=20e-6):
        """Serve ``requests`` (list of (factor, b ndarray, arrival seconds
        from the start)).  Returns (stats dict, list of Request in input
        order)."""
        torch = self.torch
        reqs = [Request(i, f, b, a) for i, (f, b, a) in enumerate(requests)]
        order = sorted(reqs, key=lambda r: r.arrival)
        # b moves to the device when the request is admitted
        host_b = {r.rid: np.ascontiguousarray(r.b, np.float64) for r in reqs}
        nf = len(self.plans)
        waiting = collections.deque()             # arrived, not admitted
        ready = [collections.deque() for _ in range(nf)]
        busy = [None] * nf                        # (event, reqs, B, X)
        in_flight = 0
        nxt = 0
        done = 0
        batches = 0
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        start.record()
        t0 = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(order) and order[nxt].arrival <= now:
                waiting.append(order[nxt])
                nxt += 1
            # admission by memory, first come first served
            while waiting:
                r = waiting[0]
                need = 16 * self.n[r.factor]
                if in_flight + need > self.mem_budget and in_flight > 0:
                    break
                waiting.popleft()
                in_flight += need
                with torch.cuda.stream(self._stream(r.factor)):
                    r.b = torch.as_tensor(host_b[r.rid], device="cuda")
                ready[r.factor].append(r)
            # retire finished batches, launch waiting work
            for f in range(nf):
                if busy[f] is not None and busy[f][0].query():
                    ev, rs, _, _ = busy[f]
                    # the batch's answers are only handed out when the
                    # plan's watchdog is clear (raises RbError otherwise)
                    self.plans[f].check(self._stream(f))
                    t = start.elapsed_time(ev)
                    for r in rs:
                        r.done_ms = t
                        in_flight -= 16 * self.n[f]
                    done += len(rs)
                    busy[f] = None
                if busy[f] is None and ready[f]:
                    k = min(len(ready[f]), self.max_batch)
                    rs = [ready[f].popleft() for _ in range(k)]
                    ev, B, X = self._launch(f, rs)
                    busy[f] = (ev, rs, B, X)
                    batches += 1
            if spin_s:
                time.sleep(spin_s)
        torch.cuda.synchronize()
        makespan_ms = max(r.done_ms for r in reqs)
        lat = np.array([r.done_ms - 1e3 * r.arrival for r in reqs])
        stats = {
            "requests": len(reqs),
            "batches": batches,
            "mean_batch": len(reqs) / max(1, batches),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()),
            "solves_per_s": len(reqs) / (makespan_ms * 1e-3),
            "makespan_ms": makespan_ms,
        }
        return stats, reqs


def poisson_requests(nfactors, sizes, count, rate, seed=0, zipf_a=1.1):
    """``count`` requests at a Poisson rate (per second); factor chosen
    Zipf(zipf_a) over ``nfactors`` (rank 1 most popular); b ~ U(-1,1)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    w = 1.0 / np.arange(1, nfactors + 1) ** zipf_a
    fac = rng.choice(nfactors, size=count, p=w / w.sum())
    return [(int(f), rng.uniform(-1, 1, sizes[f]), float(a))
            for f, a in zip(fac, arrivals)]
