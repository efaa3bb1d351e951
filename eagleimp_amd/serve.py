"""Serving loop over the device engine: the panel stays resident, request
batches of targets arrive over time and are admitted into stream contexts as
contexts and device memory free up (continuous batching); latency per request
is arrival -> last target's results collected.

Admission is by device memory twice over: the contexts' output pools are cut
from the free HBM reported by the device at start-up, and inside a context
the device planner places targets in the pool by their actual output sizes,
deferring those that do not fit; deferred targets go back to the head of the
queue.  Several contexts run on their own HIP streams, so the work of
requests admitted at different times overlaps on the device.

    python -m eagleimp_amd.serve --rate 50 --requests 200      # workload 5

The reference has no serving mode: its GPU handler drains one pre-filled
queue of all targets (src/GPUHandler.cpp:104-186, src/PBWTPhaser.cpp:478-499).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .engine import Engine, Stream


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the run
    targets: list                  # indices into the target pool
    results: dict = field(default_factory=dict)  # position -> (S, Mp, ncalls, K, dig, dig0)
    finish: float | None = None

    @property
    def latency(self):
        return None if self.finish is None else self.finish - self.arrival


class Server:
    """Continuous batching of target requests over `contexts` stream contexts."""

    def __init__(self, eng: Engine, pool, max_window: int, max_K: int, contexts: int = 3,
                 max_batch: int = 256, reserve_bytes: int = 16 << 30, pool_bytes: int | None = None):
        self.eng = eng
        self.pool = pool                    # synth.Targets: gt, s0, W, K
        self.max_batch = max_batch
        free, _ = eng.mem_info()
        if pool_bytes is None:
            pool_bytes = max((free - reserve_bytes) // contexts, 1 << 28)
        self.pool_bytes = pool_bytes
        self.ctx: list[Stream] = [eng.stream(max_batch, max_window, max_K, pool_bytes)
                                  for _ in range(contexts)]
        self.queue: deque = deque()         # (request, position)
        self.busy: dict = {}                # context -> [(request, position)]
        self.submissions = 0
        self.deferred = 0

    def close(self):
        for st in self.ctx:
            st.close()
        self.ctx = []

    def add(self, req: Request):
        self.queue.extend((req, p) for p in range(len(req.targets)))

    def idle(self) -> bool:
        return not self.queue and not self.busy

    def _submit(self, ci: int):
        n = min(self.max_batch, len(self.queue))
        items = [self.queue.popleft() for _ in range(n)]
        idx = [req.targets[p] for req, p in items]
        P = self.pool
        self.ctx[ci].submit([P.gt[i] for i in idx], P.s0[idx], P.W[idx], P.K[idx])
        self.busy[ci] = items
        self.submissions += 1

    def _collect(self, ci: int, now):
        items = self.busy.pop(ci)
        res, nd, _ = self.ctx[ci].wait()
        t = now()
        back = []
        for j, (req, p) in enumerate(items):
            if int(res.Mp[j]) == 0:        # did not fit the pool: admit again later
                back.append((req, p))
                continue
            req.results[p] = (int(res.S[j]), int(res.Mp[j]), int(res.ncalls[j]), int(res.K[j]),
                              int(res.dig_pbwt[j]), int(res.dig_count0[j]))
            if len(req.results) == len(req.targets):
                req.finish = t
        if back and len(back) == len(items):
            raise RuntimeError("a target's outputs exceed a whole context pool "
                               f"({self.pool_bytes} bytes)")
        self.deferred += len(back)
        self.queue.extendleft(reversed(back))

    def poll(self, now) -> bool:
        """Collects finished contexts and refills free ones; True if anything moved."""
        moved = False
        for ci, st in enumerate(self.ctx):
            if ci in self.busy and st.done():
                self._collect(ci, now)
                moved = True
            if ci not in self.busy and self.queue:
                self._submit(ci)
                moved = True
        return moved


def run_requests(server: Server, reqs: list[Request], realtime: bool = True) -> float:
    """Feeds requests at their arrival times (realtime) or all at once, until
    every request is complete.  Returns the wall time of the run."""
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0

    i = 0
    n = len(reqs)
    while i < n or not server.idle():
        t = now()
        while i < n and (not realtime or reqs[i].arrival <= t):
            if not realtime:  # all admitted now: latency counts from here
                reqs[i].arrival = t
            server.add(reqs[i])
            i += 1
        if not server.poll(now):
            wait = 0.0005
            if i < n and server.idle():
                wait = max(0.0, reqs[i].arrival - now())
            time.sleep(min(wait, 0.0005) if server.busy else wait)
    return now()


def make_requests(t, size, first, npool) -> list[Request]:
    return [Request(i, float(t[i]), [int((first[i] + k) % npool) for k in range(int(size[i]))])
            for i in range(len(t))]


def summarize(reqs: list[Request], wall: float) -> dict:
    lat = np.array([r.latency for r in reqs], np.float64) * 1e3
    nt = sum(len(r.targets) for r in reqs)
    return {
        "requests": len(reqs), "targets": nt,
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "mean_ms": round(float(lat.mean()), 2), "max_ms": round(float(lat.max()), 2),
        "targets_per_s": round(nt / wall, 2), "wall_s": round(wall, 3),
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--rate", type=float, nargs="+", default=[20.0, 200.0],
                    help="arrival rates in request batches per second")
    ap.add_argument("--requests", type=int, default=100, help="requests per rate")
    ap.add_argument("--contexts", type=int, default=2)
    ap.add_argument("--max-batch", type=int, default=512, help="most targets per submission")
    ap.add_argument("--pool", type=int, default=None, help="targets in the pool (default 1024)")
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    from . import synth
    c, panel, pool = synth.make_workload(5, T=args.pool)
    eng = Engine(0)
    eng.initialize(2, panel.M, panel.N, panel.N, panel.site)
    eng.set_panel_extras(panel.cM, hap=panel.hap)
    srv = Server(eng, pool, int(pool.W.max()), int(pool.K.max()), contexts=args.contexts,
                 max_batch=args.max_batch)
    out = {"workload": c["name"], "contexts": args.contexts, "runs": []}
    # warm-up pass (allocations, first launches)
    warm = make_requests([0.0], [min(64, len(pool.W))], [0], len(pool.W))
    run_requests(srv, warm, realtime=False)
    for rate in args.rate:
        t, size, first = synth.make_arrivals(args.seed, args.requests, rate, len(pool.W),
                                             c["bmin"], c["bmax"])
        reqs = make_requests(t, size, first, len(pool.W))
        wall = run_requests(srv, reqs)
        s = summarize(reqs, wall)
        s["rate_batches_per_s"] = rate
        s["offered_targets_per_s"] = round(float(size.sum()) / float(t[-1]), 1)
        out["runs"].append(s)
        print(json.dumps(s), file=sys.stderr, flush=True)
    srv.close()
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
