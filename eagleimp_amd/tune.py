"""Autotuner for the stream-context path: picks the PBWT kernel variant
(threads per workgroup x chunks per wave, which fixes the workgroup's LDS and
so the resident workgroups per CU) and the targets per submission, by timing
a sample batch of the actual workload through timed contexts (HIP events
between the kernel classes).  Results are cached per panel shape, K range,
window range and candidate set.

    from eagleimp_amd.tune import tune
    best = tune(eng, targets, sample_idx)   # sets eng's PBWT variant
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from .engine import Engine, pbwt_variants

CACHE = os.environ.get("EIMP_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "eagleimp_amd", "tune.json"))


def _key(eng: Engine, Kmax: int, Wmax: int, batches, variants) -> str:
    sig = json.dumps([eng.Nad, eng.Ncorr, eng.M, Kmax, Wmax, list(batches),
                      [(v["threads"], v["chunks_per_wave"]) for v in variants]])
    return hashlib.sha1(sig.encode()).hexdigest()[:16]


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def region_bytes(Mp, K, layout="gpu"):
    """Bytes a target takes in a context's output pool (cref blocks, count0,
    forward + backward PBWT rows, each rounded to 256 bytes)."""
    Mp = np.asarray(Mp, np.int64)
    K = np.asarray(K, np.int64)
    up = lambda x: (x + 255) // 256 * 256  # noqa: E731
    kpad = (K + 63) // 64 * 64
    rs = {"compact": 2 * ((K + 31) // 32), "raw": (K + 511) // 512 * 16}.get(
        layout, 4 * ((K + 511) // 512) * 8)
    return up((Mp + 15) // 16 * kpad * 2) + up(Mp * 4) + up(2 * Mp * rs * 4)


def context_bytes(eng: Engine, batch: int, Wmax: int, Kmax: int, layout: str = "gpu") -> int:
    """Device bytes one stream context holds besides its output pool: free
    memory before and after making a context with a 1 MiB pool."""
    f0, _ = eng.mem_info()
    st = eng.stream(batch, Wmax, Kmax, 1 << 20, timed=True, graph=False, digest=False,
                    **({"compact": True} if layout == "compact" else {}))
    f1, _ = eng.mem_info()
    st.close()
    return max(int(f0) - int(f1) - (1 << 20), 0)


def pool_budget(eng: Engine, contexts: int, batch: int, Wmax: int, Kmax: int,
                layout: str = "gpu", margin: int = 4 << 30) -> int:
    """Device bytes left for output pools once `contexts` contexts' own
    buffers (measured) and a margin are set aside."""
    ctx = context_bytes(eng, batch, Wmax, Kmax, layout)
    free, _ = eng.mem_info()
    return max(int(free) - contexts * ctx - margin, 1 << 28)


def candidates(Kmax: int):
    """PBWT variants whose capacity holds K rounded up to 64."""
    kpad = (Kmax + 63) // 64 * 64
    return [v for v in pbwt_variants() if v["max_k"] >= kpad]


def tune(eng: Engine, targets, idx, batches=(128, 256), reserve_bytes: int | None = None,
         use_cache: bool = True, log=None, layout: str = "gpu", contexts: int = 2) -> dict:
    """Times every (variant, batch) candidate on the sample targets[idx] and
    sets the fastest variant on the engine.  Returns the choice and the table
    of per-target device milliseconds.  layout: the contexts' output rows,
    "gpu" (drop-in) or "compact".  The trial pool is what `contexts` contexts
    sharing one pool would get (pool_budget), or free memory less
    reserve_bytes when that is given."""
    if layout not in ("gpu", "compact"):
        raise ValueError(f"layout {layout!r}: 'gpu' or 'compact'")
    idx = list(idx)
    Kmax = int(min(int(np.max(targets.K[idx])), eng.Ncorr))
    Wmax = int(np.max(targets.W[idx]))
    vs = candidates(Kmax)
    if not vs:
        raise ValueError(f"no PBWT variant holds K = {Kmax}")
    batches = [b for b in batches if b <= len(idx)] or [len(idx)]
    key = _key(eng, Kmax, Wmax, batches, vs) + ("" if layout == "gpu" else "/" + layout)
    db = _load() if use_cache else {}
    if key in db:
        best = db[key]
        eng.set_pbwt_variant(best["variant"])
        best = dict(best, cached=True)
        return best
    def trial_pool(b):
        # what the pool would be with b-target contexts (their own buffers grow with b)
        if reserve_bytes is None:
            return pool_budget(eng, contexts, b, Wmax, Kmax, layout)
        free, _ = eng.mem_info()
        return max(free - reserve_bytes, 1 << 28)

    table = []
    cap = len(idx)  # targets one context's pool holds (learnt from deferrals)
    estimated = False
    for v in vs:
        eng.set_pbwt_variant(v["variant"])
        todo, seen = list(batches), set()
        while todo:
            b = min(todo.pop(0), cap)
            if b in seen:
                continue
            seen.add(b)
            pool = trial_pool(b)
            st = eng.stream(b, Wmax, Kmax, pool, timed=True, graph=False, digest=False,
                            **({"compact": True} if layout == "compact" else {}))
            sub = idx[:b]
            t0 = time.perf_counter()
            st.submit([targets.gt[i] for i in sub], targets.s0[sub], targets.W[sub],
                      targets.K[sub])
            res, nd, _ = st.wait()
            ms = st.kernel_ms()
            wall = time.perf_counter() - t0
            st.close()
            fit = res.Mp > 0
            done = int(fit.sum())
            dev = sum(v_ for k, v_ in ms.items() if k != "copy")
            row = dict(variant=v["variant"], threads=v["threads"],
                       chunks_per_wave=v["chunks_per_wave"], batch=b, targets=done,
                       ms_per_target=dev / max(done, 1), pbwt_ms=ms["pbwt"], cref_ms=ms["cref"],
                       wall_s=round(wall, 3),
                       bytes_per_target=float(region_bytes(res.Mp[fit], res.K[fit],
                                                           layout).mean())
                       if done else 0.0, deferred=b - done)
            table.append(row)
            if log:
                log(f"tune: {row}")
            if 0 < done < b:  # the pool holds fewer: measure at its capacity instead
                # (a smaller submission's contexts leave more for the pool: the
                # first capacity is estimated from the targets' sizes; if that
                # still defers, what it placed is the capacity)
                cap = done
                if not estimated:
                    per = float(region_bytes(res.Mp[fit], res.K[fit], layout).max())
                    if per > 0:
                        cap = max(done, min(int(trial_pool(done) // per), b - 1))
                    estimated = True
                todo.insert(0, cap)
    # submissions that deferred targets did their preparation twice: out
    fitting = [r for r in table if r["deferred"] == 0] or table
    # near-ties (within 2%) go to the lowest variant index, so that repeated
    # runs pick the same kernel; then the fastest submission size of it
    fastest = min(r["ms_per_target"] for r in fitting)
    near = [r for r in fitting if r["ms_per_target"] <= 1.02 * fastest]
    v0 = min(r["variant"] for r in near)
    best = min((r for r in near if r["variant"] == v0), key=lambda r: r["ms_per_target"])
    choice = dict(variant=best["variant"], threads=best["threads"],
                  chunks_per_wave=best["chunks_per_wave"], batch=best["batch"],
                  ms_per_target=best["ms_per_target"],
                  bytes_per_target=max(r["bytes_per_target"] for r in table),
                  table=table, cached=False)
    eng.set_pbwt_variant(choice["variant"])
    if use_cache:
        db[key] = {k: v for k, v in choice.items() if k != "cached"}
        _store(db)
    return choice
