"""Ensemble workloads on one GPU: a fixed batch of uneven meshes (all advancing together)
and a served stream of hydro runs with continuous batching.

Serving: runs arrive at Poisson times; each is admitted into the device ensemble as soon
as a slot and enough HBM (its footprint) are free, keeps its state resident until its own
final time, and is retired when done; admitted runs advance together in batched launches,
so a run that arrives while others are on the GPU overlaps with them.
"""
import heapq
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import cases
from . import types as T


def batch_cases(n=256, seed=20261202, lo=64, hi=512, cycles=300):
    """2D bi-material tubes with seeded states; sides log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        c = cases.riemann2d(side, side, rng, stiffened=False, cycles=cycles, name=f"batch{i}_{side}")
        out.append(c)
    return out


def serve_requests(n=64, seed=20261202, rate=4.0, t_end=0.1, lo2=128, hi2=1024, lo3=32, hi3=160):
    """Poisson arrivals (rate per second) of 2D / 3D stiffened-gas Riemann runs."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() < 0.5:
            side = int(round(np.exp(rng.uniform(np.log(lo2), np.log(hi2)))))
            c = cases.riemann2d(side, side, rng, stiffened=True, t_end=t_end, name=f"req{i}_2d{side}")
        else:
            side = int(round(np.exp(rng.uniform(np.log(lo3), np.log(hi3)))))
            c = cases.riemann3d(side, rng, stiffened=True, t_end=t_end, name=f"req{i}_3d{side}")
        reqs.append((t, c))
    return reqs


def run_batch(batch, graph=True, max_slots=None, budget=200e9, poll_every=30, tune_cycles=12):
    """All cases in one ensemble, timed end to end: the creation of every case (device memory and
    initial state), the autotuner's timed iterations (the runs' own iterations), every other
    iteration, the completion polls and the release.  Returns (seconds, per-case iterations,
    per-case statuses, per-case completion times from the start, persistent blocks per CU)."""
    from .lib import Ensemble
    ncyc = batch[0].max_cycles
    t0 = time.perf_counter()
    E = Ensemble(max_slots or len(batch), budget)
    slots = [E.add(c, max_iterations=ncyc) for c in batch]
    assert all(s is not None for s in slots), "batch exceeds the HBM budget"
    E.synchronize()
    phases = {"create_s": time.perf_counter() - t0}
    bpc = E.autotune(tune_cycles)[0] if tune_cycles else None
    phases["autotune_s"] = time.perf_counter() - t0 - phases["create_s"]
    done = {}
    while len(done) < len(slots):
        E.step(poll_every, graph=graph)
        n, st, tm, it = E.poll()
        now = time.perf_counter() - t0
        for i, s in enumerate(slots):
            if i not in done and (st[s] == T.MH_STOPPED or st[s] < 0):
                done[i] = now
    iters = [int(it[s]) for s in slots]
    stats = [int(st[s]) for s in slots]
    t_run = time.perf_counter()
    E.close()
    phases["iterate_s"] = t_run - t0 - phases["create_s"] - phases["autotune_s"]
    phases["release_s"] = time.perf_counter() - t_run
    run_batch.phases = phases
    return time.perf_counter() - t0, iters, stats, [done[i] for i in range(len(slots))], bpc


def run_back_to_back(batch, graph=6):
    """Each case alone on the GPU, one after the other (the single-mesh path), timed on the
    same basis as run_batch: creation, graph capture, iterations and release of every case.
    Returns (seconds, per-case completion times from the start)."""
    from .lib import GpuCase
    t0 = time.perf_counter()
    done = []
    iterate = 0.
    for c in batch:
        g = GpuCase(c)
        g.synchronize()
        t1 = time.perf_counter()
        g.run(c.max_cycles, graph=graph)
        iterate += time.perf_counter() - t1
        g.close()
        done.append(time.perf_counter() - t0)
    run_back_to_back.iterate_s = iterate  # the iterations alone (graph captures included)
    return time.perf_counter() - t0, done


@dataclass
class Served:
    arrival: float
    admitted: float = -1.0
    done: float = -1.0
    iterations: int = 0
    slot: int = -1
    status: int = 0
    live_at_admission: int = 0  # runs already on the GPU when this one was admitted


def serve(reqs, budget=64e9, max_slots=64, step_cycles=6, graph=True, on_done=None):
    """Continuous-batching request loop; times are wall-clock seconds from the first arrival.

    on_done(i, ensemble, slot), when given, is called for request i once it has finished and
    before its slot is released (to read its fields).  A request whose footprint exceeds the
    whole budget is refused with MH_ERR_NOMEM (iterations -1) instead of waiting forever."""
    from .lib import Ensemble, footprint
    E = Ensemble(max_slots, budget)
    recs: List[Served] = [Served(a) for a, _ in reqs]
    queue = list(range(len(reqs)))  # in arrival order
    live = {}
    t0 = time.perf_counter() - reqs[0][0]  # the clock starts at the first arrival
    now = lambda: time.perf_counter() - t0
    # Slot states are read one step late: snapshot j is queued after step j and read after
    # step j+1, so a slot admitted after snapshot j was queued is not in it (admitted_at).
    nsnap = 0  # snapshots queued so far
    admitted_at = {}
    while queue or live:
        t = now()
        # admit every arrived request that fits (FIFO, by HBM footprint and slots)
        while queue and reqs[queue[0]][0] <= t:
            i = queue[0]
            slot = E.add(reqs[i][1])
            if slot is None:
                if not live and footprint(reqs[i][1]) > budget:
                    recs[i].admitted = recs[i].done = now()
                    recs[i].status, recs[i].iterations = T.MH_ERR_NOMEM, -1
                    queue.pop(0)
                    continue
                break
            recs[i].admitted = now()
            recs[i].slot = slot
            recs[i].live_at_admission = len(live)
            live[slot] = i
            admitted_at[slot] = nsnap
            queue.pop(0)
        if not live:
            if not queue:
                break
            # idle until the next arrival
            wait = reqs[queue[0]][0] - now()
            if wait > 0:
                time.sleep(min(wait, 0.05))
            continue
        # queue the next iterations, then read the state as it was before them: the GPU works
        # while the host retires and admits
        E.step(step_cycles, graph=graph)
        snap = nsnap - 1  # the snapshot in flight, if any
        if snap >= 0:
            n, st, tm, it = E.poll_result()
        E.poll_async()
        nsnap += 1
        if snap < 0:
            continue
        for slot in list(live):
            if admitted_at[slot] > snap:
                continue  # admitted after that snapshot was queued
            if st[slot] == T.MH_STOPPED or st[slot] < 0:
                i = live.pop(slot)
                recs[i].done = now()
                if on_done is not None:
                    on_done(i, E, slot)
                recs[i].iterations = int(it[slot])
                recs[i].status = int(st[slot])
                if st[slot] < 0:
                    recs[i].iterations = -1
                E.remove(slot)
    E.close()
    return recs


def serve_back_to_back(reqs, graph=6):
    """Baseline: the same stream served one run at a time on the single-mesh path."""
    from .lib import GpuCase
    t0 = time.perf_counter() - reqs[0][0]
    recs = []
    for a, c in reqs:
        while time.perf_counter() - t0 < a:
            time.sleep(0.001)
        g = GpuCase(c)
        while True:
            g.queue(60, graph=graph)
            st = g.synchronize()
            if st != T.MH_OK:
                break
        r = Served(a, admitted=a, done=time.perf_counter() - t0, iterations=g.iteration)
        g.close()
        recs.append(r)
    return recs


def latency_stats(recs):
    tts = np.array([r.done - r.arrival for r in recs])
    span = max(r.done for r in recs) - min(r.arrival for r in recs)
    return {"runs": len(recs), "runs_per_s": len(recs) / span, "p50_s": float(np.percentile(tts, 50)),
            "p99_s": float(np.percentile(tts, 99)), "span_s": span}
