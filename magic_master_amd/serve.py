"""Serving loop for scenario requests: continuous batching on one GPU.

Requests (a grid, a background column, forcing parameters, an end time in
model seconds and a snapshot cadence) arrive over wall-clock time.  They are
admitted first come first served whenever their device footprint fits the
memory budget next to the scenarios already running; every admitted scenario
is a slot of one MagicBatch (FP32 perturbation form), so all of them advance
together in each graph replay, and a request that arrives while others run
joins them at the next replay.  At every snapshot time a scenario stops its
interval (claw3's clip to the interval end), its state is copied to the host
and handed to a writer thread (HDF5 frames with the reference's out3 layout,
or kept in memory), and the next interval starts; at its end time the
scenario leaves the batch and its memory is freed.

Reported per request: arrival, admission and completion times (queue wait
and time to completion), accepted and rejected steps, and the global mass
and energy at every snapshot (compensated FP64 sums); for the run: p50/p99
of both times, and aggregate cell-updates per second over the run.
"""
from __future__ import annotations

import os
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import problem as P


@dataclass
class Request:
    rid: int
    arrival: float              # wall-clock seconds after the start of the run
    claw: P.ClawData
    prob: P.ProbData
    grid: P.Grid
    column: np.ndarray          # background aux(1:maux, k) for every k
    t_end: float                # model time to reach (s)
    cadence: float = 60.0       # model seconds between snapshots
    q0: Optional[np.ndarray] = None   # initial q (None: at rest on the background)


@dataclass
class Outcome:
    rid: int
    arrival: float
    admitted: float = float("nan")
    completed: float = float("nan")
    steps: int = 0
    rejects: int = 0
    snapshots: List[tuple] = field(default_factory=list)   # (model time, path or None)
    sums: List[tuple] = field(default_factory=list)        # (model time, mass kg, energy J) per snapshot
    final_q: Optional[np.ndarray] = None
    status: int = 0

    @property
    def queue_wait(self):
        return self.admitted - self.arrival

    @property
    def time_to_completion(self):
        return self.completed - self.arrival


def footprint_bytes(g: P.Grid) -> int:
    """Device memory of one scenario of the FP32 form in a batch (sweeps in
    turn, the batch's default): q' and its backup (2 x 4 B x meqn), the
    source pass's scratch (5 x 4 B), the FP64 staging of q for upload and
    download (8 B x meqn), per padded cell, plus the columns."""
    cells = (g.mx + 2 * g.mbc) * (g.my + 2 * g.mbc) * (g.mz + 2 * g.mbc)
    return cells * (8 * g.meqn + 20 + 8 * g.meqn) + 12 * (g.mz + 2 * g.mbc) * 8


class _PinnedPool:
    """Page-locked snapshot buffers (hip.PinnedBuffer), reused, at most
    cap_bytes in all (one buffer beyond that when none is free, so that a
    snapshot larger than the cap still runs)."""

    def __init__(self, cap_bytes: float):
        self.cap = cap_bytes
        self.free = []
        self.total = 0
        self.cv = threading.Condition()

    def get(self, nbytes: int):
        from .hip import PinnedBuffer
        with self.cv:
            while True:
                for k, b in enumerate(self.free):
                    if b.nbytes >= nbytes:
                        return self.free.pop(k)
                if self.total + nbytes <= self.cap or self.total == 0:
                    self.total += nbytes
                    return PinnedBuffer(nbytes)
                if self.free:          # too small to reuse: make room
                    b = self.free.pop(0)
                    self.total -= b.nbytes
                    b.close()
                    continue
                self.cv.wait()         # every buffer is with the writer

    def put(self, b):
        with self.cv:
            self.free.append(b)
            self.cv.notify_all()

    def close(self):
        with self.cv:
            for b in self.free:
                b.close()
            self.free = []


class _Writer(threading.Thread):
    """Finishes snapshots on the host while the GPU keeps stepping: waits for
    each snapshot's copy (queued by the serving loop with snapshot_begin),
    writes it as an out3 frame (hdf5) or drops it (download), keeps a
    request's final state when asked, and returns the buffer to the pool."""

    def __init__(self, outdir: Optional[str], pool: _PinnedPool):
        super().__init__(daemon=True)
        self.outdir = outdir
        self.pool = pool
        self.q = queue.Queue()
        self.error = None

    def run(self):
        from . import h5
        from .hip import snapshot_wait
        while True:
            item = self.q.get()
            if item is None:
                return
            handle, buf, g, t, frame, rid, keep = item
            try:
                snapshot_wait(handle)
                qh = buf.array(g.shape_q)
                if self.outdir:
                    d = os.path.join(self.outdir, f"req{rid:05d}")
                    os.makedirs(d, exist_ok=True)
                    h5.out3(g, qh, t, frame, d)
                if keep is not None:
                    keep.final_q = qh.copy()
            except Exception as e:   # reported by Server.run
                self.error = e
            finally:
                self.pool.put(buf)


class Server:
    def __init__(self, budget_bytes: float = 250e9, check_every: int = 4, snapshots: str = "hdf5",
                 outdir: Optional[str] = None, keep_final: bool = False, tile_shapes=None,
                 progress: Optional[float] = None, pinned_bytes: float = 8e9):
        """snapshots: "hdf5" (frames written under outdir), "download" (copied
        to the host and dropped) or "none".  Snapshots are copied to
        page-locked buffers (at most pinned_bytes of them) while the batch
        keeps stepping, and a writer thread finishes them.  tile_shapes: the
        sweeps' tile shapes (x, y, z) for the whole run (None: the library's
        default), fixed so that a request's result does not depend on when
        it ran.  progress: print a status line every that many wall-clock
        seconds."""
        if snapshots not in ("hdf5", "download", "none"):
            raise ValueError(snapshots)
        if snapshots == "hdf5" and not outdir:
            raise ValueError("hdf5 snapshots need outdir")
        self.budget = budget_bytes
        self.check_every = check_every
        self.snapshots = snapshots
        self.outdir = outdir
        self.keep_final = keep_final
        self.tile_shapes = tile_shapes
        self.progress = progress
        self.pinned_bytes = pinned_bytes

    def run(self, requests: List[Request]) -> dict:
        from .hip import MagicBatch
        reqs = sorted(requests, key=lambda r: r.arrival)
        if not reqs:
            return {"outcomes": []}
        g0 = reqs[0].grid
        batch = MagicBatch([], [], [], meqn=g0.meqn, mbc=g0.mbc, maux=g0.maux)
        if self.tile_shapes is not None:
            batch.set_tile_shapes(self.tile_shapes)
        pool = _PinnedPool(self.pinned_bytes)
        writer = (_Writer(self.outdir if self.snapshots == "hdf5" else None, pool)
                  if self.snapshots != "none" else None)
        if writer:
            writer.start()
        out = {r.rid: Outcome(r.rid, r.arrival) for r in reqs}
        pending = list(reqs)
        waiting: List[Request] = []
        active = {}                  # slot -> (request, frame, footprint)
        used = 0.0
        cells_steps = 0
        max_active = 0
        max_used = 0.0
        t0 = time.perf_counter()
        shown = 0.0
        try:
            while pending or waiting or active:
                now = time.perf_counter() - t0
                if self.progress and now - shown >= self.progress:
                    shown = now
                    done = len(reqs) - len(pending) - len(waiting) - len(active)
                    print(f"  [{now:7.1f} s] done {done}, running {len(active)}, waiting {len(waiting)}, "
                          f"not yet arrived {len(pending)}", flush=True)
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.pop(0))
                while waiting:
                    r = waiting[0]
                    fp = footprint_bytes(r.grid)
                    if fp > self.budget:
                        raise ValueError(f"request {r.rid} needs {fp:.3g} B, more than the budget")
                    if used + fp > self.budget:
                        break
                    waiting.pop(0)
                    s = batch.add(r.grid, r.prob, r.column, r.q0)
                    c = r.claw
                    batch.start(s, 0.0, min(r.cadence, r.t_end), c.dtv[0], c.dtv, c.cflv, 1 << 30, c.mthbc)
                    active[s] = (r, 0, fp)
                    used += fp
                    out[r.rid].admitted = time.perf_counter() - t0
                max_active = max(max_active, len(active))
                max_used = max(max_used, used)
                if not active:
                    if pending:
                        time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                    continue
                batch.run(max_attempts=self.check_every, check_every=self.check_every)
                for s in list(active):
                    r, frame, fp = active[s]
                    st = batch.state(s)
                    if st["status"] != 0:
                        out[r.rid].status = st["status"]
                        out[r.rid].completed = time.perf_counter() - t0
                        batch.remove(s)
                        del active[s]
                        used -= fp
                        continue
                    if st["active"]:
                        continue
                    # end of an interval: a snapshot
                    g = r.grid
                    cells_steps += g.mx * g.my * g.mz * st["nsteps"]
                    out[r.rid].steps += st["nsteps"]
                    out[r.rid].rejects += st["rejects"]
                    frame += 1
                    last = st["t"] >= r.t_end
                    path = None
                    if writer:
                        buf = pool.get(8 * int(np.prod(g.shape_q)))
                        handle = batch.snapshot_begin(s, buf)
                        keep = out[r.rid] if (self.keep_final and last) else None
                        writer.q.put((handle, buf, g, st["t"], frame, r.rid, keep))
                        if self.snapshots == "hdf5":
                            from . import h5
                            path = os.path.join(self.outdir, f"req{r.rid:05d}", h5.frame_name(frame))
                    elif self.keep_final and last:
                        out[r.rid].final_q = batch.get_q(s)
                    out[r.rid].snapshots.append((st["t"], path))
                    gs = batch.global_sums(s)   # compensated FP64 sums of the full state
                    out[r.rid].sums.append((st["t"], float(gs[0]), float(gs[4])))
                    if last:
                        out[r.rid].completed = time.perf_counter() - t0
                        batch.remove(s)
                        del active[s]
                        used -= fp
                    else:
                        c = r.claw
                        batch.start(s, st["t"], min(st["t"] + r.cadence, r.t_end), st["dt"], c.dtv, c.cflv,
                                    1 << 30, c.mthbc)
                        active[s] = (r, frame, fp)
            makespan = time.perf_counter() - t0
        finally:
            if writer:
                writer.q.put(None)
                writer.join()
            batch.close()
            pool.close()
        if writer and writer.error:
            raise writer.error
        oc = [out[r.rid] for r in reqs]
        ttc = np.array([o.time_to_completion for o in oc])
        qw = np.array([o.queue_wait for o in oc])
        return {
            "outcomes": oc,
            "requests": len(oc),
            "makespan_s": makespan,
            "mcell_updates_per_s": cells_steps / makespan / 1e6,
            "ttc_p50_s": float(np.percentile(ttc, 50)), "ttc_p99_s": float(np.percentile(ttc, 99)),
            "queue_p50_s": float(np.percentile(qw, 50)), "queue_p99_s": float(np.percentile(qw, 99)),
            "max_concurrent": max_active,
            "max_footprint_bytes": max_used,
        }


def poisson_requests(rate: float, window: float, seed: int = 7, t_end=(600.0, 6000.0),
                     cadence: float = 60.0, nmin=(64, 64, 96), nmax=(200, 200, 300)) -> List[Request]:
    """Requests arriving as a Poisson process of the given rate (per second)
    over window seconds; each scenario is drawn like the ensemble workload
    (problem.ensemble) with an end time uniform in t_end (model seconds)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    times = []
    t = float(rng.exponential(1.0 / rate))
    while t < window:
        times.append(t)
        t += float(rng.exponential(1.0 / rate))
    scen = P.ensemble(len(times), seed=seed + 1, nmin=nmin, nmax=nmax)
    return [Request(k, times[k], c, p, g, col, float(rng.uniform(*t_end)), cadence)
            for k, (c, p, g, col) in enumerate(scen)]


def main(argv=None):
    """Serving runs at several arrival rates: Poisson requests drawn like the
    ensemble workload, one Server run per rate, one summary line per rate
    (requests, makespan, aggregate cell-updates/s, p50/p99 of time to
    completion and queue wait, peak concurrency and footprint)."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m magic_master_amd.serve", description=main.__doc__)
    ap.add_argument("--rates", default="0.5,2,8", help="arrival rates, requests per wall-clock second")
    ap.add_argument("--window", type=float, default=600.0, help="arrival window, wall-clock seconds")
    ap.add_argument("--t-end", default="600,6000", help="request end time range, model seconds")
    ap.add_argument("--cadence", type=float, default=60.0, help="snapshot interval, model seconds")
    ap.add_argument("--budget", type=float, default=250e9, help="admission budget, bytes")
    ap.add_argument("--snapshots", default="download", choices=("hdf5", "download", "none"))
    ap.add_argument("--outdir", default=None, help="HDF5 frames go here (--snapshots hdf5)")
    ap.add_argument("--shapes", default=None, help="tile shapes of the x, y, z sweeps, e.g. 1,0,0")
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args(argv)
    t_end = tuple(float(v) for v in a.t_end.split(","))
    shapes = [int(v) for v in a.shapes.split(",")] if a.shapes else None
    for rate in [float(v) for v in a.rates.split(",")]:
        rq = poisson_requests(rate, a.window, seed=a.seed, t_end=t_end, cadence=a.cadence)
        print(f"rate {rate} req/s: {len(rq)} requests over {a.window} s", flush=True)
        res = Server(budget_bytes=a.budget, snapshots=a.snapshots, tile_shapes=shapes, progress=30.0,
                     outdir=os.path.join(a.outdir, f"rate{rate:g}") if a.outdir else None).run(rq)
        failed = [o.rid for o in res["outcomes"] if o.status != 0]
        line = {"rate": rate, "window_s": a.window, "t_end": list(t_end), "failed": failed}
        line.update({k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items() if k != "outcomes"})
        line["steps"] = int(sum(o.steps for o in res["outcomes"]))
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
