"""Continuous-batching trajectory server on one GPU.

Requests (a seed set, a resident dataset, a step budget, an integrator) arrive over time. The
server keeps every dataset's slices resident in HBM, admits queued requests in arrival order
while their particle state fits the HBM budget, and advances all admitted jobs together in
rounds: each round is one launch of the persistent batched kernel (`gr_batch_run`) in which
every active job runs up to `quantum` steps. Jobs that finish (step budget reached, slices
exhausted, or MaximumSimulationTime reached) are retired -- final positions and statuses are
copied to the host -- and their memory is released, which lets the next requests in. Particle
state stays on the device between rounds, and a job's results do not depend on which other
jobs shared its rounds (each particle's arithmetic is independent of the batch).

Latency of a request = completion time - arrival time (host clock), including queueing,
admission (state allocation + seed upload), all rounds and the final copy.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import capi


@dataclasses.dataclass
class Dataset:
    field: capi.Field
    grid_scale: list
    domain: list
    dtd: float
    nslices: int
    zlevels: object = None


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                  # seconds on the server clock
    dataset: str
    seeds: object                   # (n, 3) float32 numpy array or device tensor, or None ...
    steps: int                      # step budget
    integrator: int = 1             # 0 Euler, 1 RK4
    dt: float = 60.0
    curvature: int = 0
    exprs: tuple = ()
    allow_abs: bool = False
    n: int = 0                      # ... with seed_fn(): seeds made at admission (n of them)
    seed_fn: object = None
    # filled by the server
    admitted: float = float("nan")
    finished: float = float("nan")
    steps_done: int = 0
    positions: np.ndarray = None
    status: np.ndarray = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def count(self) -> int:
        return int(self.seeds.shape[0]) if self.seeds is not None else int(self.n)


class _Active:
    __slots__ = ("req", "job", "window", "time", "time64", "counter", "bytes")

    def __init__(self, req, job, nbytes):
        self.req = req
        self.job = job
        self.window = 0
        self.time = 0.0
        self.time64 = 0.0
        self.counter = 0
        self.bytes = nbytes


def job_bytes(desc, n: int, reorder: bool = True) -> int:
    """Device bytes of a job of `desc` with n particles as the library allocates it (state,
    scratch, one saved column; with `reorder` also the index map and shadow copies a reorder
    adds): gr_job_bytes, the same allocation list gr_job_create uses."""
    return capi.job_bytes(desc, n, 1, reorder)


class Server:
    def __init__(self, ctx: capi.Context, datasets: dict, hbm_budget: int = 64 << 30,
                 quantum: int = 250, block: int = 0, reorder: bool = True, reorder_every: int = 1):
        self.ctx = ctx
        self.datasets = datasets
        self.hbm_budget = int(hbm_budget)
        self.quantum = int(quantum)
        self.block = block
        self.reorder = reorder
        self.reorder_every = max(1, int(reorder_every))
        self.early = 0  # jobs retired once every particle had stopped
        self.queue = collections.deque()
        self.active: list[_Active] = []
        self.done: list[Request] = []
        self.in_use = 0
        self.rounds = 0
        self.particle_steps = 0
        self.device_ms = 0.0
        self.t0 = None

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request):
        self.queue.append(req)

    # -- admission ------------------------------------------------------------------------
    def _desc(self, req: Request):
        ds = self.datasets[req.dataset]
        d = capi.JobDesc()
        d.dim = 3
        d.integrator = req.integrator
        d.curvature = req.curvature
        d.windfield_mode = 1
        d.dt = req.dt
        d.dtd = ds.dtd
        for k in range(3):
            d.grid_scale[k] = ds.grid_scale[k]
            d.domain[k] = ds.domain[k]
        d.save_interval = 0
        d.nprog = len(req.exprs)
        for i, e in enumerate(req.exprs):
            d.prog[i] = capi.compile_expr(e, dim=3, curvature=req.curvature,
                                          allow_abs=req.allow_abs)
        return d

    def _make_job(self, req: Request, d):
        seeds = req.seeds if req.seeds is not None else req.seed_fn()
        job = capi.Job(self.ctx, d, seeds)
        if self.reorder and req.count >= 4096:
            job.reorder()
        return job

    def admit(self):
        while self.queue and self.queue[0].arrival <= self.now():
            req = self.queue[0]
            d = self._desc(req)
            nb = job_bytes(d, req.count, self.reorder and req.count >= 4096)
            if self.in_use + nb > self.hbm_budget and self.active:
                break  # wait for memory (FIFO: no overtaking)
            self.queue.popleft()
            job = self._make_job(req, d)
            req.admitted = self.now()
            self.active.append(_Active(req, job, nb))
            self.in_use += nb

    # -- one round ------------------------------------------------------------------------
    def round(self):
        spans = []
        for a in self.active:
            left = a.req.steps - a.req.steps_done
            spans.append((a.job, self.datasets[a.req.dataset].field,
                          dict(first_window=a.window, max_steps=min(left, self.quantum),
                               time_start=a.time, time_start64=a.time64,
                               counter_start=a.counter)))
        ends, _ = self.ctx.batch_run(spans, block=self.block, timed=False, ends=True)
        self.rounds += 1
        keep = []
        for a, e in zip(self.active, ends):
            a.req.steps_done += e["steps"]
            self.particle_steps += e["steps"] * a.req.count
            a.window, a.time, a.time64, a.counter = e["window"], e["time"], e["time64"], e["counter"]
            if e["done"] or e["steps"] == 0 or a.req.steps_done >= a.req.steps:
                self._retire(a)
            elif a.job.live_range() == 0:
                # every particle has stopped: the remaining steps change nothing
                a.req.steps_done = a.req.steps
                self.early += 1
                self._retire(a)
            else:
                keep.append(a)
        self.active = keep
        # regroup the large jobs by cell (every round by default): their launches then cover
        # only the particles still alive (the live range) with wavefronts of neighbouring
        # particles, and a job whose particles have all stopped retires at once
        if self.reorder and self.rounds % self.reorder_every == 0:
            big = [a.job for a in self.active if a.req.count >= 4096]
            if big:
                self.ctx.reorder(big)

    def _retire(self, a: _Active):
        req = a.req
        req.positions = a.job.read(capi.OUT_POSITIONS)
        req.status = a.job.read(capi.OUT_STATUS)
        a.job.close()
        self.in_use -= a.bytes
        req.finished = self.now()
        self.done.append(req)

    # -- driver ---------------------------------------------------------------------------
    def run(self, requests, deadline: float | None = None, progress=None, every: float = 10.0):
        """Serves `requests` (sorted by arrival) in real time; returns the finished ones.
        progress(text) (e.g. a print to stderr) receives a status line every `every` seconds."""
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        self.ctx.take_kernel_ms()  # device time of this run's launches, summed at the end
        self.t0 = time.perf_counter()
        next_report = every
        while self.queue or self.active:
            if deadline is not None and self.now() > deadline:
                break
            if progress is not None and self.now() >= next_report:
                progress(f"t={self.now():.0f}s done={len(self.done)} active={len(self.active)} "
                         f"queued={len(self.queue)} rounds={self.rounds}")
                next_report += every
            self.admit()
            if self.active:
                self.round()
            elif self.queue:
                wait = self.queue[0].arrival - self.now()
                if wait > 0:
                    time.sleep(min(wait, 0.05))
        self.ctx.synchronize()
        ms, _ = self.ctx.take_kernel_ms()
        self.device_ms += ms
        return self.done

    def report(self) -> dict:
        lat = np.array([r.latency for r in self.done]) if self.done else np.array([np.nan])
        wall = self.now() if self.t0 is not None else float("nan")
        return {
            "jobs": len(self.done),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(np.mean(lat)),
            "particle_steps": int(self.particle_steps),
            "particle_steps_per_s": self.particle_steps / wall if wall > 0 else float("nan"),
            "rounds": self.rounds,
            "retired_all_stopped": self.early,
            "device_busy_frac": (self.device_ms / 1e3) / wall if wall > 0 else float("nan"),
            "wall_s": wall,
        }


def poisson_arrivals(seed: int, rate: float, duration: float) -> np.ndarray:
    """Arrival times of a Poisson process (exponential gaps from SplitMix64 uniforms)."""
    from .synthetic import uniform01
    n_max = int(rate * duration * 3 + 100)
    u = uniform01(seed, n_max)
    gaps = -np.log1p(-u) / rate
    t = np.cumsum(gaps)
    return t[t < duration]
