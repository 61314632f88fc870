"""Continuous-batching ensemble server for 2D Fourier Kelvin-Helmholtz runs.

Each request is one run of SpectralSolver2D semantics (phrike/solver.py:527-812:
RK4 with positivity clamp, post-step filter every step, dt from the CFL
condition recomputed after every step, last step clipped to t_end) with its own
size, seed and t_end.  Runs of one size share a *bucket*: a fixed number of
slots whose states live in HBM as (slots, 4, N, N), plus per-slot FP64 t,
t_end, dt and an active flag on the device.  A bucket advances all of its
slots by S steps with one replay of a captured HIP graph:

    per step:  dt[r] = cfl*h/max(|ux|+a)        (ph_max_wave_speed2d)
               dt[r] = min(dt[r], t_end[r]-t[r]) (ph_runs_clip)
               U' = filter(RK4(U, dt[r]))       (ph_step2d, inactive slots untouched)
               t[r] += dt[r]; active[r] = t[r] < t_end[r] - 1e-12  (ph_runs_advance)

so no host synchronisation happens inside a replay.  Buckets replay on their
own streams, so different sizes overlap on the device.  Between replays the
host polls (non-blocking) the flags of finished replays, retires completed
runs, and admits queued requests into free slots.  The slot count of every
bucket is fixed up front from an HBM budget (admission control).

A run's arithmetic does not depend on the other slots (every kernel and
reduction is per run), so a result is bit-identical alone and in a mix.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .equations import EulerEquations2D
from .grid import Grid2D
from .initial_conditions import khi_2d_periodic
from .solver import SpectralSolver2D


@dataclass
class RunRequest:
    rid: int
    n: int
    seed: int
    t_end: float
    arrival: float = 0.0  # seconds after the server starts


@dataclass
class RunResult:
    rid: int
    n: int
    t_end: float
    arrival: float
    admitted: float
    finished: float
    steps: int
    t_final: float
    state: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def failed(self) -> bool:
        """The run diverged: a NaN wave speed gave a NaN dt, which retires the
        run with t = NaN (the reference's while t < t_end loop ends the same
        way, phrike/solver.py:764)."""
        return not np.isfinite(self.t_final)


def slot_bytes(n: int, precision: str = "single") -> int:
    """HBM per slot: two state buffers + the per-run share of the plan workspace
    (8 spectral fields of N x mxp complex, stage state, reduction partials)."""
    s = 4 if precision == "single" else 8
    mx = n // 3 + 1
    al = 16 if s == 4 else 8
    mxp = (mx + al - 1) // al * al
    return 2 * 4 * n * n * s + 8 * n * mxp * 2 * s + 4 * n * n * s + 256 * 5 * 8


def capacities_for_budget(sizes: Dict[int, float], budget_bytes: int, precision: str = "single",
                          max_slots: int = 64) -> Dict[int, int]:
    """Split the HBM budget over the sizes in proportion to weight * N^2 (their
    share of the work) and convert to slot counts (at least one each)."""
    share = {n: w * n * n for n, w in sizes.items()}
    tot = sum(share.values())
    caps = {}
    for n in sizes:
        b = budget_bytes * share[n] / tot
        caps[n] = int(max(1, min(max_slots, b // slot_bytes(n, precision))))
    return caps


class Bucket:
    def __init__(self, n: int, slots: int, precision: str, cfl: float, steps_per_replay: int,
                 filter_params: Optional[Dict], device):
        self.n = n
        self.cap = slots
        self.S = steps_per_replay + (steps_per_replay % 2)  # even: every replay starts from buffer 0
        self.dev = device
        fp = dict(filter_params or {"enabled": True, "p": 16, "alpha": 36.0})
        fp["interval"] = 1
        self.grid = Grid2D(n, n, 1.0, 1.0, dealias=True, filter_params=fp, backend="hip",
                           precision=precision, batch=slots, torch_device=str(device))
        self.solver = SpectralSolver2D(self.grid, EulerEquations2D(1.4), cfl=cfl)
        dt = self.grid.dtype
        self.U = [torch.zeros((slots, 4, n, n), dtype=dt, device=device) for _ in range(2)]
        self.t = torch.zeros(slots, dtype=torch.float64, device=device)
        self.t_end = torch.zeros(slots, dtype=torch.float64, device=device)
        self.dtv = torch.zeros(slots, dtype=torch.float64, device=device)
        self.active = torch.zeros(slots, dtype=torch.int32, device=device)
        self.steps = torch.zeros(slots, dtype=torch.int32, device=device)
        self.host_active = torch.zeros(slots, dtype=torch.int32).pin_memory()
        self.host_steps = torch.zeros(slots, dtype=torch.int32).pin_memory()
        self.host_t = torch.zeros(slots, dtype=torch.float64).pin_memory()
        # slots: free (None) or occupied by a request
        self.occupant: List[Optional[RunRequest]] = [None] * slots
        self.admitted_at = [0.0] * slots
        self.stream = torch.cuda.Stream(device=device)
        self.graph = None
        self.event = None
        self.inflight = False
        self.grid.workspace()  # allocate before capture

    # -- device program ------------------------------------------------------
    def _one_step(self, src, dst):
        plan, ws, st = self.solver._ctx()
        nb = self.cap
        _lib.call("ph_max_wave_speed2d", plan, nb, _lib.ptr(src), None, _lib.ptr(self.dtv),
                  self.solver.cfl, _lib.ptr(self.active), ws, st)
        _lib.call("ph_runs_clip", nb, _lib.ptr(self.t), _lib.ptr(self.t_end), _lib.ptr(self.dtv),
                  _lib.ptr(self.active), st)
        _lib.call("ph_step2d", plan, nb, _lib.ptr(src), _lib.ptr(dst), 0.0, _lib.ptr(self.dtv),
                  _lib.ptr(self.active), 1, ws, st)
        _lib.call("ph_runs_advance", nb, _lib.ptr(self.t), _lib.ptr(self.t_end), _lib.ptr(self.dtv),
                  _lib.ptr(self.active), _lib.ptr(self.steps), st)

    def _program(self):
        for k in range(self.S):
            self._one_step(self.U[k % 2], self.U[(k + 1) % 2])

    def capture(self):
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._program()
        torch.cuda.synchronize(self.dev)
        self.graph = g

    def launch(self, use_graph: bool = True):
        with torch.cuda.stream(self.stream):
            if use_graph:
                if self.graph is None:
                    self.capture()
                self.graph.replay()
            else:
                self._program()
            self.host_active.copy_(self.active, non_blocking=True)
            self.host_steps.copy_(self.steps, non_blocking=True)
            self.host_t.copy_(self.t, non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record(self.stream)
        self.inflight = True

    # -- slots ---------------------------------------------------------------
    def free_slots(self) -> List[int]:
        return [i for i, o in enumerate(self.occupant) if o is None]

    def admit(self, req: RunRequest, U0: np.ndarray, now: float):
        i = self.free_slots()[0]
        with torch.cuda.stream(self.stream):
            src = torch.from_numpy(np.ascontiguousarray(U0)).to(self.grid.dtype).pin_memory()
            self.U[0][i].copy_(src, non_blocking=True)
            self.t[i] = 0.0
            self.t_end[i] = float(req.t_end)
            self.steps[i] = 0
            self.active[i] = 1
        self.occupant[i] = req
        self.admitted_at[i] = now

    def n_active(self) -> int:
        return sum(o is not None for o in self.occupant)


class EnsembleServer:
    """Serve RunRequests with continuous batching; see the module docstring."""

    def __init__(self, capacities: Dict[int, int], precision: str = "single", cfl: float = 0.4,
                 steps_per_replay: int = 16, filter_params: Optional[Dict] = None,
                 device: str = "cuda:0", use_graph: bool = True, keep_states: bool = False):
        self.dev = torch.device(device)
        self.precision = precision
        self.use_graph = use_graph
        self.keep_states = keep_states
        self.buckets = {n: Bucket(n, c, precision, cfl, steps_per_replay, filter_params, self.dev)
                        for n, c in sorted(capacities.items())}

    def hbm_bytes(self) -> int:
        return sum(slot_bytes(b.n, self.precision) * b.cap for b in self.buckets.values())

    def serve(self, requests: List[RunRequest], realtime: bool = True, ic=None) -> List[RunResult]:
        """Run every request to completion.  With realtime, requests become
        visible at their arrival times (wall clock since the call); otherwise all
        are queued at once.  ``ic(req)`` gives the initial state (default: the
        periodic KHI profile with the request's seed)."""
        ic = ic or (lambda r: khi_2d_periodic(r.n, r.n, seed=r.seed))
        for r in requests:
            if r.n not in self.buckets:
                raise ValueError(f"no bucket for size {r.n}")
        arrivals = deque(sorted(requests, key=lambda r: r.arrival))
        queues = {n: deque() for n in self.buckets}
        results: List[RunResult] = []
        start = time.perf_counter()
        while arrivals or any(queues.values()) or any(b.n_active() for b in self.buckets.values()):
            now = time.perf_counter() - start
            while arrivals and (not realtime or arrivals[0].arrival <= now):
                r = arrivals.popleft()
                queues[r.n].append(r)
            progressed = False
            for n, b in self.buckets.items():
                # retire finished runs of a completed replay
                if b.inflight and b.event.query():
                    b.inflight = False
                    t_done = time.perf_counter() - start
                    for i, occ in enumerate(b.occupant):
                        if occ is not None and int(b.host_active[i]) == 0:
                            steps = int(b.host_steps[i])
                            state = None
                            if self.keep_states:
                                state = b.U[steps % 2][i].double().cpu().numpy()
                            results.append(RunResult(occ.rid, n, occ.t_end, occ.arrival, b.admitted_at[i],
                                                     t_done, steps, float(b.host_t[i]), state))
                            b.occupant[i] = None
                    progressed = True
                if not b.inflight:
                    while queues[n] and b.free_slots():
                        r = queues[n].popleft()
                        b.admit(r, ic(r), time.perf_counter() - start)
                        progressed = True
                    if b.n_active():
                        b.launch(self.use_graph)
                        progressed = True
            if not progressed:
                if arrivals and realtime and not any(b.inflight for b in self.buckets.values()):
                    time.sleep(max(0.0, min(0.002, arrivals[0].arrival - (time.perf_counter() - start))))
                else:
                    time.sleep(0.0002)
        return sorted(results, key=lambda r: r.rid)


def poisson_requests(rate: float, window: float, sizes: Dict[int, float], t_end_range=(0.5, 2.0),
                     seed: int = 0) -> List[RunRequest]:
    """Poisson arrivals (exponential gaps, mean 1/rate) over `window` seconds;
    size drawn with the given weights, own seed, t_end uniform in range."""
    rng = np.random.default_rng(seed)
    ns = np.array(sorted(sizes))
    w = np.array([sizes[n] for n in ns], dtype=float)
    w /= w.sum()
    out, t, rid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window:
            break
        out.append(RunRequest(rid, int(rng.choice(ns, p=w)), int(rng.integers(0, 2 ** 31 - 1)),
                              float(rng.uniform(*t_end_range)), t))
        rid += 1
    return out


def latency_summary(results: List[RunResult], wall: float) -> Dict[str, float]:
    lat = np.array([r.latency for r in results])
    cells = sum(r.n * r.n * r.steps for r in results)
    return {"completed": len(results), "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "max_s": float(lat.max()) if len(lat) else None,
            "mcell_updates_per_s": cells / wall / 1e6 if wall > 0 else None, "wall_s": wall}
