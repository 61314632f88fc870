"""Continuous-batching ensemble server for 2D Fourier Kelvin-Helmholtz runs.

Each request is one run of SpectralSolver2D semantics (phrike/solver.py:527-812:
RK4 with positivity clamp, post-step filter every step, dt from the CFL
condition recomputed after every step, last step clipped to t_end) with its own
size, seed and t_end.  Runs of one size share a *bucket*: slots whose states
live in HBM as (slots, 4, N, N), with per-slot FP64 t, dt, t_end and an int32
active flag and step count on the device (the ph_run2d scalars).  A bucket
advances all of its slots by S steps with one replay of a captured HIP graph
of one ph_run2d call, whose steps each do, per active slot,

    dt = min(dt, t_end - t); U' = filter(RK4(U, dt)); dt_next = CFL(U');
    t += dt; dt = dt_next; active = t < t_end - 1e-12

so no host synchronisation happens inside a replay.  Buckets replay on their
own streams, so different sizes overlap on the device.  Between replays the
host polls (non-blocking) the flags of finished replays, retires completed
runs and admits queued requests into free slots (their first dt comes from
ph_max_wave_speed2d on the admitted slots alone).

Admission by memory: with ``hbm_budget`` the buckets draw their slots from one
pool of HBM.  Between replays a bucket whose queue exceeds its free slots
grows (its buffers are reallocated with the active runs copied over and its
graph captured again) as far as the pool allows; when the free part of the
pool does not suffice, mostly idle buckets give slots back, so memory left
idle by small runs admits a queued large run.  With ``capacities`` the slot
counts are fixed instead.

A run's arithmetic does not depend on the other slots (every kernel and
reduction is per run), so a result is bit-identical alone and in a mix, in
any slot and at any bucket capacity.  A run whose wave speed turns NaN
retires with t = NaN and is reported as failed.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

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
    """HBM per slot: two state buffers + the per-run share of the plan
    workspace (8 spectral fields of N x mxp complex, a stage state,
    reduction partials) + the run scalars."""
    s = 4 if precision == "single" else 8
    mx = n // 3 + 1
    al = 16 if s == 4 else 8
    mxp = (mx + al - 1) // al * al
    return 2 * 4 * n * n * s + 8 * n * mxp * 2 * s + 4 * n * n * s + 256 * 12 * 8 + 64


def capacities_for_budget(sizes: Dict[int, float], budget_bytes: int, precision: str = "single",
                          max_slots: int = 64) -> Dict[int, int]:
    """A fixed split of the budget over the sizes in proportion to weight * N^2
    (their share of the work), as slot counts (at least one each)."""
    share = {n: w * n * n for n, w in sizes.items()}
    tot = sum(share.values())
    caps = {}
    for n in sizes:
        b = budget_bytes * share[n] / tot
        caps[n] = int(max(1, min(max_slots, b // slot_bytes(n, precision))))
    return caps


def _pow2_ceil(k: int) -> int:
    return 1 if k <= 1 else 1 << (int(k) - 1).bit_length()


class Bucket:
    """The slots of one run size; capacity 0 until the first admission when
    the server manages memory."""

    def __init__(self, n: int, cap: int, precision: str, cfl: float, steps_per_replay: int,
                 filter_params: Optional[Dict], device):
        self.n = n
        self.precision = precision
        self.cfl = cfl
        self.S = steps_per_replay + (steps_per_replay % 2)  # even: every replay starts from buffer 0
        self.dev = device
        fp = dict(filter_params or {"enabled": True, "p": 16, "alpha": 36.0})
        fp["interval"] = 1
        self.fp = fp
        self.stream = torch.cuda.Stream(device=device)
        self.cap = 0
        self.occupant: List[Optional[RunRequest]] = []
        self.admitted_at: List[float] = []
        self.graph = None
        self.event = None
        self.inflight = False
        self.resizes = 0
        self.n_fresh = 0
        if cap > 0:
            self.resize(cap)

    # -- buffers -------------------------------------------------------------
    def resize(self, cap: int) -> None:
        """New capacity (not while a replay is in flight): new buffers and
        plan, the active runs copied into the first slots, graph captured
        again on the next launch."""
        assert not self.inflight
        cap = int(cap)
        keep = [i for i, o in enumerate(self.occupant) if o is not None]
        if cap < len(keep):
            raise ValueError("capacity below the number of active runs")
        old = (getattr(self, "U", None), getattr(self, "rs", None), getattr(self, "rf", None))
        old_cap = self.cap
        occ, adm = self.occupant, self.admitted_at
        self.graph = None
        self.cap = cap
        self.occupant = [None] * cap
        self.admitted_at = [0.0] * cap
        self.resizes += 1
        if cap == 0:
            for k in ("grid", "solver", "U", "rs", "rf", "fresh"):
                setattr(self, k, None)
            return
        n = self.n
        self.grid = Grid2D(n, n, 1.0, 1.0, dealias=True, filter_params=self.fp, backend="hip",
                           precision=self.precision, batch=cap, torch_device=str(self.dev))
        self.solver = SpectralSolver2D(self.grid, EulerEquations2D(1.4), cfl=self.cfl)
        dt = self.grid.dtype
        with torch.cuda.stream(self.stream):
            self.U = [torch.zeros((cap, 4, n, n), dtype=dt, device=self.dev) for _ in range(2)]
            self.rs = torch.zeros(_lib.PH_RS_N * cap, dtype=torch.float64, device=self.dev)
            self.rf = torch.zeros(_lib.PH_RF_N * cap, dtype=torch.int32, device=self.dev)
            self.fresh = torch.zeros(cap, dtype=torch.int32, device=self.dev)
            if keep:
                oU, ors, orf = old
                src = torch.tensor(keep, dtype=torch.long, device=self.dev)
                m = len(keep)
                self.U[0][:m].copy_(oU[0].index_select(0, src))
                for f in range(_lib.PH_RS_N):
                    self.rs[f * cap:f * cap + m].copy_(ors[f * old_cap:(f + 1) * old_cap].index_select(0, src))
                for f in range(_lib.PH_RF_N):
                    self.rf[f * cap:f * cap + m].copy_(orf[f * old_cap:(f + 1) * old_cap].index_select(0, src))
        for j, i in enumerate(keep):
            self.occupant[j] = occ[i]
            self.admitted_at[j] = adm[i]
        self.host_rf = torch.zeros(_lib.PH_RF_N * cap, dtype=torch.int32).pin_memory()
        self.host_rs = torch.zeros(_lib.PH_RS_N * cap, dtype=torch.float64).pin_memory()
        self.grid.workspace()  # allocate before capture
        self.stream.synchronize()
        self.host_rf.copy_(self.rf)  # the carried runs' step counts
        self.host_rs.copy_(self.rs)

    def rs_field(self, f):
        return self.rs[f * self.cap:(f + 1) * self.cap]

    def rf_field(self, f):
        return self.rf[f * self.cap:(f + 1) * self.cap]

    # -- device program ------------------------------------------------------
    def _program(self):
        plan, ws, st = self.solver._ctx()
        _lib.call("ph_run2d", plan, self.cap, _lib.ptr(self.U[0]), _lib.ptr(self.U[1]), self.S,
                  _lib.ptr(self.rs), _lib.ptr(self.rf), self.cfl, ws, st)

    def capture(self):
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._program()
        torch.cuda.synchronize(self.dev)
        self.graph = g

    def launch(self, use_graph: bool = True):
        self.steps_before = [int(x) for x in self.host_rf[_lib.PH_RF_STEPS * self.cap:(_lib.PH_RF_STEPS + 1) *
                                                          self.cap]]
        with torch.cuda.stream(self.stream):
            if use_graph:
                if self.graph is None:
                    self.capture()
                self.graph.replay()
            else:
                self._program()
            self.host_rf.copy_(self.rf, non_blocking=True)
            self.host_rs.copy_(self.rs, non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record(self.stream)
        self.inflight = True

    # -- slots ---------------------------------------------------------------
    def free_slots(self) -> List[int]:
        return [i for i, o in enumerate(self.occupant) if o is None]

    def admit(self, req: RunRequest, U0: np.ndarray, now: float):
        i = self.free_slots()[0]
        cap = self.cap
        with torch.cuda.stream(self.stream):
            src = torch.from_numpy(np.ascontiguousarray(U0)).to(self.grid.dtype).pin_memory()
            self.U[0][i].copy_(src, non_blocking=True)
            self.rs[_lib.PH_RS_T * cap + i] = 0.0
            self.rs[_lib.PH_RS_TEND * cap + i] = float(req.t_end)
            self.rs[_lib.PH_RS_TSTOP * cap + i] = float("inf")
            self.rf[_lib.PH_RF_STEPS * cap + i] = 0
            self.rf[_lib.PH_RF_ACTIVE * cap + i] = 1
            self.fresh[i] = 1
        self.n_fresh += 1
        self.host_rf[_lib.PH_RF_STEPS * cap + i] = 0
        self.occupant[i] = req
        self.admitted_at[i] = now

    def start_admitted(self):
        """First dt (CFL of the initial state) of the slots admitted since the
        last replay; the other slots' scalars are left as they are."""
        if self.n_fresh == 0:
            return
        plan, ws, st = self.solver._ctx()
        with torch.cuda.stream(self.stream):
            _lib.call("ph_max_wave_speed2d", plan, self.cap, _lib.ptr(self.U[0]), None,
                      _lib.ptr(self.rs_field(_lib.PH_RS_DT)), self.cfl, _lib.ptr(self.fresh), ws,
                      _lib.stream_handle(torch, self.dev))
            self.fresh.zero_()
        self.n_fresh = 0

    def n_active(self) -> int:
        return sum(o is not None for o in self.occupant)


class EnsembleServer:
    """Serve RunRequests with continuous batching; see the module docstring.

    Either ``capacities`` ({N: slots}, fixed) or ``hbm_budget`` (bytes; buckets
    for ``sizes`` grow and shrink between replays within it)."""

    def __init__(self, capacities: Optional[Dict[int, int]] = None, precision: str = "single", cfl: float = 0.4,
                 steps_per_replay: int = 16, filter_params: Optional[Dict] = None, device: str = "cuda:0",
                 use_graph: bool = True, keep_states: bool = False, hbm_budget: Optional[int] = None,
                 sizes: Sequence[int] = (128, 256, 512, 1024), max_slots: int = 64, min_grow: int = 4):
        self.dev = torch.device(device)
        self.precision = precision
        self.use_graph = use_graph
        self.keep_states = keep_states
        self.budget = hbm_budget
        self.max_slots = int(max_slots)
        self.min_grow = max(1, int(min_grow))
        if capacities is None and hbm_budget is None:
            raise ValueError("give capacities or hbm_budget")
        if capacities is not None:
            self.fixed = True
            items = sorted(capacities.items())
        else:
            self.fixed = False
            items = [(int(n), 0) for n in sorted(sizes)]
        self.buckets = {n: Bucket(n, c, precision, cfl, steps_per_replay, filter_params, self.dev)
                        for n, c in items}
        self.peak_bytes = self.hbm_bytes()

    def hbm_bytes(self) -> int:
        return sum(slot_bytes(b.n, self.precision) * b.cap for b in self.buckets.values())

    def _reclaim(self, need: int, skip: int, queues) -> int:
        """Shrink idle buckets (no queue, no replay in flight, at most a
        quarter occupied) until `need` bytes are free; returns the bytes
        freed.  Memory is only taken back under pressure, so buckets do not
        shrink and regrow (each resize captures the bucket's graph again)."""
        freed = 0
        for n, b in sorted(self.buckets.items(), key=lambda kv: -slot_bytes(kv[0], self.precision) * kv[1].cap):
            if freed >= need:
                break
            if n == skip or b.inflight or queues[n] or b.cap == 0:
                continue
            occ = b.n_active()
            new = 0 if occ == 0 else _pow2_ceil(occ)
            if new < b.cap and occ <= b.cap // 4:
                freed += (b.cap - new) * slot_bytes(n, self.precision)
                b.resize(new)
        return freed

    def _rebalance(self, queues) -> bool:
        """Pool mode: a bucket whose queue exceeds its free slots grows
        (powers of two, at least ``min_grow`` slots, at most max_slots) as
        far as the budget allows, taking memory back from idle buckets only
        when the free part of the budget does not suffice.  Only buckets
        without a replay in flight are resized."""
        if self.fixed:
            return False
        changed = False
        for n, b in sorted(self.buckets.items(), key=lambda kv: -len(queues[kv[0]])):
            if b.inflight or not queues[n]:
                continue
            occ = b.n_active()
            need = occ + len(queues[n])
            if need <= b.cap:
                continue
            per = slot_bytes(n, self.precision)
            target = min(self.max_slots, max(self.min_grow, _pow2_ceil(need)))
            if target <= b.cap:
                continue
            free = self.budget - self.hbm_bytes()
            if (target - b.cap) * per > free:
                free += self._reclaim((target - b.cap) * per - free, n, queues)
            while target > b.cap and (target - b.cap) * per > free:
                target //= 2
            if target <= b.cap:
                # no power of two fits: any growth that does
                target = min(need, self.max_slots, b.cap + int(free // per))
            if target > b.cap:
                b.resize(target)
                changed = True
        self.peak_bytes = max(self.peak_bytes, self.hbm_bytes())
        return changed

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
        RF_ACT, RF_STEPS, RS_T = _lib.PH_RF_ACTIVE, _lib.PH_RF_STEPS, _lib.PH_RS_T
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
                    cap = b.cap
                    for i, occ in enumerate(b.occupant):
                        if occ is not None and int(b.host_rf[RF_ACT * cap + i]) == 0:
                            steps = int(b.host_rf[RF_STEPS * cap + i])
                            state = None
                            if self.keep_states:
                                state = b.U[(steps - b.steps_before[i]) % 2][i].double().cpu().numpy()
                            results.append(RunResult(occ.rid, n, occ.t_end, occ.arrival, b.admitted_at[i],
                                                     t_done, steps, float(b.host_rs[RS_T * cap + i]), state))
                            b.occupant[i] = None
                    progressed = True
            if self._rebalance(queues):
                progressed = True
            for n, b in self.buckets.items():
                if not b.inflight:
                    while queues[n] and b.free_slots():
                        r = queues[n].popleft()
                        b.admit(r, ic(r), time.perf_counter() - start)
                        progressed = True
                    if b.n_active():
                        b.start_admitted()
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
    return {"completed": len(results), "failed": sum(r.failed for r in results),
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "max_s": float(lat.max()) if len(lat) else None,
            "mcell_updates_per_s": cells / wall / 1e6 if wall > 0 else None, "wall_s": wall}
