"""Snapshots and restart files in the reference's ``.npz`` layout.

Keys follow phrike/io.py:103-236: ``t, U, x, y[, z], rho, ux, uy[, uz], p,
meta{Nx, Ny[, Nz], Lx, Ly[, Lz], gamma}, created``, file name
``snapshot_t{t:.6f}[_{tag}].npz``.  ``meta`` is stored as the reference stores
it (a dict, which numpy pickles) so the reference's ``load_checkpoint`` reads
these files; a JSON copy ``meta_json`` is stored as well, and
``load_checkpoint`` here reads only plain arrays (``allow_pickle=False``).

The state is staged by a device-to-device copy on the compute stream; the
stage goes to pinned host buffers on a side stream and the file is written by
a background thread (``blocking=False``), so the compute stream is not
stalled and may overwrite U at once; ``SnapshotWriter.wait()`` joins pending
writes.
"""
from __future__ import annotations

import json
import os
import threading
from datetime import datetime
from typing import Any, Dict, List, Optional

import numpy as np
import torch


def _host(a):
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def snapshot_name(outdir: str, t: float, tag: Optional[str] = None) -> str:
    ts = f"t{t:.6f}"
    if tag:
        ts = f"{ts}_{tag}"
    return os.path.join(outdir, f"snapshot_{ts}.npz")


def _meta(grid, equations) -> Dict[str, Any]:
    m = {"Nx": grid.Nx, "Ny": grid.Ny, "Lx": grid.Lx, "Ly": grid.Ly, "gamma": equations.gamma}
    if hasattr(grid, "Nz"):
        m.update({"Nz": grid.Nz, "Lz": grid.Lz})
    # reference key order: Nx, Ny, Nz, Lx, Ly, Lz, gamma
    order = ["Nx", "Ny", "Nz", "Lx", "Ly", "Lz", "gamma"]
    return {k: m[k] for k in order if k in m}


class SnapshotWriter:
    """Asynchronous snapshot writer (one side stream, background threads)."""

    def __init__(self, device=None):
        self.stream = torch.cuda.Stream(device=device)
        self.threads: List[threading.Thread] = []

    def save(self, outdir: str, t: float, U, grid, equations, tag: Optional[str] = None,
             blocking: bool = True) -> str:
        os.makedirs(outdir, exist_ok=True)
        fname = snapshot_name(outdir, t, tag)
        is3 = U.shape[0] == 5
        # stage on the compute stream (device copy), so later work on that
        # stream may overwrite U while the side stream copies the stage out
        stage = U.detach().clone()
        prim = [v.contiguous() for v in equations.primitive(stage)]
        ev = torch.cuda.Event()
        ev.record()
        host = {}
        with torch.cuda.stream(self.stream):
            self.stream.wait_event(ev)
            names = ["rho", "ux", "uy", "uz", "p"] if is3 else ["rho", "ux", "uy", "p"]
            for k, v in zip(names, prim):
                h = torch.empty(v.shape, dtype=v.dtype, pin_memory=True)
                h.copy_(v, non_blocking=True)
                host[k] = h
            hU = torch.empty(U.shape, dtype=U.dtype, pin_memory=True)
            hU.copy_(stage, non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.stream)
        # keep the device stage alive until the copies have run
        for t_ in [stage, *prim]:
            t_.record_stream(self.stream)
        coords = {"x": _host(grid.x), "y": _host(grid.y)}
        if is3:
            coords["z"] = _host(grid.z)
        meta = _meta(grid, equations)

        def write():
            done.synchronize()
            np.savez(fname, t=t, U=hU.numpy(), **coords, **{k: v.numpy() for k, v in host.items()},
                     meta=meta, meta_json=json.dumps(meta), created=str(datetime.utcnow()))

        if blocking:
            write()
        else:
            th = threading.Thread(target=write, daemon=True)
            th.start()
            self.threads.append(th)
        return fname

    def wait(self):
        for th in self.threads:
            th.join()
        self.threads.clear()


_default_writer: Optional[SnapshotWriter] = None


def save_solution_snapshot(outdir: str, t: float, U, grid, equations, tag: Optional[str] = None,
                           tracers=None, blocking: bool = True) -> str:
    """phrike/io.py:103-236 for the hip backend (no tracers on this path)."""
    global _default_writer
    if tracers is not None:
        raise NotImplementedError("tracers are not part of the hip backend")
    if _default_writer is None:
        _default_writer = SnapshotWriter(getattr(grid, "device", None))
    return _default_writer.save(outdir, t, U, grid, equations, tag, blocking)


def load_checkpoint(path: str) -> Dict[str, Any]:
    """Read a snapshot written here (or by the reference, when it carries
    plain arrays) without unpickling anything: t, U, x/y/z, primitive
    variables, and meta from ``meta_json`` when present
    (phrike/io.py:32-83 layout of the returned dict)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Checkpoint file not found: {path}")
    with np.load(path, allow_pickle=False) as d:
        out: Dict[str, Any] = {"t": float(d["t"]), "U": d["U"]}
        out["meta"] = json.loads(str(d["meta_json"])) if "meta_json" in d.files else {}
        for k in ("x", "y", "z"):
            if k in d.files:
                out[k] = d[k]
        out["primitive_vars"] = {k: d[k] for k in ("rho", "u", "ux", "uy", "uz", "p") if k in d.files}
    return out
