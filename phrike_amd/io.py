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


def _meta(grid, equations, mhd: bool = False) -> Dict[str, Any]:
    m = {"Nx": grid.Nx, "Ny": grid.Ny, "Lx": grid.Lx, "Ly": grid.Ly, "gamma": equations.gamma}
    if hasattr(grid, "Nz"):
        m.update({"Nz": grid.Nz, "Lz": grid.Lz})
    if mhd:  # phrike/io.py:128-159 key order
        m["mhd"] = True
        order = ["mhd", "gamma", "Lx", "Nx", "Ny", "Nz", "Ly", "Lz"]
    else:  # phrike/io.py:171-207: Nx, Ny, Nz, Lx, Ly, Lz, gamma
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
        is_mhd = U.shape[0] == 8
        is3 = U.shape[0] == 5 or is_mhd
        # stage on the compute stream (device copy), so later work on that
        # stream may overwrite U while the side stream copies the stage out
        stage = U.detach().clone()
        prim = [v.contiguous() for v in equations.primitive(stage)]
        extra = {}
        if is_mhd:  # phrike/io.py:122-167: B and a div(B) diagnostic
            extra["div_b"] = (grid.dx1(stage[5]) + grid.dy1(stage[6]) + grid.dz1(stage[7])).contiguous()
        ev = torch.cuda.Event()
        ev.record()
        host = {}
        with torch.cuda.stream(self.stream):
            self.stream.wait_event(ev)
            names = ["rho", "ux", "uy", "uz", "p"] if is3 else ["rho", "ux", "uy", "p"]
            if is_mhd:
                names += ["Bx", "By", "Bz"]
            for k, v in extra.items():
                prim.append(v)
                names.append(k)
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
        meta = _meta(grid, equations, is_mhd)

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


def _recover_meta(d, files) -> Dict[str, Any]:
    """Grid and gamma of a snapshot whose meta is a pickled dict (the
    reference's files), from its plain arrays: sizes from U's shape, lengths
    from the uniform x/y/z coordinates (x = linspace(0, L, N, endpoint=False)),
    gamma from E = p/(gamma-1) + kinetic (+ magnetic) energy."""
    U = d["U"]
    meta: Dict[str, Any] = {}
    spatial = U.shape[1:]
    names = ["Nx", "Ny", "Nz"][:len(spatial)]
    for k, n in zip(names, reversed(spatial)):
        meta[k] = int(n)
    for ax, (k, L) in zip("xyz", (("Nx", "Lx"), ("Ny", "Ly"), ("Nz", "Lz"))):
        if ax in files and k in meta and meta[k] > 1:
            c = np.asarray(d[ax], dtype=np.float64)
            meta[L] = float(meta[k] * (c[1] - c[0]))
    if "p" in files and "rho" in files:
        rho, p = np.asarray(d["rho"], np.float64), np.asarray(d["p"], np.float64)
        ncomp = U.shape[0]
        ekey = {3: 2, 4: 3, 5: 4, 8: 4}.get(ncomp)
        if ekey is not None:
            vel = [np.asarray(d[k], np.float64) for k in ("ux", "uy", "uz", "u") if k in files]
            kin = 0.5 * rho * sum(v * v for v in vel) if vel else 0.0
            mag = 0.0
            if ncomp == 8:
                mag = 0.5 * (U[5].astype(np.float64) ** 2 + U[6] ** 2 + U[7] ** 2)
                meta["mhd"] = True
            internal = U[ekey].astype(np.float64) - kin - mag
            ok = np.abs(internal) > 1e-12 * np.abs(U[ekey]).max()
            if ok.any():
                meta["gamma"] = float(1.0 + np.median(p[ok] / internal[ok]))
    return meta


def load_checkpoint(path: str) -> Dict[str, Any]:
    """Read a snapshot (phrike/io.py:32-83 layout of the returned dict) without
    unpickling anything (``allow_pickle=False``): t, U, x/y/z, primitive
    variables, and meta from ``meta_json`` (files written here).  The
    reference's own files store meta as a pickled dict; for those, meta is
    recovered from the plain arrays (_recover_meta) and
    ``meta["recovered"] = True``."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Checkpoint file not found: {path}")
    with np.load(path, allow_pickle=False) as d:
        files = set(d.files)
        out: Dict[str, Any] = {"t": float(d["t"]), "U": d["U"]}
        if "meta_json" in files:
            out["meta"] = json.loads(str(d["meta_json"]))
        else:
            out["meta"] = _recover_meta(d, files)
            out["meta"]["recovered"] = True
        for k in ("x", "y", "z"):
            if k in files:
                out[k] = d[k]
        out["primitive_vars"] = {k: d[k] for k in ("rho", "u", "ux", "uy", "uz", "p") if k in files}
        for k in ("Bx", "By", "Bz", "div_b"):
            if k in files:
                out[k] = d[k]
    return out


def load_restart_data(path: str) -> Dict[str, Any]:
    """BaseProblem.load_restart_data (phrike/problems/base.py:238-255): the
    checkpoint, whose t becomes the run's t0."""
    data = load_checkpoint(path)
    data["t0"] = data["t"]
    return data


def validate_restart_data(data: Dict[str, Any], grid, equations) -> List[str]:
    """BaseProblem.validate_restart_data (phrike/problems/base.py:257-290):
    a grid size mismatch is an error (ValueError); a gamma mismatch beyond
    1e-10 is a warning, returned (and printed) as the reference prints it.
    The state's shape must also match the grid (the reference leaves that to
    the first step)."""
    warnings: List[str] = []
    if data is None:
        return warnings
    meta = data.get("meta", {})
    if "gamma" in meta and abs(float(meta["gamma"]) - equations.gamma) > 1e-10:
        w = f"Warning: Gamma mismatch - restart: {meta['gamma']}, current: {equations.gamma}"
        print(w)
        warnings.append(w)
    if "N" in meta:
        if meta["N"] != getattr(grid, "N", None):
            raise ValueError(f"Grid size mismatch - restart: {meta['N']}, current: {getattr(grid, 'N', None)}")
    elif "Nx" in meta:
        for k in ("Nx", "Ny", "Nz"):
            if k in meta and meta[k] != getattr(grid, k, None):
                raise ValueError(f"Grid {k} mismatch - restart: {meta[k]}, current: {getattr(grid, k, None)}")
    dims = tuple(getattr(grid, k) for k in ("Nz", "Ny", "Nx") if hasattr(grid, k))
    U = np.asarray(data["U"])
    if tuple(U.shape[1:]) != dims:
        raise ValueError(f"restart state has shape {U.shape}, the grid is {dims}")
    return warnings


def restart_state(path: str, grid, equations):
    """(U0, t0) for a run that continues from a checkpoint: the state as a
    tensor on the grid's device and dtype, validated against the grid and
    gamma first."""
    data = load_restart_data(path)
    validate_restart_data(data, grid, equations)
    U0 = torch.as_tensor(np.ascontiguousarray(data["U"]), dtype=grid.dtype, device=grid.device)
    return U0, float(data["t0"])
