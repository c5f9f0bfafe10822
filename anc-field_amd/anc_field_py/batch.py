"""Batched impulse-response identification: many independent rooms stepped
together in one handle (one fused launch per step covers every room),
against the same rooms run one after another.

The reference runs one room per process and one source per run
(examples/ie224-impulse.py:79-106); every room here carries its own
source(s) and receivers, and each receiver trace is one impulse response.
"""
import time

import numpy as np

from . import _afx


def make_batch(workloads, precision="fp32", compat=False, **kw):
    """One handle holding every workload as its own room (shared c, dx, dt)."""
    w0 = workloads[0]
    s = _afx.Sim(c=w0.c, dx=w0.dx, dt=w0.dt, res=w0.dx, precision=precision, compat=compat, **kw)
    s.set_voxels(w0.vox)
    s.set_materials(w0.a, w0.b)
    ids = [0] + [s.add_room(w.vox, w.a, w.b) for w in workloads[1:]]
    for rid, w in zip(ids, workloads):
        for cell, sig in zip(w.src_cells, w.src):
            s.add_source_cell(cell, sig, room=rid)
    for rid, w in zip(ids, workloads):
        for cell in w.mic_cells:
            s.add_receiver_cell(cell, room=rid)
    return s


def run_batched(workloads, steps, precision="fp32", tune=False, **kw):
    """Returns (traces [sum of receivers, steps], wall seconds incl. setup,
    tile shape).  With tune, afx_autotune picks the tile shape and the steps
    per pass first (every candidate checked bit-exact; the tile returned is
    then (tx, ty, tz, steps per pass)); tuning time is not included."""
    t0 = time.perf_counter()
    s = make_batch(workloads, precision, **kw)
    tile = None
    if tune:
        t1 = time.perf_counter()
        tile = s.autotune(30) + (s.time_block()[0],)
        t0 += time.perf_counter() - t1
    s.run(steps)
    tr = s.receivers()
    s.close()
    return tr, time.perf_counter() - t0, tile


STORAGE_BYTES = {"fp32": 4, "fp64": 8, "fp16": 2, "bf16": 2}


def cache_groups(workloads, precision="fp32", budget=192 << 20, fields=4):
    """Rooms split into consecutive groups whose pressure fields (`fields`
    of them, per cell) fit `budget` bytes, so a group's whole working set
    stays in the 256 MB Infinity Cache while it is stepped.  Returns lists of
    indices into workloads; a room larger than the budget is a group alone."""
    per_cell = fields * STORAGE_BYTES[precision]
    groups, cur, used = [], [], 0
    for i, w in enumerate(workloads):
        b = w.cells * per_cell
        if cur and used + b > budget:
            groups.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += b
    if cur:
        groups.append(cur)
    return groups


def run_grouped(workloads, steps, precision="fp32", tune=False, budget=192 << 20, **kw):
    """The batch stepped group by group (cache_groups): each group is one
    handle run for all `steps` before the next starts.  Returns (traces in
    workload order, wall seconds excluding tuning, groups)."""
    groups = cache_groups(workloads, precision, budget)
    out = [None] * len(workloads)
    wall = 0.0
    for g in groups:
        ws = [workloads[i] for i in g]
        t0 = time.perf_counter()
        s = make_batch(ws, precision, **kw)
        if tune:
            t1 = time.perf_counter()
            s.autotune(20)
            t0 += time.perf_counter() - t1
        s.run(steps)
        tr = s.receivers()
        s.close()
        wall += time.perf_counter() - t0
        k = 0
        for i, w in zip(g, ws):
            n = len(w.mic_cells)
            out[i] = tr[k:k + n]
            k += n
    return np.concatenate(out, axis=0), wall, groups


def run_sequential(workloads, steps, precision="fp32", **kw):
    """One handle per room, run back to back."""
    from .rooms import make_sim
    t0 = time.perf_counter()
    out = []
    for w in workloads:
        s = make_sim(w, _afx.Sim, precision=precision, **kw)
        s.run(steps)
        out.append(s.receivers())
        s.close()
    return np.concatenate(out, axis=0), time.perf_counter() - t0


def rel_l2(a, b):
    """Relative L2 error of a against b over all traces."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
