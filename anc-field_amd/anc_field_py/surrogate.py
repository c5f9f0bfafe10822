"""Learned secondary-path surrogate for one fixed room, with FDTD as fallback.

The reference identifies every secondary path (loudspeaker -> error
microphone) by running the full simulation and estimating the impulse
response from it (anc_field_py/ancutil.py:49-79, examples/identification*.py,
which keeps every 4th sample, identification.py:36).  Here:

1. ``simulate`` runs many such identifications at once: copies of the room
   share one afx handle (batched rooms, one fused launch per step), each copy
   holding one unit-impulse source and its receivers.  Each trace is low-passed
   (zero-phase windowed sinc at ``fc``) and decimated by ``decim`` to ``taps``
   samples: the band-limited impulse response.
2. ``fit`` trains the afx MLP (BF16 MFMA kernels) to map the normalised
   (source cell, receiver cell) pair to that response, with the loss weighted
   so that it is the mean squared relative L2 error.
3. ``validate`` measures the relative L2 error of the network on pairs whose
   FDTD response is known; pairs within ``bound`` become validated.
4. ``query`` answers a list of pairs: validated pairs from the network, all
   others by FDTD, which also judges the network's answer for those pairs
   (and validates them when it is within ``bound``).

Positions are cell indices (x, y, z); the network sees (cell + 0.5) / dim.
"""
import time

import numpy as np
from scipy.signal import oaconvolve

from . import _afx, rooms


def lowpass_decimate(traces, dt, taps, decim, fc, ntaps=129):
    """Zero-phase FIR low-pass at fc (Hamming-windowed sinc), then every decim-th sample."""
    traces = np.atleast_2d(np.asarray(traces, np.float64))
    h = rooms.lowpass_fir(fc, 1.0 / dt, ntaps)
    half = (ntaps - 1) // 2
    y = oaconvolve(traces, h[None, :], mode="full", axes=1)[:, half:half + traces.shape[1]]
    out = y[:, ::decim][:, :taps]
    if out.shape[1] < taps:
        raise ValueError(f"{traces.shape[1]} samples give fewer than {taps} taps after decimation by {decim}")
    return out.astype(np.float32)


def rel_l2(y, t):
    """Per-row relative L2 error |y - t| / |t|."""
    y = np.asarray(y, np.float64)
    t = np.asarray(t, np.float64)
    return np.linalg.norm(y - t, axis=-1) / np.maximum(np.linalg.norm(t, axis=-1), 1e-300)


def free_cells(w, rng, k, margin=2):
    """k distinct air cells with air neighbours, outside the PML slabs."""
    v = w.vox.copy()
    if w.pml_width:
        nz, ny, nx = v.shape
        depth = w.pml_width + 1 + margin
        f = w.pml_faces
        if f & 1:
            v[:, :, :depth] = 0
        if f & 2:
            v[:, :, nx - depth:] = 0
        if f & 4:
            v[:, :depth, :] = 0
        if f & 8:
            v[:, ny - depth:, :] = 0
        if f & 16:
            v[:depth] = 0
        if f & 32:
            v[nz - depth:] = 0
    return rooms.interior_cells(v, rng, k, margin)


class Surrogate:
    def __init__(self, w, taps=1024, decim=4, fc=2000.0, bound=5e-2, rooms_per_handle=16, mlp=None, **mlp_kw):
        self.w = w
        self.taps, self.decim, self.fc, self.bound = taps, decim, fc, bound
        self.steps = taps * decim
        self.rooms_per_handle = rooms_per_handle
        kw = dict(in_dim=6, out_dim=taps)
        kw.update(mlp_kw)
        self.net = mlp if mlp is not None else _afx.Mlp(**kw)
        nz, ny, nx = w.vox.shape
        self.dims = np.array([nx, ny, nz], np.float64)
        self.validated = {}  # (src cell, rcv cell) -> measured relative L2 error
        self.scale = None    # global RMS of the training responses
        self.fdtd_runs = 0
        self.fdtd_seconds = 0.0

    def close(self):
        self.net.close()

    # ------------------------------------------------------------ inputs
    def encode(self, src, rcv):
        src = np.atleast_2d(np.asarray(src, np.float64))
        rcv = np.atleast_2d(np.asarray(rcv, np.float64))
        return np.concatenate([(src + 0.5) / self.dims, (rcv + 0.5) / self.dims], axis=1).astype(np.float32)

    # ------------------------------------------------------------ FDTD
    def simulate(self, src_cells, rcv_lists, precision="fp32"):
        """Band-limited impulse responses: for source i, one row per cell in
        rcv_lists[i].  Returns [sum len(rcv_lists[i])][taps] in that order."""
        w = self.w
        src_cells = np.asarray(src_cells, np.int32).reshape(-1, 3)
        impulse = np.array([1.0], np.float32)
        out = []
        t0 = time.perf_counter()
        for c0 in range(0, len(src_cells), self.rooms_per_handle):
            chunk = range(c0, min(c0 + self.rooms_per_handle, len(src_cells)))
            s = _afx.Sim(c=w.c, dx=w.dx, dt=w.dt, res=w.dx, compat=False, precision=precision)
            try:
                s.set_voxels(w.vox)
                s.set_materials(w.a, w.b)
                if w.pml_width:
                    s.set_pml(w.pml_width, w.pml_faces, w.pml_profile)
                ids = [0] + [s.add_room(w.vox, w.a, w.b) for _ in list(chunk)[1:]]
                for rid, i in zip(ids, chunk):
                    s.add_source_cell(src_cells[i], impulse, room=rid)
                for rid, i in zip(ids, chunk):
                    for cell in np.asarray(rcv_lists[i], np.int32).reshape(-1, 3):
                        s.add_receiver_cell(cell, room=rid)
                s.run(self.steps)
                tr = s.receivers()
            finally:
                s.close()
            out.append(lowpass_decimate(tr, w.dt, self.taps, self.decim, self.fc))
            self.fdtd_runs += len(chunk)
        self.fdtd_seconds += time.perf_counter() - t0
        return np.concatenate(out, axis=0)

    # ------------------------------------------------------------ network
    def fit(self, x, ir, steps, lr0=1e-3, lr1=1e-5, seed=0):
        """Trains on responses divided by their global RMS (kept as self.scale);
        the loss is the mean squared relative L2 error of the batch."""
        ir = np.asarray(ir, np.float64)
        if self.scale is None:
            self.scale = float(np.sqrt(np.mean(ir ** 2)))
        t = (ir / self.scale).astype(np.float32)
        wgt = (1.0 / (self.net.cfg.batch * np.sum(t.astype(np.float64) ** 2, axis=1))).astype(np.float32)
        self.net.set_data(x, t, wgt)
        return self.net.train(steps, lr0, lr1, seed)

    def predict(self, src, rcv):
        return self.net.forward(self.encode(src, rcv)) * np.float32(self.scale if self.scale else 1.0)

    def validate(self, src, rcv, ir):
        """Network error on pairs with known responses; records them."""
        err = rel_l2(self.predict(src, rcv), ir)
        for s_, r_, e in zip(np.atleast_2d(src), np.atleast_2d(rcv), err):
            self.validated[(tuple(int(v) for v in s_), tuple(int(v) for v in r_))] = float(e)
        return err

    def trusted(self, s_, r_):
        e = self.validated.get((tuple(int(v) for v in s_), tuple(int(v) for v in r_)))
        return e is not None and e <= self.bound

    def query(self, src, rcv):
        """(responses [n][taps], from_fdtd [n] bool, network error [n] or nan)."""
        src = np.atleast_2d(np.asarray(src, np.int32))
        rcv = np.atleast_2d(np.asarray(rcv, np.int32))
        n = len(src)
        out = self.predict(src, rcv)
        fb = np.array([not self.trusted(s_, r_) for s_, r_ in zip(src, rcv)], bool)
        err = np.full(n, np.nan)
        if fb.any():
            idx = np.flatnonzero(fb)
            keys = sorted({tuple(int(v) for v in src[i]) for i in idx})
            rows = {k: [i for i in idx if tuple(int(v) for v in src[i]) == k] for k in keys}
            ir = self.simulate(np.array(keys, np.int32), [rcv[rows[k]] for k in keys])
            order = [i for k in keys for i in rows[k]]
            err[order] = self.validate(src[order], rcv[order], ir)
            out[order] = ir
        return out, fb, err


def pairs(sources, rcv_lists):
    """Flatten per-source receiver lists into (src [n][3], rcv [n][3])."""
    s = np.concatenate([np.repeat(np.asarray(c, np.int32)[None], len(r), 0) for c, r in zip(sources, rcv_lists)])
    r = np.concatenate([np.asarray(x, np.int32).reshape(-1, 3) for x in rcv_lists])
    return s, r


def nearest_pair_baseline(train_src, train_rcv, train_ir, src, rcv):
    """Response of the nearest training pair (6-D Euclidean distance in cells):
    what the training set gives without a network."""
    a = np.concatenate([train_src, train_rcv], axis=1).astype(np.float64)
    b = np.concatenate([src, rcv], axis=1).astype(np.float64)
    d = ((b[:, None, :] - a[None, :, :]) ** 2).sum(-1)
    return np.asarray(train_ir)[np.argmin(d, axis=1)]
