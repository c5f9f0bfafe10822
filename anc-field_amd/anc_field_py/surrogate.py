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
5. Source zones (``SourceZone``, ``fit_zone``, ``calibrate_zone``): a box of
   candidate loudspeaker cells with fixed receivers.  FDTD runs from a
   lattice of the box train an ensemble of networks (``members``); held-out
   zone sources calibrate a threshold on the ensemble's spread.  A query from
   a new cell of the box is then answered by the ensemble mean, without an
   FDTD run, when the spread of every receiver of that source is within it.
   Receivers need no such zone: one FDTD run gives every cell.

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
    def __init__(self, w, taps=1024, decim=4, fc=2000.0, bound=5e-2, rooms_per_handle=16, mlp=None, members=1,
                 **mlp_kw):
        """members > 1 trains an ensemble of that many networks (initial
        weights seed, seed+1, ...; minibatch order seed, seed+1, ...) on the
        same data: the answer is their mean and their spread estimates its
        error."""
        self.w = w
        self.taps, self.decim, self.fc, self.bound = taps, decim, fc, bound
        self.steps = taps * decim
        self.rooms_per_handle = rooms_per_handle
        kw = dict(in_dim=6, out_dim=taps)
        kw.update(mlp_kw)
        if mlp is not None:
            self.nets = [mlp]
        else:
            seed = int(kw.pop("seed", 1))
            self.nets = []
            for k in range(members):
                self.nets.append(_afx.Mlp(seed=seed + k, **kw))
        self.net = self.nets[0]
        nz, ny, nx = w.vox.shape
        self.dims = np.array([nx, ny, nz], np.float64)
        self.validated = {}  # (src cell, rcv cell) -> measured relative L2 error
        self.zones = []      # SourceZone: queries from inside a zone may be accepted unvalidated
        self.scale = None    # global RMS of the training responses
        self.fdtd_runs = 0
        self.fdtd_seconds = 0.0

    def close(self):
        for n in self.nets:
            n.close()

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
        loss = []
        for k, n in enumerate(self.nets):
            n.set_data(x, t, wgt)
            loss.append(n.train(steps, lr0, lr1, seed + k))
        return np.mean(loss, axis=0)

    def members(self, src, rcv):
        """Each network's responses, [members][n][taps]."""
        x = self.encode(src, rcv)
        sc = np.float32(self.scale if self.scale else 1.0)
        return np.stack([n.forward(x) * sc for n in self.nets])

    def predict(self, src, rcv):
        return self.members(src, rcv).mean(axis=0)

    @staticmethod
    def spread(ys):
        """Relative sample standard deviation of the members' responses
        [members][n][taps] about their mean, per row (needs 2 members)."""
        m = ys.mean(axis=0)
        dev = np.sqrt(np.sum((ys - m) ** 2, axis=(0, 2)) / (len(ys) - 1))
        return dev / np.maximum(np.linalg.norm(m, axis=1), 1e-30)

    def validate(self, src, rcv, ir):
        """Network error on pairs with known responses; records them."""
        err = rel_l2(self.predict(src, rcv), ir)
        for s_, r_, e in zip(np.atleast_2d(src), np.atleast_2d(rcv), err):
            self.validated[(tuple(int(v) for v in s_), tuple(int(v) for v in r_))] = float(e)
        return err

    def trusted(self, s_, r_, spread=None):
        """Validated pairs; or a source inside a calibrated zone, a receiver of
        that zone, and the ensemble spread within the zone's threshold."""
        e = self.validated.get((tuple(int(v) for v in s_), tuple(int(v) for v in r_)))
        if e is not None:
            return e <= self.bound
        if spread is None:
            return False
        for z in self.zones:
            if z.delta < 0 or not z.inside(s_)[0]:
                continue
            if np.all(z.receivers == np.asarray(r_), axis=1).any():
                return spread <= z.delta
        return False

    def query(self, src, rcv):
        """(responses [n][taps], from_fdtd [n] bool, network error [n] or nan).
        One FDTD run gives every receiver of a source, so when any pair of a
        source is not trusted, that source is simulated and the run answers
        all of its pairs."""
        src = np.atleast_2d(np.asarray(src, np.int32))
        rcv = np.atleast_2d(np.asarray(rcv, np.int32))
        n = len(src)
        ys = self.members(src, rcv)
        out = ys.mean(axis=0)
        sp = self.spread(ys) if len(ys) > 1 else np.full(n, np.nan)
        fb = np.array([not self.trusted(s_, r_, None if np.isnan(d) else d) for s_, r_, d in zip(src, rcv, sp)],
                      bool)
        for k in {tuple(int(v) for v in src[i]) for i in np.flatnonzero(fb)}:
            fb[np.all(src == np.asarray(k), axis=1)] = True
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


# ---------------------------------------------------------------- source zones
class SourceZone:
    """A box of candidate source cells (a loudspeaker region) with fixed
    receivers.  FDTD runs from every `step`-th cell of the box train the
    ensemble; a query from any cell of the box is accepted when the ensemble
    spread is at most `delta`, a threshold calibrated on held-out zone
    sources so that every accepted validation answer was within a margin of
    the bound.  Trilinear interpolation of the training responses is the
    same data's answer without a network (the comparison baseline)."""

    def __init__(self, lo, size, step, receivers):
        self.lo = np.asarray(lo, np.int32)
        self.size, self.step = int(size), int(step)
        self.receivers = np.asarray(receivers, np.int32).reshape(-1, 3)
        self.delta = -1.0  # no query accepted before calibration
        self.train_ir = None  # [n_lattice^3][n_receivers][taps]

    def lattice(self):
        """Training source cells: the box's cells at multiples of step, x fastest."""
        k = np.arange(0, self.size, self.step)
        zz, yy, xx = np.meshgrid(k, k, k, indexing="ij")
        return (self.lo + np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)).astype(np.int32)

    def inside(self, src):
        """Within the hull of the training lattice (where interpolation holds)."""
        off = np.atleast_2d(src) - self.lo
        hi = (self.size - 1) // self.step * self.step
        return np.all((off >= 0) & (off <= hi), axis=1)

    def interpolate(self, src, rcv_index):
        """Trilinear interpolation of the training responses at source cells."""
        n = (self.size - 1) // self.step + 1
        g = self.train_ir.reshape(n, n, n, len(self.receivers), -1)
        out = []
        for s_, r_ in zip(np.atleast_2d(src), np.atleast_1d(rcv_index)):
            f = (s_ - self.lo) / self.step
            i0 = np.minimum(np.floor(f).astype(int), n - 2)
            t = f - i0
            acc = 0.0
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        wgt = (t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2])
                        acc = acc + wgt * g[i0[2] + dz, i0[1] + dy, i0[0] + dx, r_]
            out.append(acc)
        return np.asarray(out, np.float32)


def zone_box(w, rng, size, margin=2):
    """Origin of a size^3 box of air cells (with air neighbours), outside the PML."""
    v = w.vox
    for _ in range(10000):
        c = free_cells(w, rng, 1, margin)[0] - size // 2
        blk = v[c[2] - 1:c[2] + size + 1, c[1] - 1:c[1] + size + 1, c[0] - 1:c[0] + size + 1]
        if blk.shape == (size + 2,) * 3 and (blk == rooms.AIR).all() and c.min() >= margin \
                and c[0] + size < v.shape[2] - (w.pml_width + 1 + margin):
            return c.astype(np.int32)
    raise RuntimeError("no free box of that size")


def fit_zone(sur, zone, steps, lr0=1e-3, lr1=1e-5, seed=0):
    """FDTD runs from the zone's lattice, then training on those pairs."""
    lat = zone.lattice()
    ir = sur.simulate(lat, [zone.receivers] * len(lat))
    zone.train_ir = ir.reshape(len(lat), len(zone.receivers), -1)
    s, r = pairs(lat, [zone.receivers] * len(lat))
    loss = sur.fit(sur.encode(s, r), ir, steps, lr0, lr1, seed)
    sur.zones.append(zone)
    return loss


def spread_threshold(spread, err, limit):
    """The largest spread of a validation answer with error <= limit that is
    below the smallest spread of one above it (-1 when there is none): every
    validation answer at or below the threshold was within the limit."""
    ok = err <= limit
    lim = spread[~ok].min() if (~ok).any() else np.inf
    below = spread[ok & (spread < lim)]
    return float(below.max()) if len(below) else -1.0


def calibrate_zone(sur, zone, val_sources, margin=0.8):
    """FDTD from held-out zone sources judges the ensemble; delta is the
    spread threshold below which every validation answer was within
    margin * bound (the margin covers queries the validation set did not
    see).  Returns (ensemble error, interpolation error, spread) per
    validation pair."""
    if len(sur.nets) < 2:
        raise ValueError("zone calibration needs an ensemble (members >= 2)")
    val_sources = np.asarray(val_sources, np.int32).reshape(-1, 3)
    ir = sur.simulate(val_sources, [zone.receivers] * len(val_sources))
    s, r = pairs(val_sources, [zone.receivers] * len(val_sources))
    ridx = np.tile(np.arange(len(zone.receivers)), len(val_sources))
    ys = sur.members(s, r)
    e_net, sp = rel_l2(ys.mean(axis=0), ir), sur.spread(ys)
    e_tri = rel_l2(zone.interpolate(s, ridx), ir)
    zone.delta = spread_threshold(sp, e_net, margin * sur.bound)
    for s_, r_, e in zip(s, r, e_net):
        sur.validated[(tuple(int(v) for v in s_), tuple(int(v) for v in r_))] = float(e)
    return e_net, e_tri, sp
