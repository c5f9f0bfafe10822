"""Seeded room, material, source and receiver generators for the workloads.

All randomness comes from numpy's PCG64 seeded per workload, so every
consumer (library, tests, oracle, benchmark) sees identical inputs.

Workloads (grid sizes in cells, x, y, z):
  shoebox_small   64 x 48 x 32, dx 0.02, 3 IIR materials on the walls, one
                  face replaced by an 8-cell PML, Ricker 1 kHz at (10,12,8),
                  4 receivers, 2000 steps
  furnished       256 x 192 x 128, 12 seeded boxes (sizes 8..40), 4 materials,
                  12-cell PML on one face, 4 kHz band-limited noise, 16
                  receivers, 10 000 steps
  batch_rooms     48 rooms 64^3 .. 256 x 192 x 128, 2-4 materials, Ricker,
                  8 receivers each, 4096 steps
  large           640 x 480 x 320, dx 0.01, 6 materials, 16-cell PML on two
                  faces, band-limited noise, 32 receivers, 20 000 steps
"""
import math

import numpy as np

AIR = -1
C_AIR = 343.0


def courant_dt(dx, c=C_AIR, courant=0.99):
    """dt = 0.99 dx / (c sqrt 3) in single precision."""
    return float(np.float32(courant * dx / (c * math.sqrt(3.0))))


# ---------------------------------------------------------------- geometry
def shoebox(nx, ny, nz, face_mats, open_faces=()):
    """Voxels [nz, ny, nx]: a one-cell solid shell on the grid faces.

    face_mats: material per face in order x-, x+, y-, y+, z-, z+.
    open_faces: faces left as air (their plane is the zero-pressure edge
    behind a PML).
    """
    v = np.full((nz, ny, nx), AIR, np.int32)
    order = ["x-", "x+", "y-", "y+", "z-", "z+"]
    sl = {
        "x-": (slice(None), slice(None), 0), "x+": (slice(None), slice(None), nx - 1),
        "y-": (slice(None), 0, slice(None)), "y+": (slice(None), ny - 1, slice(None)),
        "z-": (0, slice(None), slice(None)), "z+": (nz - 1, slice(None), slice(None)),
    }
    for f, m in zip(order, face_mats):
        if f in open_faces:
            continue
        v[sl[f]] = m
    for f in open_faces:  # an open face clears the shell edges lying in it
        v[sl[f]] = AIR
    return v


def add_boxes(v, rng, n, lo, hi, nmat, margin=2):
    """n axis-aligned solid boxes with edge lengths U[lo, hi] cells."""
    nz, ny, nx = v.shape
    for _ in range(n):
        sx, sy, sz = (int(rng.integers(lo, hi + 1)) for _ in range(3))
        sx, sy, sz = min(sx, nx - 2 * margin), min(sy, ny - 2 * margin), min(sz, nz - 2 * margin)
        x0 = int(rng.integers(margin, nx - margin - sx + 1))
        y0 = int(rng.integers(margin, ny - margin - sy + 1))
        z0 = int(rng.integers(margin, nz - margin - sz + 1))
        v[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = int(rng.integers(0, nmat))
    return v


def repair(v):
    """Fill air cells that would see solid on two opposite faces.

    The boundary scheme models locally reacting walls on one side per axis
    only (such cells are rejected as unphysical); a one-cell slot between
    two solids is closed with the material of the +side neighbour.
    """
    v = v.copy()
    while True:
        s = v != AIR
        a = ~s[1:-1, 1:-1, 1:-1]
        bad = np.zeros_like(a)
        fill = np.full(a.shape, AIR, np.int32)
        for ax, (p, m) in enumerate([((1, 1, 2), (1, 1, 0)), ((1, 2, 1), (1, 0, 1)), ((2, 1, 1), (0, 1, 1))]):
            sp = s[p[0]:p[0] + a.shape[0], p[1]:p[1] + a.shape[1], p[2]:p[2] + a.shape[2]]
            sm = s[m[0]:m[0] + a.shape[0], m[1]:m[1] + a.shape[1], m[2]:m[2] + a.shape[2]]
            vp = v[p[0]:p[0] + a.shape[0], p[1]:p[1] + a.shape[1], p[2]:p[2] + a.shape[2]]
            hit = a & sp & sm & ~bad
            fill = np.where(hit, vp, fill)
            bad |= hit
        if not bad.any():
            return v
        inner = v[1:-1, 1:-1, 1:-1]
        inner[bad] = fill[bad]


# ---------------------------------------------------------------- materials
def iir_material(rng, r_dc=None):
    """Second-order admittance filter beta(z) = A(z)/B(z), passive.

    A = g (1 - 2 r cos(th) z^-1 + r^2 z^-2): conjugate poles of the impedance
    B/A with radius r ~ U[0.5, 0.9] and angle th ~ U[0, pi/2].
    B = 1 - 2 rho cos(th) z^-1 + rho^2 z^-2 with rho ~ r U[0.2, 0.8].
    g scales the filter so the DC reflection (sum B - sum A)/(sum B + sum A)
    is r_dc ~ U[0.6, 0.95].  Draws are repeated until Re beta >= 0 on the
    unit circle (passivity).
    """
    w = np.linspace(0, math.pi, 512)
    z1 = np.exp(-1j * w)
    for _ in range(1000):
        r = rng.uniform(0.5, 0.9)
        th = rng.uniform(0.0, math.pi / 2)
        rho = r * rng.uniform(0.2, 0.8)
        R = rng.uniform(0.6, 0.95) if r_dc is None else r_dc
        a0 = np.array([1.0, -2 * r * math.cos(th), r * r])
        b = np.array([1.0, -2 * rho * math.cos(th), rho * rho])
        g = (1 - R) / (1 + R) * b.sum() / a0.sum()
        a = g * a0
        beta = (a[0] + a[1] * z1 + a[2] * z1 ** 2) / (b[0] + b[1] * z1 + b[2] * z1 ** 2)
        if beta.real.min() >= 0:
            return a.astype(np.float32), b.astype(np.float32)
    raise RuntimeError("no passive filter drawn")


def materials(rng, nmat):
    ab = [iir_material(rng) for _ in range(nmat)]
    return np.stack([x[0] for x in ab]), np.stack([x[1] for x in ab])


def reflection_dc(a, b):
    return (b.sum() - a.sum()) / (b.sum() + a.sum())


# ---------------------------------------------------------------- signals
def ricker(n, dt, f0=1000.0, t0=None, amp=1.0):
    """Ricker wavelet, peak at t0 (default 1.2/f0)."""
    t0 = 1.2 / f0 if t0 is None else t0
    t = np.arange(n) * dt - t0
    x = (math.pi * f0 * t) ** 2
    return (amp * (1 - 2 * x) * np.exp(-x)).astype(np.float32)


def lowpass_fir(fc, fs, ntaps=101):
    """Hamming-windowed sinc low-pass, unit DC gain."""
    m = np.arange(ntaps) - (ntaps - 1) / 2
    h = 2 * fc / fs * np.sinc(2 * fc / fs * m) * np.hamming(ntaps)
    return h / h.sum()


def band_noise(rng, n, dt, fc=4000.0, centre=0.25, width=0.1):
    """Unit-variance white noise, low-passed at fc, under a Gaussian window
    centred at `centre` of the duration with standard deviation `width`."""
    x = rng.standard_normal(n + 200)
    y = np.convolve(x, lowpass_fir(fc, 1.0 / dt), mode="same")[100:100 + n]
    y /= y.std() if y.std() > 0 else 1.0
    t = np.arange(n) / max(n, 1)
    return (y * np.exp(-0.5 * ((t - centre) / width) ** 2)).astype(np.float32)


def interior_cells(v, rng, k, margin=2):
    """k distinct air cells whose 6 neighbours are all air."""
    s = v != AIR
    nz, ny, nx = v.shape
    ok = ~s.copy()
    ok[:margin] = ok[-margin:] = False
    ok[:, :margin] = ok[:, -margin:] = False
    ok[:, :, :margin] = ok[:, :, -margin:] = False
    inner = np.zeros_like(ok)
    inner[1:-1, 1:-1, 1:-1] = (~s[1:-1, 1:-1, 1:-1] & ~s[1:-1, 1:-1, 2:] & ~s[1:-1, 1:-1, :-2] & ~s[1:-1, 2:, 1:-1]
                               & ~s[1:-1, :-2, 1:-1] & ~s[2:, 1:-1, 1:-1] & ~s[:-2, 1:-1, 1:-1])
    idx = np.flatnonzero(ok & inner)
    pick = rng.choice(idx, size=k, replace=False)
    z, rem = np.divmod(pick, ny * nx)
    y, x = np.divmod(rem, nx)
    return np.stack([x, y, z], axis=1).astype(np.int32)


# ---------------------------------------------------------------- workloads
def pml_profile(width, dx, c=C_AIR, order=3.0, r0=1e-3):
    """Graded PML damping (1/s) at cell-centre depths 0.5 .. width-0.5:
    zeta_max (d/width)^order with zeta_max = (order+1) c ln(1/r0) / (2 width dx)."""
    zmax = (order + 1.0) * c * math.log(1.0 / r0) / (2.0 * width * dx)
    return (zmax * ((np.arange(width) + 0.5) / width) ** order).astype(np.float32)


class Workload:
    """Everything one room run needs; cells are (x, y, z)."""

    def __init__(self, name, vox, a, b, dx, dt, src_cells, src, mic_cells, steps, pml_width=0, pml_faces=0, c=C_AIR):
        self.name, self.vox, self.a, self.b = name, vox, a, b
        self.dx, self.dt, self.c = dx, dt, c
        self.src_cells, self.src, self.mic_cells = src_cells, src, mic_cells
        self.steps, self.pml_width, self.pml_faces = steps, pml_width, pml_faces
        self.pml_profile = pml_profile(pml_width, dx, c) if pml_width else None

    @property
    def dims(self):
        nz, ny, nx = self.vox.shape
        return nx, ny, nz

    @property
    def cells(self):
        return int(self.vox.size)


FACE_BITS = {"x-": 1, "x+": 2, "y-": 4, "y+": 8, "z-": 16, "z+": 32}


def shoebox_small(seed=1, steps=2000):
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz, dx = 64, 48, 32, 0.02
    dt = courant_dt(dx)
    a, b = materials(rng, 3)
    face_mats = [int(rng.integers(0, 3)) for _ in range(6)]
    v = repair(shoebox(nx, ny, nz, face_mats, open_faces=("x+",)))
    src = ricker(steps, dt, 1000.0)[None, :]
    mics = interior_cells(v, rng, 4)
    return Workload("shoebox_small", v, a, b, dx, dt, np.array([[10, 12, 8]], np.int32), src, mics, steps,
                    pml_width=8, pml_faces=FACE_BITS["x+"])


def furnished(seed=2, steps=10000, dims=(256, 192, 128), nbox=12, nmic=16):
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = dims
    dx = 0.02
    dt = courant_dt(dx)
    a, b = materials(rng, 4)
    face_mats = [int(rng.integers(0, 4)) for _ in range(6)]
    v = shoebox(nx, ny, nz, face_mats, open_faces=("x+",))
    v = repair(add_boxes(v, rng, nbox, 8, 40, 4))
    src_cell = interior_cells(v, rng, 1)
    src = band_noise(rng, steps, dt, 4000.0)[None, :]
    mics = interior_cells(v, rng, nmic)
    return Workload("furnished", v, a, b, dx, dt, src_cell, src, mics, steps, pml_width=12, pml_faces=FACE_BITS["x+"])


def batch_rooms(seed=3, nrooms=48, steps=4096, lo=(64, 64, 64), hi=(256, 192, 128), nmic=8):
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 0.02
    dt = courant_dt(dx)
    out = []
    for k in range(nrooms):
        nx, ny, nz = (int(rng.integers(l, h + 1)) for l, h in zip(lo, hi))
        nmat = int(rng.integers(2, 5))
        a, b = materials(rng, nmat)
        face_mats = [int(rng.integers(0, nmat)) for _ in range(6)]
        v = shoebox(nx, ny, nz, face_mats)
        nb = int(rng.integers(2, 7))
        v = repair(add_boxes(v, rng, nb, 6, max(8, min(nx, ny, nz) // 4), nmat))
        src_cell = interior_cells(v, rng, 1)
        src = ricker(steps, dt, 1000.0)[None, :]
        mics = interior_cells(v, rng, nmic)
        out.append(Workload(f"batch_room_{k}", v, a, b, dx, dt, src_cell, src, mics, steps))
    return out


def large(seed=4, steps=20000, dims=(640, 480, 320), nbox=24, nmic=32):
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = dims
    dx = 0.01
    dt = courant_dt(dx)
    a, b = materials(rng, 6)
    face_mats = [int(rng.integers(0, 6)) for _ in range(6)]
    v = shoebox(nx, ny, nz, face_mats, open_faces=("x+", "y+"))
    v = repair(add_boxes(v, rng, nbox, 16, 80, 6))
    src_cell = interior_cells(v, rng, 1)
    src = band_noise(rng, steps, dt, 4000.0)[None, :]
    mics = interior_cells(v, rng, nmic)
    return Workload("large", v, a, b, dx, dt, src_cell, src, mics, steps, pml_width=16,
                    pml_faces=FACE_BITS["x+"] | FACE_BITS["y+"])


def make_sim(w, sim_cls, compat=False, precision="fp32", **kw):
    """Configure a Sim (from _afx) for workload w."""
    s = sim_cls(c=w.c, dx=w.dx, dt=w.dt, res=w.dx, precision=precision, compat=compat, **kw)
    s.set_voxels(w.vox)
    s.set_materials(w.a, w.b)
    if w.pml_width:
        s.set_pml(w.pml_width, w.pml_faces, None if compat else w.pml_profile)
    for cell, sig in zip(w.src_cells, w.src):
        s.add_source_cell(cell, sig)
    for cell in w.mic_cells:
        s.add_receiver_cell(cell)
    return s
