"""ctypes wrapper of the CPU oracle (oracle/liborc.so) — test infrastructure.

Mirrors the library's inputs (voxels, materials, sources, receivers, PML)
so a test can run the same workload through both and compare.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


class Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("vox", _i32p), ("ncoef", C.c_int32), ("nmat", C.c_int32),
        ("a", _f32p), ("b", _f32p),
        ("c", C.c_float), ("dx", C.c_float), ("dt", C.c_float),
        ("compat", C.c_int32),
        ("nsrc", C.c_int32), ("src_cell", _i32p), ("src", _f32p), ("src_len", C.c_int64),
        ("nmic", C.c_int32), ("mic_cell", _i32p),
        ("pml_mode", C.c_int32), ("pml_width", C.c_int32),
        ("ksi", _f64p * 3), ("ksi_h", _f64p * 3),
        ("pml_boxes", _i32p), ("pml_nbox", C.c_int32),
        ("store", C.c_int32), ("scale", C.c_double),
        ("vol", C.c_int32 * 3),
    ]


def load_lib(path=ORC_PATH):
    """The oracle library at `path` with its argument types declared (a
    variant built from edited sources loads the same way)."""
    if path == ORC_PATH and not os.path.exists(ORC_PATH):
        subprocess.run([os.path.join(ROOT, "oracle", "build.sh")], check=True)
    lib = C.CDLL(path)
    P = C.c_void_p
    for sfx, T in (("f32", C.c_float), ("f64", C.c_double)):
        Tp = C.POINTER(T)
        getattr(lib, f"orc_create_{sfx}").restype = P
        getattr(lib, f"orc_create_{sfx}").argtypes = [C.POINTER(Problem)]
        getattr(lib, f"orc_run_{sfx}").restype = C.c_int
        getattr(lib, f"orc_run_{sfx}").argtypes = [P, C.c_int64, Tp]
        getattr(lib, f"orc_field_{sfx}").argtypes = [P, C.c_int, Tp]
        getattr(lib, f"orc_set_field_{sfx}").argtypes = [P, C.c_int, Tp]
        getattr(lib, f"orc_destroy_{sfx}").argtypes = [P]
        getattr(lib, f"orc_nbound_{sfx}").restype = C.c_int64
        getattr(lib, f"orc_nbound_{sfx}").argtypes = [P]
        getattr(lib, f"orc_rings_{sfx}").argtypes = [P, Tp, Tp]
        getattr(lib, f"orc_ring_index_{sfx}").restype = C.c_int32
        getattr(lib, f"orc_ring_index_{sfx}").argtypes = [P]
    lib.orc_classify.restype = C.c_int64
    lib.orc_classify_range.restype = C.c_int64
    lib.orc_classify_range.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32p, C.c_int32, C.c_int32, C.c_int32,
                                       _i32p, _i32p, C.c_int64]
    lib.orc_classify.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p, _i32p, C.c_int64]
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_round_store.restype = C.c_float
    lib.orc_round_store.argtypes = [C.c_float, C.c_int32]
    return lib


lib = load_lib()

STORE_MODES = {None: 0, "fp16": 1, "bf16": 2}


def round_store(x, storage):
    """float32 values rounded to the storage type and widened back."""
    mode = STORE_MODES[storage]
    x = np.asarray(x, np.float32)
    return np.array([lib.orc_round_store(float(v), mode) for v in x.ravel()], np.float32).reshape(x.shape)


def storage_scale(src):
    """The power-of-two scale the library picks for FP16/BF16 storage: the
    smallest 2^k at or above the sources' peak amplitude."""
    mx = float(np.max(np.abs(np.asarray(src, np.float32)))) if np.size(src) else 0.0
    return math.ldexp(1.0, math.ceil(math.log2(mx))) if mx > 0 else 1.0


def classify(vox, vol=None):
    """Reference-order boundary list: (pos[n,6], group[n]) or raises on an
    unphysical pattern.  vol = (vx, vy, vz) limits the scan to that voxel
    volume inside the grid."""
    v = np.ascontiguousarray(vox, np.int32)
    nz, ny, nx = v.shape
    cap = max(int(v.size), 1)
    pos = np.zeros((cap, 6), np.int32)
    grp = np.zeros(cap, np.int32)
    lx, ly, lz = vol if vol is not None else (nx, ny, nz)
    n = lib.orc_classify_range(nx, ny, nz, v.ctypes.data_as(_i32p), lx, ly, lz, pos.ctypes.data_as(_i32p),
                               grp.ctypes.data_as(_i32p), cap)
    if n < 0:
        raise ValueError(f"unphysical pattern at cell {tuple(pos[0, :3])}")
    return pos[:n].copy(), grp[:n].copy()


def compat_ksi(dims, profile):
    """Per-axis kappa arrays as SetPml builds them (profile reversed at both ends)."""
    dmax = max(dims) + 2
    out = []
    for n in dims:
        k = np.zeros(dmax, np.float64)
        L = len(profile)
        for i in range(min(L, n)):
            k[i] = profile[L - 1 - i]
            k[n - i - 1] = profile[L - 1 - i]
        out.append(k)
    return out


def subtract_box(a, b):
    if a[0] >= b[1] or a[1] <= b[0] or a[2] >= b[3] or a[3] <= b[2] or a[4] >= b[5] or a[5] <= b[4]:
        return [a]
    r = list(a)
    out = []
    for d in range(3):
        if r[2 * d] < b[2 * d]:
            c = list(r); c[2 * d + 1] = b[2 * d]; out.append(c); r[2 * d] = b[2 * d]
        if r[2 * d + 1] > b[2 * d + 1]:
            c = list(r); c[2 * d] = b[2 * d + 1]; out.append(c); r[2 * d + 1] = b[2 * d + 1]
    return out


def phys_pml(dims, width, faces, profile):
    """Zeta arrays (cell centres, +1/2 points) and disjoint slab boxes of the
    physical PML for an explicit per-depth profile (length == width)."""
    w = width
    prof = np.asarray(profile, np.float32).astype(np.float64)

    def at(k):
        if k < 0:
            return 0.0
        return float(prof[w - 1]) if k >= w else float(prof[k])

    def pr(depth):
        if depth <= 0:
            return 0.0
        u = depth - 0.5
        lo = int(np.floor(u))
        f = u - lo
        return (1.0 - f) * at(lo) + f * at(lo + 1)

    zc, zh, boxes = [], [], []
    for d in range(3):
        n = dims[d]
        lo_on, hi_on = (faces >> (2 * d)) & 1, (faces >> (2 * d + 1)) & 1
        c_arr = np.zeros(n + 1, np.float64)
        h_arr = np.zeros(n + 1, np.float64)
        for i in range(n):
            vc = vh = 0.0
            if hi_on:
                vc = max(vc, pr(i - (n - 1 - w) + 0.5))
                vh = max(vh, pr(i - (n - 1 - w) + 1.0))
            if lo_on:
                vc = max(vc, pr((w + 0.5) - i))
                vh = max(vh, pr(float(w - i)))
            c_arr[i] = vc
            h_arr[i] = vh
        zc.append(c_arr)
        zh.append(h_arr)
        full = [1, dims[0] - 1, 1, dims[1] - 1, 1, dims[2] - 1]
        for side in (0, 1):
            if not (hi_on if side else lo_on):
                continue
            slab = list(full)
            if side:
                slab[2 * d] = n - 1 - w
            else:
                slab[2 * d + 1] = 1 + w
            pieces = [slab]
            for b in boxes:
                nxt = []
                for p in pieces:
                    nxt += subtract_box(p, b)
                pieces = nxt
            boxes += [p for p in pieces if p[1] > p[0] and p[3] > p[2] and p[5] > p[4]]
    return zc, zh, np.array(boxes, np.int32).reshape(-1, 6)


class Oracle:
    """One oracle simulation; precision 'f32' or 'f64'."""

    def __init__(self, vox, a, b, c, dx, dt, compat, src_cells=(), src=None, mic_cells=(), precision="f32",
                 pml_mode=0, pml_width=0, pml_faces=0, pml_profile=None, dims=None, threads=8,
                 storage=None, scale=None, vol=None, lib=None, compat_kappa=None):
        self.lib = lib = lib if lib is not None else globals()["lib"]
        self.sfx = precision
        self.T = np.float32 if precision == "f32" else np.float64
        self._keep = []
        pb = Problem()
        if vox is not None:
            v = np.ascontiguousarray(vox, np.int32)
            nz, ny, nx = v.shape
            pb.vox = v.ctypes.data_as(_i32p)
            self._keep.append(v)
        else:
            nz, ny, nx = dims
            pb.vox = None
        self.dims = (nz, ny, nx)
        pb.nx, pb.ny, pb.nz = nx, ny, nz
        if a is not None:
            A = np.ascontiguousarray(a, np.float32)
            B = np.ascontiguousarray(b, np.float32)
            pb.nmat, pb.ncoef = A.shape
            pb.a = A.ctypes.data_as(_f32p)
            pb.b = B.ctypes.data_as(_f32p)
            self._keep += [A, B]
        pb.c, pb.dx, pb.dt = c, dx, dt
        pb.compat = 1 if compat else 0
        sc = np.ascontiguousarray(np.asarray(src_cells, np.int32).reshape(-1, 3))
        pb.nsrc = sc.shape[0]
        if pb.nsrc:
            S = np.ascontiguousarray(np.asarray(src, np.float32).reshape(pb.nsrc, -1))
            pb.src_cell = sc.ctypes.data_as(_i32p)
            pb.src = S.ctypes.data_as(_f32p)
            pb.src_len = S.shape[1]
            self._keep += [sc, S]
        mc = np.ascontiguousarray(np.asarray(mic_cells, np.int32).reshape(-1, 3))
        pb.nmic = mc.shape[0]
        self.nmic = pb.nmic
        if pb.nmic:
            pb.mic_cell = mc.ctypes.data_as(_i32p)
            self._keep.append(mc)
        pb.pml_mode = pml_mode
        pb.pml_width = pml_width
        if pml_mode == 1:
            ks = compat_ksi((nx, ny, nz), [] if pml_profile is None else list(np.asarray(pml_profile, np.float32)))
            if compat_kappa is not None:  # explicit per-axis kappa arrays (x, y, z)
                ks = [np.ascontiguousarray(np.pad(np.asarray(k, np.float64), (0, max(nx, ny, nz) + 2 - len(k))))
                      for k in compat_kappa]
            for d in range(3):
                pb.ksi[d] = ks[d].ctypes.data_as(_f64p)
            self._keep += ks
        elif pml_mode == 2:
            zc, zh, boxes = phys_pml((nx, ny, nz), pml_width, pml_faces, pml_profile)
            for d in range(3):
                pb.ksi[d] = zc[d].ctypes.data_as(_f64p)
                pb.ksi_h[d] = zh[d].ctypes.data_as(_f64p)
            boxes = np.ascontiguousarray(boxes)
            pb.pml_boxes = boxes.ctypes.data_as(_i32p)
            pb.pml_nbox = boxes.shape[0]
            self._keep += zc + zh + [boxes]
        # FP16/BF16 field storage: f32 arithmetic, values rounded on every
        # store, fields in units of pressure / scale
        pb.store = STORE_MODES[storage]
        self.scale = float(scale) if scale else (storage_scale(src) if storage else 1.0)
        pb.scale = self.scale
        if vol is not None:
            pb.vol[0], pb.vol[1], pb.vol[2] = vol
        self._pb = pb
        lib.orc_set_threads(threads)
        self._h = getattr(lib, f"orc_create_{self.sfx}")(C.byref(pb))
        if not self._h:
            raise ValueError("oracle rejected the geometry (unphysical pattern)")
        self.traces = [np.zeros((self.nmic, 0), self.T)]

    def run(self, steps):
        tr = np.zeros((max(self.nmic, 1), steps), self.T)
        ptr = tr.ctypes.data_as(C.POINTER(C.c_float if self.T == np.float32 else C.c_double))
        getattr(self.lib, f"orc_run_{self.sfx}")(self._h, steps, ptr)
        self.traces.append(tr[:self.nmic])
        return tr[:self.nmic]

    def receivers(self):
        return np.concatenate(self.traces, axis=1)

    def field(self, which=0):
        out = np.zeros(self.dims, self.T)
        ptr = out.ctypes.data_as(C.POINTER(C.c_float if self.T == np.float32 else C.c_double))
        getattr(self.lib, f"orc_field_{self.sfx}")(self._h, which, ptr)
        return out * self.T(self.scale)  # pressure; the scale is a power of two

    def set_field(self, which, values):
        """Overwrites pressure field `which` (0 newest, 1 p^n, 2 p^{n-1}) with
        `values` [nz][ny][nx] in pressure units."""
        v = np.ascontiguousarray(np.asarray(values, np.float64) / self.scale, self.T)
        assert v.shape == self.dims
        ptr = v.ctypes.data_as(C.POINTER(C.c_float if self.T == np.float32 else C.c_double))
        getattr(self.lib, f"orc_set_field_{self.sfx}")(self._h, which, ptr)

    def nbound(self):
        return getattr(self.lib, f"orc_nbound_{self.sfx}")(self._h)

    def rings(self):
        """IIR state (ring_in, ring_out), each [nb][3*ncoef], boundary cells in classify order."""
        nb = self.nbound()
        n3 = 3 * max(int(self._pb.ncoef), 1)
        rin = np.zeros((nb, n3), self.T)
        rout = np.zeros((nb, n3), self.T)
        P = C.POINTER(C.c_float if self.T == np.float32 else C.c_double)
        getattr(self.lib, f"orc_rings_{self.sfx}")(self._h, rin.ctypes.data_as(P), rout.ctypes.data_as(P))
        return rin, rout

    def ring_index(self):
        """Ring slot the last step wrote."""
        return int(getattr(self.lib, f"orc_ring_index_{self.sfx}")(self._h))

    def close(self):
        if self._h:
            getattr(self.lib, f"orc_destroy_{self.sfx}")(self._h)
            self._h = None

    def __del__(self):
        self.close()


def from_workload(w, compat=False, precision="f32", pml_profile=None, steps=None, threads=8):
    pml_mode = 0
    if w.pml_width:
        pml_mode = 1 if compat else 2
        if pml_profile is None and not compat:
            pml_profile = w.pml_profile
    return Oracle(w.vox, w.a, w.b, w.c, w.dx, w.dt, compat, w.src_cells, w.src, w.mic_cells, precision,
                  pml_mode, w.pml_width, w.pml_faces, pml_profile, threads=threads)
