"""ctypes binding of libafx.so (the C ABI declared in include/afx.h).

There is no fallback: if the compiled library is missing, importing this
module raises.  Build it with ``python anc-field_amd/build.py``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("AFX_LIB", os.path.join(_HERE, "..", "lib", "libafx.so"))

AFX_OK = 0
AFX_E_ARG = -1
AFX_E_STATE = -2
AFX_E_GEOMETRY = -3
AFX_E_IO = -4
AFX_E_HIP = -5
AFX_E_NOMEM = -6
AFX_E_RANGE = -7

FP32, FP64, FP16, BF16 = 0, 1, 2, 3
PRECISIONS = {"fp32": FP32, "fp64": FP64, "fp16": FP16, "bf16": BF16}
COMPAT = 1
NO_GRAPH = 2
SPLIT = 4
LIST_BLOCKS = 8
PML_BLOCKS = 16
FACE_XM, FACE_XP, FACE_YM, FACE_YP, FACE_ZM, FACE_ZP = 1, 2, 4, 8, 16, 32


class AfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"afx error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("c", C.c_float), ("dx", C.c_float), ("dt", C.c_float), ("res", C.c_float),
        ("precision", C.c_int32), ("flags", C.c_uint32), ("device", C.c_int32),
        ("pml_order", C.c_float), ("pml_r0", C.c_float), ("scale", C.c_float),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("cells", C.c_int64), ("boundary", C.c_int64), ("faces", C.c_int64), ("edges", C.c_int64),
        ("corners", C.c_int64), ("group_count", C.c_int64 * 26), ("steps", C.c_int64),
        ("rooms", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("tile_x", C.c_int32), ("tile_y", C.c_int32), ("tile_z", C.c_int32), ("fused", C.c_int32),
        ("last_run_ms", C.c_double), ("last_run_gpu_ms", C.c_double), ("last_run_steps", C.c_int64),
    ]


class Profile(C.Structure):
    _fields_ = [("step_ms", C.c_double), ("kernel_ms", C.c_double * 4), ("launches", C.c_int64 * 4),
                ("steps", C.c_int64)]


if not os.path.exists(LIB_PATH):
    raise ImportError(f"libafx.so not found at {LIB_PATH}: build it with python anc-field_amd/build.py")
lib = C.CDLL(LIB_PATH)

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_P = C.c_void_p

_SIGS = {
    "afx_config_default": (C.c_int, [C.POINTER(Config)]),
    "afx_device_count": (C.c_int, [_i32p]),
    "afx_device_name": (C.c_int, [C.c_int32, C.c_char_p, C.c_size_t]),
    "afx_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "afx_destroy": (None, [_P]),
    "afx_last_error": (C.c_char_p, []),
    "afx_load_kerkythea": (C.c_int, [_P, C.c_char_p]),
    "afx_set_min_size": (C.c_int, [_P, C.c_float, C.c_float, C.c_float]),
    "afx_set_voxels": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _i32p]),
    "afx_set_materials": (C.c_int, [_P, _f32p, _f32p, C.c_int32, C.c_int32]),
    "afx_set_pml": (C.c_int, [_P, C.c_int32, C.c_uint32, _f32p, C.c_int32]),
    "afx_add_source": (C.c_int, [_P, _f32p, _f32p, C.c_int64]),
    "afx_add_receiver": (C.c_int, [_P, _f32p]),
    "afx_add_source_cell": (C.c_int, [_P, C.c_int32, _i32p, _f32p, C.c_int64]),
    "afx_add_receiver_cell": (C.c_int, [_P, C.c_int32, _i32p]),
    "afx_add_room": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _i32p, _f32p, _f32p, C.c_int32, C.c_int32, _i32p]),
    "afx_run": (C.c_int, [_P, C.c_int64]),
    "afx_synchronize": (C.c_int, [_P]),
    "afx_voxelize_kerkythea": (C.c_int, [C.c_char_p, C.c_float, _i32p, _i32p, C.c_int64, _f32p, _f32p, C.c_int32,
                                         _i32p, _i32p]),
    "afx_classify": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _i32p, C.c_int32, C.POINTER(C.c_int64), _i32p]),
    "afx_get_receivers": (C.c_int, [_P, _f32p, C.c_int64, C.POINTER(C.c_int64)]),
    "afx_get_pressure": (C.c_int, [_P, C.c_int32, C.c_int32, _f32p, C.c_int64]),
    "afx_get_plane": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _f32p, C.c_int64]),
    "afx_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "afx_autotune": (C.c_int, [_P, C.c_int32, _i32p]),
    "afx_profile_run": (C.c_int, [_P, C.c_int64, C.POINTER(Profile)]),
    "afx_set_tile": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "afx_set_time_block": (C.c_int, [_P, C.c_int32]),
    "afx_get_time_block": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "afx_set_tile_split": (C.c_int, [_P, C.c_int32]),
    "afx_get_tile_split": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "afx_copy_probe": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def _check(rc):
    if rc != AFX_OK:
        raise AfxError(rc, lib.afx_last_error().decode(errors="replace"))
    return rc


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_f32p)


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def device_count():
    n = C.c_int32(0)
    _check(lib.afx_device_count(C.byref(n)))
    return n.value


def device_name(i):
    buf = C.create_string_buffer(256)
    _check(lib.afx_device_name(i, buf, 256))
    return buf.value.decode()


def copy_probe(nbytes, reps=20, device=0):
    """Device copy rate in GB/s (read + write) of a plain copy of nbytes."""
    g = C.c_double(0.0)
    _check(lib.afx_copy_probe(device, int(nbytes), int(reps), C.byref(g)))
    return g.value


def voxelize_kerkythea(path, res=0.02):
    """(vox[nz, ny, nx] int32, a[nmat, ncoef], b[nmat, ncoef]) that `load` builds."""
    dims = (C.c_int32 * 3)()
    nc, nm = C.c_int32(0), C.c_int32(0)
    _check(lib.afx_voxelize_kerkythea(os.fsencode(path), res, dims, None, 0, None, None, 0, C.byref(nc), C.byref(nm)))
    nx, ny, nz = dims
    vox = np.empty((nz, ny, nx), np.int32)
    a = np.empty((nm.value, nc.value), np.float32)
    b = np.empty_like(a)
    _check(lib.afx_voxelize_kerkythea(os.fsencode(path), res, dims, vox.ctypes.data_as(_i32p), vox.size,
                                      a.ctypes.data_as(_f32p), b.ctypes.data_as(_f32p), a.size, C.byref(nc),
                                      C.byref(nm)))
    return vox, a, b


def classify_counts(vox, compat=True):
    """Boundary cells per reference group id (26 counts)."""
    v, vp = _i32(vox)
    nz, ny, nx = v.shape
    counts = (C.c_int64 * 26)()
    bad = (C.c_int32 * 3)()
    _check(lib.afx_classify(nx, ny, nz, vp, 1 if compat else 0, counts, bad))
    return np.array(counts[:], np.int64)


def default_config():
    cfg = Config()
    _check(lib.afx_config_default(C.byref(cfg)))
    return cfg


class Sim:
    """One simulation handle (one or more rooms stepped together)."""

    def __init__(self, c=None, dx=None, dt=None, res=None, precision="fp32", compat=True, device=0,
                 graph=True, fused=True, pml_order=None, pml_r0=None, scale=None, list_blocks=False,
                 pml_blocks=False):
        cfg = default_config()
        if c is not None:
            cfg.c = c
        if dx is not None:
            cfg.dx = dx
        if dt is not None:
            cfg.dt = dt
        if res is not None:
            cfg.res = res
        cfg.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg.flags = ((COMPAT if compat else 0) | (0 if graph else NO_GRAPH) | (0 if fused else SPLIT)
                     | (LIST_BLOCKS if list_blocks else 0) | (PML_BLOCKS if pml_blocks else 0))
        cfg.device = device
        if pml_order is not None:
            cfg.pml_order = pml_order
        if pml_r0 is not None:
            cfg.pml_r0 = pml_r0
        if scale is not None:
            cfg.scale = scale
        self.cfg = cfg
        h = _P()
        _check(lib.afx_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.nmic = 0
        self.rooms = 1

    def close(self):
        if getattr(self, "_h", None):
            lib.afx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def load_kerkythea(self, path):
        _check(lib.afx_load_kerkythea(self._h, os.fsencode(path)))

    def set_min_size(self, x, y, z):
        _check(lib.afx_set_min_size(self._h, x, y, z))

    def set_voxels(self, vox=None, dims=None):
        if vox is None:
            nz, ny, nx = dims
            _check(lib.afx_set_voxels(self._h, nx, ny, nz, None))
            return
        v, p = _i32(vox)
        nz, ny, nx = v.shape
        _check(lib.afx_set_voxels(self._h, nx, ny, nz, p))

    def set_materials(self, a, b):
        a, pa = _f32(a)
        b, pb = _f32(b)
        nmat, ncoef = a.shape
        _check(lib.afx_set_materials(self._h, pa, pb, ncoef, nmat))

    def set_pml(self, width, faces=0, profile=None):
        if profile is None:
            _check(lib.afx_set_pml(self._h, width, faces, None, 0))
        else:
            pr, pp = _f32(profile)
            _check(lib.afx_set_pml(self._h, width, faces, pp, pr.size))

    def add_source(self, pos, samples):
        p, pp = _f32(pos)
        s, sp = _f32(samples)
        _check(lib.afx_add_source(self._h, pp, sp, s.size))

    def add_receiver(self, pos):
        p, pp = _f32(pos)
        _check(lib.afx_add_receiver(self._h, pp))
        self.nmic += 1

    def add_source_cell(self, cell, samples, room=0):
        c, cp = _i32(cell)
        s, sp = _f32(samples)
        _check(lib.afx_add_source_cell(self._h, room, cp, sp, s.size))

    def add_receiver_cell(self, cell, room=0):
        c, cp = _i32(cell)
        _check(lib.afx_add_receiver_cell(self._h, room, cp))
        self.nmic += 1

    def add_room(self, vox, a=None, b=None, dims=None):
        rid = C.c_int32(0)
        if vox is None:
            nz, ny, nx = dims
            _check(lib.afx_add_room(self._h, nx, ny, nz, None, None, None, 0, 0, C.byref(rid)))
        else:
            v, vp = _i32(vox)
            a, pa = _f32(a)
            b, pb = _f32(b)
            nz, ny, nx = v.shape
            _check(lib.afx_add_room(self._h, nx, ny, nz, vp, pa, pb, a.shape[1], a.shape[0], C.byref(rid)))
        self.rooms += 1
        return rid.value

    def run(self, steps):
        _check(lib.afx_run(self._h, int(steps)))

    def synchronize(self):
        _check(lib.afx_synchronize(self._h))

    def receivers(self):
        n = C.c_int64(0)
        _check(lib.afx_get_receivers(self._h, None, 0, C.byref(n)))
        if self.nmic == 0 or n.value == 0:
            return np.zeros((self.nmic, 0), np.float32)
        out = np.empty(n.value, np.float32)
        _check(lib.afx_get_receivers(self._h, out.ctypes.data_as(_f32p), out.size, C.byref(n)))
        return out.reshape(self.nmic, -1)

    def pressure(self, which=0, room=0):
        st = self.stats()
        dims = self.room_dims(room) if room else (st.nz, st.ny, st.nx)
        out = np.empty(int(np.prod(dims)), np.float32)
        _check(lib.afx_get_pressure(self._h, room, which, out.ctypes.data_as(_f32p), out.size))
        return out.reshape(dims)

    def plane(self, axis, index, dims, which=0, room=0):
        """FP32 plane of room `room` (dims = (nz, ny, nx)): axis 0 -> [nz][ny],
        1 -> [nz][nx] (the reference's cross section), 2 -> [ny][nx]."""
        nz, ny, nx = dims
        shape = {0: (nz, ny), 1: (nz, nx), 2: (ny, nx)}[axis]
        out = np.empty(shape, np.float32)
        _check(lib.afx_get_plane(self._h, room, axis, index, which, out.ctypes.data_as(_f32p), out.size))
        return out

    def room_dims(self, room):
        raise NotImplementedError("pass dims for rooms other than 0")

    def pressure_dims(self, dims, which=0, room=0):
        out = np.empty(int(np.prod(dims)), np.float32)
        _check(lib.afx_get_pressure(self._h, room, which, out.ctypes.data_as(_f32p), out.size))
        return out.reshape(dims)

    def stats(self):
        st = Stats()
        _check(lib.afx_get_stats(self._h, C.byref(st)))
        return st

    def profile(self, steps):
        pr = Profile()
        _check(lib.afx_profile_run(self._h, int(steps), C.byref(pr)))
        return pr

    def set_tile(self, tx, ty, tz):
        _check(lib.afx_set_tile(self._h, tx, ty, tz))

    def autotune(self, steps_per_trial=30):
        """Fastest tile shape (and steps per pass, see time_block) for this
        handle's rooms; returns the tile (tx, ty, tz)."""
        t = (C.c_int32 * 3)()
        _check(lib.afx_autotune(self._h, steps_per_trial, t))
        return tuple(t)

    def set_time_block(self, steps):
        """1: one step per pass; 2: two steps per pass where tiles allow it
        (bit-identical results)."""
        _check(lib.afx_set_time_block(self._h, steps))

    def set_tile_split(self, split):
        """0: whole tiles; 1 / 2: tiles in the top quarter / half of each
        room's planes run as two halves (bit-identical results)."""
        _check(lib.afx_set_tile_split(self._h, split))

    def tile_split(self):
        k = C.c_int32()
        _check(lib.afx_get_tile_split(self._h, C.byref(k)))
        return k.value

    def time_block(self):
        """(steps per pass, tiles taking two steps per pass)."""
        k = C.c_int32()
        n = C.c_int64()
        _check(lib.afx_get_time_block(self._h, C.byref(k), C.byref(n)))
        return k.value, n.value


# ---- surrogate MLP (include/afx_mlp.h) ------------------------------------

class MlpConfig(C.Structure):
    _fields_ = [
        ("in_dim", C.c_int32), ("n_freq", C.c_int32), ("width", C.c_int32), ("depth", C.c_int32),
        ("out_dim", C.c_int32), ("batch", C.c_int32), ("device", C.c_int32), ("seed", C.c_uint32),
        ("freq_min", C.c_float), ("freq_max", C.c_float),
        ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("weight_decay", C.c_float),
        ("flags", C.c_uint32),
    ]


MLP_NO_GRAPH = 1


class MlpStats(C.Structure):
    _fields_ = [
        ("params", C.c_int64), ("train_steps", C.c_int64), ("flop_per_step", C.c_double),
        ("last_train_ms", C.c_double), ("last_train_gpu_ms", C.c_double),
        ("last_forward_ms", C.c_double), ("last_forward_gpu_ms", C.c_double), ("last_forward_rows", C.c_int64),
    ]


_MLP_SIGS = {
    "afx_mlp_config_default": (C.c_int, [C.POINTER(MlpConfig)]),
    "afx_mlp_create": (C.c_int, [C.POINTER(MlpConfig), C.POINTER(_P)]),
    "afx_mlp_destroy": (None, [_P]),
    "afx_mlp_param_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "afx_mlp_get_params": (C.c_int, [_P, _f32p, C.c_int64]),
    "afx_mlp_set_params": (C.c_int, [_P, _f32p, C.c_int64]),
    "afx_mlp_get_frequencies": (C.c_int, [_P, _f32p, C.c_int32]),
    "afx_mlp_forward": (C.c_int, [_P, _f32p, C.c_int64, _f32p]),
    "afx_mlp_forward_device": (C.c_int, [_P, C.c_void_p, C.c_int64, C.c_void_p]),
    "afx_mlp_gradients": (C.c_int, [_P, _f32p, _f32p, _f32p, C.c_int64, _f32p, C.POINTER(C.c_double), _f32p]),
    "afx_mlp_set_data": (C.c_int, [_P, _f32p, _f32p, _f32p, C.c_int64]),
    "afx_mlp_train": (C.c_int, [_P, C.c_int64, C.c_float, C.c_float, C.c_uint32, _f32p]),
    "afx_mlp_train_noalloc": (C.c_int, [_P, C.c_int64, C.c_float, C.c_float, C.c_uint32, _f32p]),
    "afx_mlp_get_stats": (C.c_int, [_P, C.POINTER(MlpStats)]),
}
for _name, (_res, _args) in _MLP_SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = EXPORTED + tuple(_MLP_SIGS)


def mlp_default_config():
    cfg = MlpConfig()
    _check(lib.afx_mlp_config_default(C.byref(cfg)))
    return cfg


class Mlp:
    """One surrogate network on one device (BF16 MFMA kernels, FP32 master weights)."""

    def __init__(self, **kw):
        cfg = mlp_default_config()
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.cfg = cfg
        h = _P()
        _check(lib.afx_mlp_create(C.byref(cfg), C.byref(h)))
        self._h = h
        n = C.c_int64(0)
        _check(lib.afx_mlp_param_count(h, C.byref(n)))
        self.nparam = n.value

    def close(self):
        if getattr(self, "_h", None):
            lib.afx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self):
        out = np.empty(self.nparam, np.float32)
        _check(lib.afx_mlp_get_params(self._h, out.ctypes.data_as(_f32p), out.size))
        return out

    def set_params(self, p):
        p, pp = _f32(p)
        _check(lib.afx_mlp_set_params(self._h, pp, p.size))

    def frequencies(self):
        out = np.empty(self.cfg.n_freq, np.float32)
        _check(lib.afx_mlp_get_frequencies(self._h, out.ctypes.data_as(_f32p), out.size))
        return out

    def forward(self, x):
        x, xp = _f32(x)
        rows = x.shape[0]
        y = np.empty((rows, self.cfg.out_dim), np.float32)
        _check(lib.afx_mlp_forward(self._h, xp, rows, y.ctypes.data_as(_f32p)))
        return y

    def forward_device(self, x_dev, rows, y_dev):
        """Outputs for `rows` queries already in device memory (DeviceBuffer
        or raw device addresses); y_dev receives [rows][out_dim] FP32."""
        _check(lib.afx_mlp_forward_device(self._h, int(x_dev), int(rows), int(y_dev)))

    def gradients(self, x, target, weight):
        x, xp = _f32(x)
        t, tp = _f32(target)
        w, wp = _f32(weight)
        rows = x.shape[0]
        g = np.empty(self.nparam, np.float32)
        y = np.empty((rows, self.cfg.out_dim), np.float32)
        loss = C.c_double(0.0)
        _check(lib.afx_mlp_gradients(self._h, xp, tp, wp, rows, g.ctypes.data_as(_f32p), C.byref(loss),
                                     y.ctypes.data_as(_f32p)))
        return g, loss.value, y

    def set_data(self, x, target, weight):
        x, xp = _f32(x)
        t, tp = _f32(target)
        w, wp = _f32(weight)
        _check(lib.afx_mlp_set_data(self._h, xp, tp, wp, x.shape[0]))

    def train(self, steps, lr0, lr1=0.0, seed=0):
        loss = np.empty(int(steps), np.float32)
        _check(lib.afx_mlp_train(self._h, int(steps), lr0, lr1, seed, loss.ctypes.data_as(_f32p)))
        return loss

    def train_noalloc(self, steps, lr0, lr1=0.0, seed=0):
        """train() without device allocation (fixed per-GPU scratch, eager steps)."""
        loss = np.empty(int(steps), np.float32)
        _check(lib.afx_mlp_train_noalloc(self._h, int(steps), lr0, lr1, seed, loss.ctypes.data_as(_f32p)))
        return loss

    def stats(self):
        st = MlpStats()
        _check(lib.afx_mlp_get_stats(self._h, C.byref(st)))
        return st


# ---- device memory for the device-pointer entry points ----------------------
# The library's own HIP runtime (the one libafx.so links) allocates it, so no
# second runtime enters the process.

_hip = None


def _hip_lib():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipSetDevice.argtypes = [C.c_int]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipSetDevice):
            f.restype = C.c_int
        _hip = h
    return _hip


class DeviceBuffer:
    """`nbytes` of device memory on `device`; int(buf) is its address."""

    def __init__(self, nbytes, device=0):
        h = _hip_lib()
        if h.hipSetDevice(device) != 0:
            raise AfxError(-1, f"hipSetDevice({device}) failed")
        p = C.c_void_p()
        if h.hipMalloc(C.byref(p), nbytes) != 0:
            raise AfxError(-1, f"hipMalloc of {nbytes} bytes failed")
        self.ptr, self.nbytes = p.value, nbytes

    def __int__(self):
        return self.ptr

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if _hip_lib().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:  # host to device
            raise AfxError(-1, "hipMemcpy to device failed")

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        if _hip_lib().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:  # device to host
            raise AfxError(-1, "hipMemcpy to host failed")
        return out

    def close(self):
        if self.ptr:
            _hip_lib().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
