"""Edge cases against the CPU oracle, bit for bit, through the C ABI.

- Collisions: several sources on one cell add in list order, and several
  receivers on one cell (or on a source cell) read the same value.  This is
  checked on interior cells, a wall entry, a cell on the grid's edge plane and
  a cell inside the physical PML, on the fused, split, FP64 and two-step paths.
- Empty lists: no sources, no receivers, zero steps.
- Ragged source signals.  The reference keeps all signals in one [n][len]
  buffer (clh-vector.hpp:216-231 rejects a row of another length) and
  zero-pads it when a run outlasts it (simulation.hpp:436-441 with the
  value-initialised resize of clh-vector.hpp:237-250).  The C ABI accepts
  signals of any length and reads each one's tail as zeros, which is the
  reference's result for all signals padded to the longest; a signal of
  length 0 contributes nothing.
- The smallest grids (3 cells per axis, tiles larger than the grid).
- Argument errors: cells outside the grid, grids below 3 cells per axis.
"""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")

DX = 0.02


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _assert_bitexact(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want).astype(got.dtype)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        d = np.abs(got.astype(np.float64) - want.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} values differ, max |d| {d.max():.3e}")


def _pad(sigs, steps):
    """Ragged signals as the oracle's [nsrc][len] matrix, zero-padded."""
    n = max([steps] + [len(s) for s in sigs])
    out = np.zeros((len(sigs), n), np.float32)
    for i, s in enumerate(sigs):
        out[i, :len(s)] = s
    return out


def _run(v, a, b, sc, sigs, mc, steps, compat=False, pml=None, precision="fp32", tblock=1, split=None, **kw):
    dt = 3.125e-05 if compat else rooms.courant_dt(DX)
    c = 340.0 if compat else rooms.C_AIR
    s = afx.Sim(c=c, dx=DX, dt=dt, res=DX, precision=precision, compat=compat, **kw)
    s.set_voxels(v)
    s.set_materials(a, b)
    pml_mode = 0
    if pml is not None:
        s.set_pml(*pml)
        pml_mode = 1 if compat else 2
    s.set_time_block(tblock)
    for cell, sig in zip(sc, sigs):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    for k in (split or [steps]):
        s.run(k)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    o = orc.Oracle(v, a, b, c, DX, dt, compat, np.asarray(sc, np.int32).reshape(-1, 3),
                   _pad(sigs, steps) if len(sigs) else None, np.asarray(mc, np.int32).reshape(-1, 3),
                   "f64" if precision == "fp64" else "f32", pml_mode,
                   pml[0] if pml else 0, pml[1] if pml else 0, pml[2] if pml else None)
    o.run(steps)
    return got, o.receivers(), field, o.field(0)


def _open_room(seed):
    """136x30x36 shoebox with three boxes and an open x+ face behind a
    6-cell physical PML."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b = rooms.materials(rng, 3)
    v = rooms.shoebox(136, 30, 36, [int(rng.integers(0, 3)) for _ in range(6)], open_faces=("x+",))
    v = rooms.repair(rooms.add_boxes(v, rng, 3, 3, 7, 3))
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    return v, a, b, rng, prof


@pytest.mark.parametrize("variant", ["fused", "split", "fp64", "two_steps", "compat"])
def test_shared_cells_match_oracle(variant):
    _case(f"sources and receivers sharing cells, {variant}")
    v, a, b, rng, prof = _open_room(71)
    inner = rooms.interior_cells(v, rng, 2)
    wall = np.array([1, 15, 18], np.int32)      # air next to the x- wall: a boundary entry
    assert v[18, 15, 0] != rooms.AIR and v[18, 15, 1] == rooms.AIR
    edge = np.array([135, 12, 20], np.int32)    # on the open x+ edge plane
    in_pml = np.array([132, 14, 17], np.int32)  # inside the x+ PML slab
    steps = 133
    dt = rooms.courant_dt(DX)
    r = [rooms.ricker(steps, dt, f, amp=m) for f, m in ((1000.0, 1.0), (700.0, 0.6), (1300.0, 0.4), (900.0, 0.8))]
    sc = [inner[0], inner[0], wall, wall, in_pml, in_pml, inner[1]]
    sigs = [r[0], r[1], r[2], -r[2], r[3], r[1], r[2]]
    mc = [inner[0], inner[0], wall, edge, in_pml, inner[1], wall]
    kw = {}
    if variant == "compat":
        sigs[4], sigs[5] = r[3], r[0]
        sc[4] = sc[5] = edge  # compat mode here has no PML: use the edge plane twice instead
        res = _run(v, a, b, sc, sigs, mc, steps, compat=True)
    else:
        if variant == "split":
            kw["fused"] = False
        res = _run(v, a, b, sc, sigs, mc, steps, pml=(6, rooms.FACE_BITS["x+"], prof),
                   precision="fp64" if variant == "fp64" else "fp32",
                   tblock=2 if variant == "two_steps" else 1, **kw)
    got, want, f, wf = res
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")
    assert np.array_equal(got[0], got[1])  # one cell, one value
    assert np.abs(got).max() > 0


def test_empty_lists_and_zero_steps():
    _case("no sources, no receivers, zero steps")
    v, a, b, rng, prof = _open_room(72)
    # nothing at all: the field stays zero
    got, want, f, wf = _run(v, a, b, [], [], [], 25, pml=(6, 2, prof))
    assert got.shape == (0, 0) and want.shape == (0, 25)
    assert not np.any(f) and not np.any(wf)
    # receivers, no sources
    mc = rooms.interior_cells(v, rng, 2)
    got, want, f, wf = _run(v, a, b, [], [], mc, 25, pml=(6, 2, prof))
    assert got.shape == (2, 25) and not np.any(got)
    _assert_bitexact(got, want, "traces")
    # sources, no receivers; zero-step runs in between change nothing
    sc = rooms.interior_cells(v, rng, 1)
    sig = [rooms.ricker(40, rooms.courant_dt(DX))]
    got, want, f, wf = _run(v, a, b, sc, sig, [], 40, pml=(6, 2, prof), split=[0, 17, 0, 23, 0])
    assert got.shape == (0, 0)
    _assert_bitexact(f, wf, "field")
    assert np.any(f)


@pytest.mark.parametrize("tblock", [1, 2])
def test_ragged_source_signals(tblock):
    _case(f"source signals of length 0, 9, steps-1 and steps+40, {tblock} step(s) per pass")
    v, a, b, rng, prof = _open_room(73)
    cells = rooms.interior_cells(v, rng, 5)
    steps = 150
    full = rooms.ricker(steps + 40, rooms.courant_dt(DX), 600.0)
    sigs = [np.zeros(0, np.float32), full[:9] * 3.0, full[:steps - 1], full, full[:60]]
    got, want, f, wf = _run(v, a, b, cells[:4], sigs[:4], [cells[4], cells[1]], steps,
                            pml=(6, 2, prof), tblock=tblock)
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


@pytest.mark.parametrize("dims,walls,compat", [((3, 3, 3), (), False), ((3, 3, 3), (), True),
                                               ((4, 3, 5), (), False), ((7, 6, 5), ("x-",), False),
                                               ((9, 7, 8), ("x-", "y+"), True)])
def test_smallest_grids_match_oracle(dims, walls, compat):
    _case(f"smallest grids: {dims}, walls {walls}, compat {compat}")
    nx, ny, nz = dims
    open_faces = tuple(f for f in ("x-", "x+", "y-", "y+", "z-", "z+") if f not in walls)
    rng = np.random.Generator(np.random.PCG64(74))
    a, b = rooms.materials(rng, 2)
    v = rooms.repair(rooms.shoebox(nx, ny, nz, [1] * 6, open_faces=open_faces))
    air = np.argwhere(v == rooms.AIR)[:, ::-1].astype(np.int32)  # (x, y, z)
    centre = np.array([nx // 2, ny // 2, nz // 2], np.int32)
    corner = np.array([0, 0, 0], np.int32)
    steps = 60
    sig = [rooms.ricker(steps, 3.125e-05 if compat else rooms.courant_dt(DX))]
    mc = [centre, corner, air[-1]]
    got, want, f, wf = _run(v, a, b, [centre], sig, mc, steps, compat=compat, tblock=2)
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")
    assert np.abs(got[0]).max() > 0


def test_argument_errors():
    _case("cells outside the grid, grid too small, unknown room")
    v, a, b, rng, prof = _open_room(75)
    nz, ny, nx = v.shape
    sig = np.ones(4, np.float32)
    for bad_src, bad_mic in (([nx, 1, 1], None), ([-1, 2, 2], None), (None, [1, ny, 1]), (None, [1, 1, nz + 3])):
        s = afx.Sim(c=rooms.C_AIR, dx=DX, dt=rooms.courant_dt(DX), res=DX, compat=False)
        s.set_voxels(v)
        s.set_materials(a, b)
        with pytest.raises(afx.AfxError):
            if bad_src is not None:
                s.add_source_cell(np.array(bad_src, np.int32), sig)
            if bad_mic is not None:
                s.add_receiver_cell(np.array(bad_mic, np.int32))
            s.run(2)
        s.close()
    s = afx.Sim(c=rooms.C_AIR, dx=DX, dt=rooms.courant_dt(DX), res=DX, compat=False)
    with pytest.raises(afx.AfxError):
        s.set_voxels(np.full((3, 2, 3), rooms.AIR, np.int32))
    s.set_voxels(v)
    s.set_materials(a, b)
    with pytest.raises(afx.AfxError):
        s.add_source_cell(np.array([2, 2, 2], np.int32), sig, room=3)
        s.run(1)
    s.close()
