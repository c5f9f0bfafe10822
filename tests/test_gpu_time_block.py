"""Two steps per pass (time-step blocking) against the CPU oracle and against
one step per pass, through the C ABI.

A tile takes two steps in one pass only when its cells and one-cell halo are
all interior cells; every other cell takes them in two launches, and four
pressure fields rotate.  Everything must stay bit-identical to single steps:
odd step counts (a single step after the pairs), graph chunks, a source and
receiver on the grid's edge plane (whose stored value is carried from the
buffer a single step would write into), the physical PML, batched rooms,
every storage precision, plane snapshots, and switching mid-run.
"""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")

PAIR_TILES = [(64, 4, 16), (64, 4, 32), (64, 8, 16), (64, 2, 32), (128, 2, 16), (64, 4, 8)]


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _assert_bitexact(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want).astype(got.dtype)
    if not np.array_equal(got, want):
        d = np.abs(got.astype(np.float64) - want.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} values differ, max |d| {d.max():.3e}")


def _room():
    """264x40x70 shoebox, three boxes, open x+ face with a 6-cell PML: large
    enough that every tile shape has tiles whose halo avoids walls and PML."""
    rng = np.random.Generator(np.random.PCG64(41))
    nx, ny, nz = 264, 40, 70
    a, b = rooms.materials(rng, 3)
    v = rooms.shoebox(nx, ny, nz, [int(rng.integers(0, 3)) for _ in range(6)])
    v = rooms.add_boxes(v, rng, 3, 4, 9, 3)
    v[:, :, -1] = -1
    v = rooms.repair(v)
    inner = rooms.interior_cells(v, rng, 4)
    plane = np.array([0, 5, 5], np.int32)  # on the x = 0 edge plane
    sc = np.stack([inner[0], plane])
    mc = np.stack([inner[1], inner[2], inner[3], plane])
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    return v, a, b, sc, mc, prof


def _sim(v, a, b, sc, mc, prof, src, tile=None, tblock=1, precision="fp32", **kw):
    dt = rooms.courant_dt(0.02)
    s = afx.Sim(c=rooms.C_AIR, dx=0.02, dt=dt, res=0.02, compat=False, precision=precision, **kw)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_pml(6, 2, prof)
    if tile:
        s.set_tile(*tile)
    s.set_time_block(tblock)
    for cell, sig in zip(sc, src):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    return s


def _src(steps):
    dt = rooms.courant_dt(0.02)
    return np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 800.0, amp=0.5)])


@pytest.mark.parametrize("tile", PAIR_TILES)
def test_two_steps_per_pass_match_oracle(tile):
    _case(f"two steps per pass, tile {tile}, 121 steps, PML x+, edge-plane source")
    v, a, b, sc, mc, prof = _room()
    steps = 121
    src = _src(steps)
    s = _sim(v, a, b, sc, mc, prof, src, tile=tile, tblock=2)
    s.run(steps)
    k, npair = s.time_block()
    got, f = s.receivers(), s.pressure(0)
    s.close()
    print(f"tiles taking two steps per pass: {npair}")
    assert k == 2 and npair > 0
    dt = rooms.courant_dt(0.02)
    o = orc.Oracle(v, a, b, rooms.C_AIR, 0.02, dt, False, sc, src, mc, "f32", 2, 6, 2, prof)
    o.run(steps)
    _assert_bitexact(got, o.receivers(), "traces")
    _assert_bitexact(f, o.field(0), "field")
    assert np.abs(got).max() > 0 and np.count_nonzero(f) > 10000  # the waves spread over many tiles


@pytest.mark.parametrize("precision", ["fp64", "fp16", "bf16"])
def test_two_steps_per_pass_every_precision(precision):
    _case(f"two vs one step per pass, {precision}, 141 steps, no graphs and graphs")
    v, a, b, sc, mc, prof = _room()
    steps = 141  # one 128-step graph of two-step passes, then 6 passes and a single step
    src = _src(steps)
    out = []
    for tblock, graph in ((1, True), (2, True), (2, False)):
        s = _sim(v, a, b, sc, mc, prof, src, tblock=tblock, precision=precision, graph=graph)
        s.run(steps)
        out.append((s.receivers(), [s.plane(2, 20, v.shape, which=w) for w in range(3)], s.time_block()[1]))
        s.close()
    (t1, p1, _), (t2, p2, n2), (t3, p3, n3) = out
    assert n2 > 0 and n3 > 0
    _assert_bitexact(t2, t1, "traces, graphs")
    _assert_bitexact(t3, t1, "traces, no graphs")
    for w in range(3):  # newest, previous and the one before: same roles
        _assert_bitexact(p2[w], p1[w], f"plane which={w}, graphs")
        _assert_bitexact(p3[w], p1[w], f"plane which={w}, no graphs")


def test_switching_mid_run_batched_with_autotune():
    """Batched rooms with the PML on every room: single steps, then two steps
    per pass from an odd step, then an autotune pass, then more steps, equal
    one straight single-step run."""
    _case("time block switch mid-run and autotune, batched rooms with PML")
    ws = rooms.batch_rooms(seed=12, nrooms=3, steps=150, lo=(140, 30, 40), hi=(160, 40, 50), nmic=3)

    def make():
        s = afx.Sim(c=ws[0].c, dx=ws[0].dx, dt=ws[0].dt, res=ws[0].dx, compat=False)
        s.set_voxels(ws[0].vox)
        s.set_materials(ws[0].a, ws[0].b)
        s.set_pml(4, 2 | 16)
        ids = [0] + [s.add_room(w.vox, w.a, w.b) for w in ws[1:]]
        for rid, w in zip(ids, ws):
            for cell, sig in zip(w.src_cells, w.src):
                s.add_source_cell(cell, sig, room=rid)
        for rid, w in zip(ids, ws):
            for cell in w.mic_cells:
                s.add_receiver_cell(cell, room=rid)
        return s

    s = make()
    s.run(150)
    want = s.receivers()
    s.close()
    s = make()
    s.run(37)
    s.set_time_block(2)
    s.run(41)
    assert s.time_block()[1] > 0
    tile = s.autotune(10)
    k, _ = s.time_block()
    print(f"autotuned: tile {tile}, {k} step(s) per pass")
    s.run(40)
    s.set_time_block(1)
    s.run(32)
    got = s.receivers()
    s.close()
    _assert_bitexact(got, want, "traces")


def test_time_block_argument_checks():
    _case("time block argument checks")
    v, a, b, sc, mc, prof = _room()
    s = _sim(v, a, b, sc, mc, prof, _src(4), tile=(256, 1, 16), tblock=2)
    s.run(4)
    assert s.time_block() == (2, 0)  # one-row tiles: single steps
    with pytest.raises(afx.AfxError):
        s.set_time_block(3)
    s.close()


@pytest.mark.parametrize("tile,tblock,split", [((64, 4, 16), 1, 1), ((64, 4, 16), 1, 2), ((64, 4, 16), 2, 2),
                                               ((128, 4, 8), 1, 2), ((64, 2, 32), 2, 1)])
def test_tile_split_matches_oracle(tile, tblock, split):
    """Tiles in the top quarter (split 1) or half (split 2) of the room's
    planes run as two work items of half the planes each; with one or two
    steps per pass and vector tiles the result stays the oracle's, bit for
    bit, including a switch of the split mid-run."""
    _case(f"tile split {split}, tile {tile}, steps per pass {tblock}, 60 + 61 steps")
    v, a, b, sc, mc, prof = _room()
    steps = 121
    src = _src(steps)
    s = _sim(v, a, b, sc, mc, prof, src, tile=tile, tblock=tblock)
    s.set_tile_split(split)
    s.run(60)
    assert s.tile_split() == split
    s.set_tile_split(0)
    s.run(30)
    s.set_tile_split(split)
    s.run(31)
    got, f = s.receivers(), s.pressure(0)
    s.close()
    dt = rooms.courant_dt(0.02)
    o = orc.Oracle(v, a, b, rooms.C_AIR, 0.02, dt, False, sc, src, mc, "f32", 2, 6, 2, prof)
    o.run(steps)
    _assert_bitexact(got, o.receivers(), "traces")
    _assert_bitexact(f, o.field(0), "field")
    assert np.abs(got).max() > 0


def test_tile_split_argument_checks():
    _case("tile split argument checks")
    v, a, b, sc, mc, prof = _room()
    s = _sim(v, a, b, sc, mc, prof, _src(4))
    assert s.tile_split() == 0
    for bad in (-1, 3):
        with pytest.raises(afx.AfxError):
            s.set_tile_split(bad)
    s.set_tile_split(2)
    s.run(4)
    assert s.tile_split() == 2
    s.close()


def test_profiled_steps_with_two_steps_per_pass():
    """afx_profile_run (events around every launch, no graphs) with two steps
    per pass: 33 steps (16 passes and one single step) continue the state
    exactly like afx_run, and every step counts two launches per pass."""
    _case("profile run with two steps per pass, 33 + 40 steps")
    v, a, b, sc, mc, prof = _room()
    steps = 73
    src = _src(steps)
    s = _sim(v, a, b, sc, mc, prof, src)
    s.run(steps)
    want = s.receivers()
    s.close()
    s = _sim(v, a, b, sc, mc, prof, src, tblock=2)
    pr = s.profile(33)
    assert pr.steps == 33 and pr.launches[0] == 33  # 16 passes of two launches, one single step
    s.run(40)
    got = s.receivers()
    s.close()
    _assert_bitexact(got, want, "traces")


@pytest.mark.parametrize("ncoef", [1, 3, 4])
def test_two_steps_per_pass_compat_mode_match_oracle(ncoef):
    """The reference's compat mode (lambda^2 at the boundary, the no-update
    corner group) with IIR filters of 1, 3 and 4 coefficients, no PML: two
    steps per pass against the oracle, bit for bit, over 131 steps (one
    128-step graph, a pass and a single step)."""
    _case(f"two steps per pass, compat mode, ncoef={ncoef}, 131 steps")
    rng = np.random.Generator(np.random.PCG64(50 + ncoef))
    nx, ny, nz = 264, 40, 70
    if ncoef == 3:
        a, b = rooms.materials(rng, 3)
    else:
        a = (rng.uniform(0.02, 0.3, (3, ncoef)) * np.array([1.0] + [0.3] * (ncoef - 1))).astype(np.float32)
        b = np.zeros((3, ncoef), np.float32)
        b[:, 0] = 1.0
        if ncoef > 1:
            b[:, 1:] = rng.uniform(-0.2, 0.2, (3, ncoef - 1))
    v = rooms.shoebox(nx, ny, nz, [int(rng.integers(0, 3)) for _ in range(6)])
    v = rooms.repair(rooms.add_boxes(v, rng, 3, 4, 9, 3))
    inner = rooms.interior_cells(v, rng, 3)
    sc = np.stack([inner[0], np.array([0, 5, 5], np.int32)])
    mc = np.stack([inner[1], inner[2], np.array([0, 5, 5], np.int32)])
    steps = 131
    dt, c = 3.125e-05, 340.0
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 800.0, amp=0.5)])
    s = afx.Sim(c=c, dx=0.02, dt=dt, res=0.02, compat=True)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_time_block(2)
    for cell, sig in zip(sc, src):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    s.run(steps)
    got, f, npair = s.receivers(), s.pressure(0), s.time_block()[1]
    s.close()
    assert npair > 0
    o = orc.Oracle(v, a, b, c, 0.02, dt, True, sc, src, mc, "f32", 0, 0, 0, None)
    o.run(steps)
    _assert_bitexact(got, o.receivers(), "traces")
    _assert_bitexact(f, o.field(0), "field")


@pytest.mark.parametrize("kw", [dict(fused=False), dict(list_blocks=True)])
def test_two_steps_per_pass_falls_back_where_it_cannot_apply(kw):
    """The split path and the list-blocks launch take single steps even with
    two steps per pass requested (no tile pairs), with unchanged results."""
    _case(f"two steps per pass requested with {kw}: single steps, same traces")
    v, a, b, sc, mc, prof = _room()
    steps = 40
    src = _src(steps)
    out = []
    for tblock in (1, 2):
        s = _sim(v, a, b, sc, mc, prof, src, tblock=tblock, **kw)
        s.run(steps)
        out.append((s.receivers(), s.time_block()))
        s.close()
    assert out[1][1] == (2, 0)
    _assert_bitexact(out[1][0], out[0][0], "traces")
