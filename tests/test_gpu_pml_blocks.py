"""The physical PML stepped by z-marching PML blocks (AFX_PML_BLOCKS)
against the per-cell PML records (the fused kernel's default) and the oracle,
bit for bit: PML on several faces (boxes meeting at edges and corners),
boundary cells and furniture inside the layer, sources and receivers in it,
every storage precision, one- and several-cell-per-lane tiles, two steps
per pass and batched rooms."""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _room(seed, nx=72, ny=44, nz=40, faces=1 | 2 | 8 | 32):
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b = rooms.materials(rng, 3)
    names = {1: "x-", 2: "x+", 4: "y-", 8: "y+", 16: "z-", 32: "z+"}
    v = rooms.shoebox(nx, ny, nz, [int(rng.integers(0, 3)) for _ in range(6)],
                      open_faces=tuple(names[f] for f in names if faces & f))
    # a box reaching into the x+ layer and one in the z+ layer corner
    v[10:20, 15:25, nx - 9:nx - 3] = 1
    v[nz - 8:nz - 3, ny - 9:ny - 2, 20:30] = 2
    v = rooms.repair(rooms.add_boxes(v, rng, 2, 3, 6, 3))
    return v, a, b, rng


def _run(v, a, b, sc, src, mc, steps, precision, width, faces, prof, tile=None, tblock=1, **kw):
    dt = rooms.courant_dt(0.02)
    s = afx.Sim(c=rooms.C_AIR, dx=0.02, dt=dt, res=0.02, precision=precision, compat=False, **kw)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_pml(width, faces, prof)
    if tile:
        s.set_tile(*tile)
    s.set_time_block(tblock)
    for cell, sig in zip(sc, src):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    s.run(steps)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    return got, field


@pytest.mark.parametrize("precision,tile,tblock", [
    ("fp32", None, 1), ("fp64", None, 1), ("fp16", None, 1), ("bf16", (128, 4, 8), 1),
    ("fp32", (256, 2, 8), 1), ("fp32", (64, 4, 16), 2), ("fp32", (256, 1, 16), 1)])
def test_pml_blocks_match_records_and_oracle(precision, tile, tblock):
    _case(f"PML blocks, {precision}, tile {tile}, {tblock} step(s) per pass")
    faces = 1 | 2 | 8 | 32
    v, a, b, rng = _room(91, faces=faces)
    width = 6
    prof = (3e4 * ((np.arange(width) + 0.5) / width) ** 3).astype(np.float32)
    inner = rooms.interior_cells(v, rng, 3)
    steps = 140
    dt = rooms.courant_dt(0.02)
    # sources: interior, inside the x+ layer, in the x-/y+ edge of the layer
    sc = np.stack([inner[0], np.array([68, 30, 17], np.int32), np.array([3, 41, 12], np.int32)])
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 800.0, amp=0.5),
                    rooms.ricker(steps, dt, 1300.0, amp=0.7)])
    mc = np.concatenate([inner[1:], np.array([[68, 30, 17], [70, 42, 38], [2, 2, 37], [40, 40, 36]], np.int32)])
    got, field = _run(v, a, b, sc, src, mc, steps, precision, width, faces, prof, tile, tblock, pml_blocks=True)
    rec, rfield = _run(v, a, b, sc, src, mc, steps, precision, width, faces, prof, tile, tblock)
    assert np.abs(got).max() > 0
    assert np.array_equal(got, rec), "traces differ from the record path"
    assert np.array_equal(field, rfield), "fields differ from the record path"
    o = orc.Oracle(v, a, b, rooms.C_AIR, 0.02, dt, False, sc, src, mc, "f64" if precision == "fp64" else "f32", 2,
                   width, faces, prof, storage=precision if precision in ("fp16", "bf16") else None)
    o.run(steps)
    assert np.array_equal(got, np.asarray(o.receivers()).astype(got.dtype)), "traces differ from the oracle"
    assert np.array_equal(field, np.asarray(o.field(0)).astype(field.dtype)), "fields differ from the oracle"
    o.close()


def test_pml_blocks_batched_rooms_match_single_runs():
    _case("PML blocks, 3 batched rooms, PML on x+ and z-")
    ws = rooms.batch_rooms(seed=19, nrooms=3, steps=150, lo=(30, 24, 22), hi=(70, 40, 36), nmic=4)
    from anc_field_py import batch
    prof = (3e4 * ((np.arange(5) + 0.5) / 5) ** 3).astype(np.float32)

    def one(ws_, **kw):
        s = batch.make_batch(ws_, **kw)
        s.set_pml(5, 2 | 16, prof)
        s.run(150)
        tr = s.receivers()
        s.close()
        return tr

    got = one(ws, pml_blocks=True)
    assert np.array_equal(got, one(ws))
    for k, w in enumerate(ws):
        assert np.array_equal(got[4 * k:4 * (k + 1)], one([w], pml_blocks=True))
