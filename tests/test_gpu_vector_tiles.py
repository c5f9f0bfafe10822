"""Interior tiles with several cells per lane (vector loads and stores along
x, neighbours across lanes by shuffles) against the oracle, bit for bit:
every vector shape in every storage precision, the split path (compat PML),
the list-blocks layout, rooms narrower and wider than a tile, batched rooms,
and the width rule (room widths must be multiples of the lane's cell count)."""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")

VEC_TILES = [(256, 4, 8), (256, 2, 8), (128, 4, 16), (128, 4, 8)]


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _room(seed, nx, ny, nz, open_faces=("x+",)):
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b = rooms.materials(rng, 3)
    v = rooms.shoebox(nx, ny, nz, [int(rng.integers(0, 3)) for _ in range(6)], open_faces=open_faces)
    v = rooms.repair(rooms.add_boxes(v, rng, 3, 3, 8, 3))
    return v, a, b, rng


def _run(v, a, b, sc, src, mc, steps, precision, compat, pml, tile, **kw):
    dx = 0.02
    dt = 3.125e-05 if compat else rooms.courant_dt(dx)
    c = 340.0 if compat else rooms.C_AIR
    s = afx.Sim(c=c, dx=dx, dt=dt, res=dx, precision=precision, compat=compat, **kw)
    s.set_voxels(v)
    s.set_materials(a, b)
    if pml is not None:
        s.set_pml(*pml)
    s.set_tile(*tile)
    for cell, sig in zip(sc, src):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    s.run(steps)
    st = s.stats()
    assert (st.tile_x, st.tile_y, st.tile_z) == tuple(tile)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    o = orc.Oracle(v, a, b, c, dx, dt, compat, sc, src, mc, "f64" if precision == "fp64" else "f32",
                   0 if pml is None else (1 if compat else 2), pml[0] if pml else 0, pml[1] if pml else 0,
                   pml[2] if pml else None, storage=precision if precision in ("fp16", "bf16") else None)
    o.run(steps)
    want, wfield = o.receivers(), o.field(0)
    o.close()
    assert np.abs(want).max() > 0
    assert np.array_equal(got, np.asarray(want).astype(got.dtype)), "traces differ"
    assert np.array_equal(field, np.asarray(wfield).astype(field.dtype)), "fields differ"


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp16", "bf16"])
@pytest.mark.parametrize("tile", VEC_TILES)
def test_vector_tile_matches_oracle(tile, precision):
    """A 72x40x36 room (narrower than the tile: lanes past the row only
    relay their neighbours), physical PML on x+ and y-, sources on an
    interior cell and an outer-plane cell, receivers inside the PML."""
    _case(f"vector tile {tile}, {precision}, 72x40x36 with PML x+ y-")
    v, a, b, rng = _room(61, 72, 40, 36, open_faces=("x+", "y-"))
    inner = rooms.interior_cells(v, rng, 3)
    steps = 110
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 800.0, amp=0.5)])
    sc = np.stack([inner[0], np.array([0, 20, 17], np.int32)])
    mc = np.concatenate([inner[1:], np.array([[69, 20, 15], [30, 2, 20], [1, 1, 1]], np.int32)])
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    _run(v, a, b, sc, src, mc, steps, precision, False, (6, 2 | 4, prof), tile)


@pytest.mark.parametrize("tile", [(256, 4, 8), (128, 4, 16)])
def test_vector_tile_wider_room_and_list_blocks(tile):
    """A 300x24x20 room: two tiles along x, the second one partly outside
    the row; fused with the lists in their own workgroups (AFX_LIST_BLOCKS)."""
    _case(f"vector tile {tile}, 300x24x20, list blocks")
    v, a, b, rng = _room(62, 300, 24, 20, open_faces=())
    inner = rooms.interior_cells(v, rng, 4)
    steps = 150
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt, 1500.0)])
    _run(v, a, b, inner[:1], src, inner[1:], steps, "fp32", False, None, tile, list_blocks=True)


@pytest.mark.parametrize("tile", [(256, 4, 8), (128, 4, 8)])
def test_vector_tile_split_path_compat_pml(tile):
    """Compat mode with the compat PML takes the split launches: k_interior
    with vector lanes, then the boundary kernel and pml.cl's three phases."""
    _case(f"vector tile {tile}, compat PML, split path")
    v, a, b, rng = _room(63, 64, 40, 32, open_faces=())
    inner = rooms.interior_cells(v, rng, 3)
    steps = 120
    src = np.stack([rooms.ricker(steps, 3.125e-05, 1000.0, amp=0.3)])
    prof = (8000.0 * np.linspace(0.05, 1.0, 8)).astype(np.float32)
    _run(v, a, b, inner[:1], src, inner[1:], steps, "fp32", True, (8, 0, prof), tile)


def test_vector_tiles_batched_rooms_match_single_runs():
    """Three rooms of widths 64, 132 and 260 (multiples of 4) in one handle
    with vector tiles: each room equals its own single-handle run with the
    default scalar tile, and the oracle."""
    _case("vector tile (256, 4, 8), 3 batched rooms")
    ws = []
    for k, (nx, ny, nz) in enumerate(((64, 30, 26), (132, 22, 28), (260, 18, 20))):
        v, a, b, rng = _room(70 + k, nx, ny, nz, open_faces=())
        cells = rooms.interior_cells(v, rng, 4)
        ws.append(rooms.Workload(f"r{k}", v, a, b, 0.02, rooms.courant_dt(0.02), cells[:1],
                                 rooms.ricker(130, rooms.courant_dt(0.02))[None, :], cells[1:], 130))
    from anc_field_py import batch
    s = batch.make_batch(ws)
    s.set_tile(256, 4, 8)
    s.run(130)
    got = s.receivers()
    s.close()
    seq, _ = batch.run_sequential(ws, 130)
    assert np.array_equal(got, seq)
    for k, w in enumerate(ws):
        o = orc.from_workload(w)
        o.run(130)
        assert np.array_equal(got[3 * k:3 * (k + 1)], o.receivers())
        o.close()


def test_vector_tile_needs_width_multiple():
    v, a, b, _ = _room(64, 70, 20, 20, open_faces=())
    s = afx.Sim(c=rooms.C_AIR, dx=0.02, dt=rooms.courant_dt(0.02), res=0.02, compat=False)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_tile(256, 4, 8)
    with pytest.raises(afx.AfxError) as ei:
        s.run(1)
    assert ei.value.code == afx.AFX_E_ARG and "multiples of 4" in str(ei.value)
    s.close()
    # after the first run the shape is checked at the call; 70 is a
    # multiple of 2, so the two-cell shapes are accepted
    s = afx.Sim(c=rooms.C_AIR, dx=0.02, dt=rooms.courant_dt(0.02), res=0.02, compat=False)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.run(2)
    with pytest.raises(afx.AfxError):
        s.set_tile(256, 2, 8)
    s.set_tile(128, 4, 8)
    s.run(2)
    st = s.stats()
    assert (st.tile_x, st.tile_y, st.tile_z) == (128, 4, 8)
    s.close()
