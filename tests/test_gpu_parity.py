"""HIP path against the CPU oracle, through the C ABI, on seeded inputs.

FP32 and FP64 are compared bit for bit (both sides evaluate the reference's
operations in the same order with IEEE arithmetic and no contraction).
"""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")


def _closed_room(seed, dims=(40, 36, 28), ncoef=3, nbox=3, nmat=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = dims
    if ncoef == 3:
        a, b = rooms.materials(rng, nmat)
    else:
        # frequency-dependent admittances of other orders: stable, passive-ish
        a = rng.uniform(0.02, 0.3, (nmat, ncoef)).astype(np.float32) * np.array([1.0] + [0.3] * (ncoef - 1), np.float32)
        b = np.zeros((nmat, ncoef), np.float32)
        b[:, 0] = 1.0
        if ncoef > 1:
            # sum |b_i| < b_0 keeps every root of B(z) inside the unit circle
            lim = 0.2 if ncoef <= 4 else 0.6 / (ncoef - 1)
            b[:, 1:] = rng.uniform(-lim, lim, (nmat, ncoef - 1))
    face_mats = [int(rng.integers(0, nmat)) for _ in range(6)]
    v = rooms.shoebox(nx, ny, nz, face_mats)
    v = rooms.repair(rooms.add_boxes(v, rng, nbox, 4, 9, nmat))
    return v, a, b, rng


def _special_cells(v, rng):
    """Source and receiver cells: interior, a wall face, the X_M_Y_M_Z_M
    corner (1,1,1), an edge cell and an outer-plane cell."""
    inner = rooms.interior_cells(v, rng, 3)
    pos, grp = orc.classify(v)
    face = pos[grp == 2][0, :3]
    edge = pos[grp == 11][0, :3]
    corner = np.array([1, 1, 1], np.int32)
    assert v[1, 1, 1] == -1 and (grp[(pos[:, 0] == 1) & (pos[:, 1] == 1) & (pos[:, 2] == 1)] == 25).all()
    plane = np.array([0, 5, 5], np.int32)
    src_cells = np.stack([inner[0], face, corner, inner[0], plane])  # two sources share inner[0]
    mic_cells = np.stack([inner[1], inner[2], face, edge, corner, inner[0], plane])
    return src_cells, mic_cells


def _run_pair(v, a, b, compat, src_cells, src, mic_cells, steps, precision="fp32", pml=None, lib_kw=None,
              split_runs=None):
    dx = 0.02
    dt = rooms.courant_dt(dx) if not compat else 3.125e-05
    c = rooms.C_AIR if not compat else 340.0
    lib_kw = dict(lib_kw or {})
    s = afx.Sim(c=c, dx=dx, dt=dt, res=dx, precision=precision, compat=compat, **lib_kw)
    s.set_voxels(v)
    s.set_materials(a, b)
    pml_mode = 0
    if pml is not None:
        width, faces, profile = pml
        s.set_pml(width, faces, profile)
        pml_mode = 1 if compat else 2
    for cell, sig in zip(src_cells, src):
        s.add_source_cell(cell, sig)
    for cell in mic_cells:
        s.add_receiver_cell(cell)
    for k in (split_runs or [steps]):
        s.run(k)
    got = s.receivers()
    field = s.pressure(0)
    o = orc.Oracle(v, a, b, c, dx, dt, compat, src_cells, src, mic_cells,
                   "f64" if precision == "fp64" else "f32", pml_mode,
                   pml[0] if pml else 0, pml[1] if pml else 0, pml[2] if pml else None)
    o.run(steps)
    want = o.receivers()
    wfield = o.field(0)
    s.close()
    return got, want, field, wfield


def _assert_bitexact(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want).astype(got.dtype)
    if not np.array_equal(got, want):
        d = np.abs(got.astype(np.float64) - want.astype(np.float64))
        idx = np.unravel_index(np.argmax(d), d.shape)
        rel = np.linalg.norm(d) / max(np.linalg.norm(want.astype(np.float64)), 1e-300)
        raise AssertionError(f"{what}: {np.count_nonzero(d)} values differ, max |d| {d.max():.3e} at {idx}, rel L2 {rel:.3e}")


def _case(name):
    print(f"\n[case] {name}", flush=True)


@pytest.mark.parametrize("compat", [True, False])
@pytest.mark.parametrize("ncoef", [1, 3, 4, 6, 8])  # 6 and 8: the ring slots take two batches of loads
def test_room_bitexact(compat, ncoef):
    _case(f"room_bitexact compat={compat} ncoef={ncoef}")
    v, a, b, rng = _closed_room(10 + ncoef, ncoef=ncoef)
    sc, mc = _special_cells(v, rng)
    steps = 300
    dt = 3.125e-05 if compat else rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt, 1000.0), rooms.ricker(steps, dt, 1500.0, amp=0.5),
                    rooms.ricker(steps, dt, 800.0, amp=0.7), rooms.ricker(steps, dt, 2000.0, amp=0.3),
                    rooms.ricker(steps, dt, 1200.0)])
    got, want, f, wf = _run_pair(v, a, b, compat, sc, src, mc, steps)
    assert np.abs(want).max() > 0
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


@pytest.mark.parametrize("flags", [dict(fused=False), dict(graph=False), dict(fused=False, graph=False),
                                   dict(list_blocks=True)])
def test_launch_variants_bitexact(flags):
    _case(f"launch variants {flags}")
    v, a, b, rng = _closed_room(21)
    sc, mc = _special_cells(v, rng)
    steps = 200
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt)] * len(sc))
    got, want, f, wf = _run_pair(v, a, b, False, sc, src, mc, steps, lib_kw=flags)
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


def test_resumable_runs():
    _case("resumable runs 37+50+13")
    v, a, b, rng = _closed_room(22)
    sc, mc = _special_cells(v, rng)
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(100, dt)] * len(sc))
    got, want, f, wf = _run_pair(v, a, b, False, sc, src, mc, 100, split_runs=[37, 50, 13])
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


@pytest.mark.parametrize("fused,list_blocks,precision", [(True, False, "fp32"), (True, True, "fp32"),
                                                          (False, False, "fp32"), (True, False, "fp64"),
                                                          (False, False, "fp64")])
def test_physical_pml_bitexact(fused, list_blocks, precision):
    """FP64 keeps the zeta tables in double, as the oracle's f64 build does."""
    _case(f"physical PML x+ and y- faces fused={fused} list_blocks={list_blocks} {precision}")
    v, a, b, rng = _closed_room(23, dims=(48, 40, 30))
    v[:, :, -1] = -1   # open x+ face
    v[:, 0, :] = -1    # open y- face
    v = rooms.repair(v)
    inner = rooms.interior_cells(v, rng, 4)
    # sources/receivers also inside the PML slabs and on a wall cell in a slab
    in_pml = np.array([[44, 20, 15], [20, 3, 12]], np.int32)
    wall_in_pml = np.array([44, 20, 1], np.int32)  # z-  wall face inside the x+ slab
    sc = np.stack([inner[0], in_pml[0], wall_in_pml])
    mc = np.concatenate([inner[1:], in_pml, wall_in_pml[None]])
    steps = 250
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 900.0, amp=0.5),
                    rooms.ricker(steps, dt, 1300.0, amp=0.3)])
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    got, want, f, wf = _run_pair(v, a, b, False, sc, src, mc, steps, precision=precision, pml=(6, 2 | 4, prof),
                                 lib_kw=dict(fused=fused, list_blocks=list_blocks))
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_physical_pml_three_face_corner(precision):
    """PML on x+, y+ and z+: one-face slabs (one non-zero zeta: the fused
    chunks keep no psi there), two-face edges and the three-face corner (psi
    read and stored, and s3 = zeta_x zeta_y zeta_z non-zero), with a source in
    the corner and receivers in each kind of region."""
    _case(f"physical PML x+ y+ z+ corner {precision}")
    v, a, b, rng = _closed_room(29, dims=(48, 40, 30))
    v[:, :, -1] = -1   # open x+ face
    v[:, -1, :] = -1   # open y+ face
    v[-1, :, :] = -1   # open z+ face
    v = rooms.repair(v)
    inner = rooms.interior_cells(v, rng, 2)
    corner = np.array([44, 36, 26], np.int32)    # all three zeta non-zero (width 6)
    edge = np.array([44, 36, 12], np.int32)      # x+ and y+
    slab = np.array([20, 20, 27], np.int32)      # z+ only
    sc = np.stack([inner[0], corner])
    mc = np.stack([inner[1], corner, edge, slab])
    steps = 300
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 900.0, amp=0.5)])
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    got, want, f, wf = _run_pair(v, a, b, False, sc, src, mc, steps, precision=precision, pml=(6, 2 | 8 | 32, prof))
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")
    assert np.all(np.abs(np.asarray(got)).max(axis=1) > 0)  # the wave reached every receiver


@pytest.mark.parametrize("profile", [None, "shipped"])
def test_compat_pml_bitexact(profile):
    _case(f"compat PML profile={profile}")
    v, a, b, rng = _closed_room(24, dims=(36, 32, 28))
    sc = rooms.interior_cells(v, rng, 1)
    mc = rooms.interior_cells(v, rng, 3)
    steps = 120
    src = rooms.ricker(steps, 3.125e-05)[None]
    prof = None
    if profile == "shipped":
        k = np.arange(16)
        prof = (8000 * (k / 16 - np.sin(2 * np.pi * k / 16) / (2 * np.pi)))[::-1].astype(np.float32)
    got, want, f, wf = _run_pair(v, a, b, True, sc, src, mc, steps, pml=(8, 0, prof))
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


def test_fp64_bitexact():
    _case("fp64 vs oracle f64")
    v, a, b, rng = _closed_room(25)
    sc, mc = _special_cells(v, rng)
    steps = 150
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt)] * len(sc))
    got, want, f, wf = _run_pair(v, a, b, False, sc, src, mc, steps, precision="fp64")
    # the library returns FP32 copies of FP64 state
    _assert_bitexact(got, want.astype(np.float32), "traces")
    _assert_bitexact(f, wf.astype(np.float32), "field")


def test_batched_rooms_match_single():
    _case("batched rooms vs one handle per room")
    ws = rooms.batch_rooms(seed=7, nrooms=3, steps=120, lo=(24, 20, 18), hi=(44, 38, 30), nmic=3)
    # single runs
    single = []
    for w in ws:
        s = rooms.make_sim(w, afx.Sim)
        s.run(w.steps)
        single.append(s.receivers())
        s.close()
    # one handle, all rooms
    s = afx.Sim(c=ws[0].c, dx=ws[0].dx, dt=ws[0].dt, res=ws[0].dx, compat=False)
    s.set_voxels(ws[0].vox)
    s.set_materials(ws[0].a, ws[0].b)
    ids = [0] + [s.add_room(w.vox, w.a, w.b) for w in ws[1:]]
    for rid, w in zip(ids, ws):
        for cell, sig in zip(w.src_cells, w.src):
            s.add_source_cell(cell, sig, room=rid)
    for rid, w in zip(ids, ws):
        for cell in w.mic_cells:
            s.add_receiver_cell(cell, room=rid)
    s.run(ws[0].steps)
    tr = s.receivers()
    s.close()
    k = 0
    for w, want in zip(ws, single):
        _assert_bitexact(tr[k:k + len(w.mic_cells)], want, w.name)
        k += len(w.mic_cells)


@pytest.mark.parametrize("precision,tblock,compat,split", [("fp32", 1, False, 0), ("fp32", 2, False, 0),
                                                          ("fp16", 1, False, 0), ("fp32", 2, True, 0),
                                                          ("fp16", 1, True, 0), ("fp32", 1, False, 2),
                                                          ("fp32", 2, True, 1)])
def test_batched_rooms_match_oracle(precision, tblock, compat, split):
    """Each room of one handle against its own oracle run, bit for bit:
    physical mode with a PML on x+ and z- of every room, or compat mode (the
    reference's constants; its PML is single-room only, so none); whole tiles
    or the rooms' top tiles split in halves."""
    _case(f"batched rooms vs the oracle per room, {precision}, {tblock} step(s) per pass, compat {compat}, "
          f"split {split}")
    ws = rooms.batch_rooms(seed=11, nrooms=3, steps=130, lo=(70, 24, 24), hi=(140, 40, 40), nmic=3)
    faces = afx.FACE_XP | afx.FACE_ZM
    width = 5
    prof = (3e4 * ((np.arange(width) + 0.5) / width) ** 3).astype(np.float32)
    c, dt = (340.0, 3.125e-05) if compat else (ws[0].c, ws[0].dt)
    s = afx.Sim(c=c, dx=ws[0].dx, dt=dt, res=ws[0].dx, compat=compat, precision=precision)
    s.set_voxels(ws[0].vox)
    s.set_materials(ws[0].a, ws[0].b)
    if not compat:
        s.set_pml(width, faces, prof)
    s.set_time_block(tblock)
    s.set_tile_split(split)
    ids = [0] + [s.add_room(w.vox, w.a, w.b) for w in ws[1:]]
    for rid, w in zip(ids, ws):
        for cell, sig in zip(w.src_cells, w.src):
            s.add_source_cell(cell, sig, room=rid)
    for rid, w in zip(ids, ws):
        for cell in w.mic_cells:
            s.add_receiver_cell(cell, room=rid)
    s.run(ws[0].steps)
    tr = s.receivers()
    s.close()
    k = 0
    for w in ws:
        o = orc.Oracle(w.vox, w.a, w.b, c, w.dx, dt, compat, w.src_cells, w.src, w.mic_cells, "f32",
                       0 if compat else 2, 0 if compat else width, 0 if compat else faces, None if compat else prof,
                       storage=precision if precision != "fp32" else None)
        want = o.run(w.steps)
        _assert_bitexact(tr[k:k + len(w.mic_cells)], want, w.name)
        assert np.abs(want).max() > 0
        k += len(w.mic_cells)


@pytest.mark.parametrize("fused", [True, False])
def test_batched_rooms_with_pml_match_single(fused):
    """Every room of a handle gets the PML; each room equals its own run, with
    sources and receivers inside the PML slabs too."""
    _case(f"batched rooms with physical PML vs one handle per room, fused={fused}")
    ws = rooms.batch_rooms(seed=8, nrooms=3, steps=150, lo=(26, 22, 20), hi=(40, 36, 30), nmic=3)
    faces = afx.FACE_XP | afx.FACE_YM
    width = 4
    extra = []  # one source and one receiver cell inside each room's x+ slab
    for w in ws:
        nz, ny, nx = w.vox.shape
        extra.append((np.array([nx - 3, ny // 2, nz // 2], np.int32), np.array([nx - 2, ny // 3, nz // 3], np.int32)))

    def setup(s, rid, w, ex):
        for cell, sig in zip(w.src_cells, w.src):
            s.add_source_cell(cell, sig, room=rid)
        s.add_source_cell(ex[0], 0.5 * w.src[0], room=rid)

    def mics(s, rid, w, ex):
        for cell in w.mic_cells:
            s.add_receiver_cell(cell, room=rid)
        s.add_receiver_cell(ex[1], room=rid)

    single = []
    for w, ex in zip(ws, extra):
        s = afx.Sim(c=w.c, dx=w.dx, dt=w.dt, res=w.dx, compat=False, fused=fused)
        s.set_voxels(w.vox)
        s.set_materials(w.a, w.b)
        s.set_pml(width, faces)
        setup(s, 0, w, ex)
        mics(s, 0, w, ex)
        s.run(w.steps)
        single.append(s.receivers())
        s.close()
    s = afx.Sim(c=ws[0].c, dx=ws[0].dx, dt=ws[0].dt, res=ws[0].dx, compat=False, fused=fused)
    s.set_voxels(ws[0].vox)
    s.set_materials(ws[0].a, ws[0].b)
    s.set_pml(width, faces)
    ids = [0] + [s.add_room(w.vox, w.a, w.b) for w in ws[1:]]
    for rid, w, ex in zip(ids, ws, extra):
        setup(s, rid, w, ex)
    for rid, w, ex in zip(ids, ws, extra):
        mics(s, rid, w, ex)
    s.run(ws[0].steps)
    tr = s.receivers()
    s.close()
    k = 0
    for w, want in zip(ws, single):
        assert np.abs(want[-1]).max() > 0  # the PML receiver sees the field
        _assert_bitexact(tr[k:k + len(want)], want, w.name)
        k += len(want)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_reduced_precision_error(precision):
    _case(f"reduced precision {precision}")
    v, a, b, rng = _closed_room(26)
    sc = rooms.interior_cells(v, rng, 1)
    mc = rooms.interior_cells(v, rng, 4)
    steps = 400
    dt = rooms.courant_dt(0.02)
    src = rooms.ricker(steps, dt)[None]
    got, _, _, _ = _run_pair(v, a, b, False, sc, src, mc, steps, precision=precision)
    ref, want64, _, _ = _run_pair(v, a, b, False, sc, src, mc, steps, precision="fp64")
    err = np.linalg.norm(got - want64) / np.linalg.norm(want64)
    # storage rounding (11-bit / 8-bit significands) accumulates over the run;
    # bounds record the measured error with margin (fp16 ~1e-2, bf16 ~1.2e-1)
    bound = 3e-2 if precision == "fp16" else 2.5e-1
    print(f"{precision} rel L2 vs fp64: {err:.3e}")
    assert err < bound


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_plane_snapshot_matches_field(precision):
    _case(f"plane snapshots, {precision}")
    v, a, b, rng = _closed_room(27, dims=(36, 30, 26))
    sc = rooms.interior_cells(v, rng, 1)
    steps = 60
    dt = rooms.courant_dt(0.02)
    s = afx.Sim(c=rooms.C_AIR, dx=0.02, dt=dt, res=0.02, compat=False, precision=precision)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.add_source_cell(sc[0], rooms.ricker(steps, dt))
    s.run(steps)
    dims = v.shape
    for which in (0, 1):
        f = s.pressure(which)
        assert np.array_equal(s.plane(1, 13, dims, which), f[:, 13, :])   # cross section, [z][x]
        assert np.array_equal(s.plane(0, 7, dims, which), f[:, :, 7])
        assert np.array_equal(s.plane(2, 11, dims, which), f[11])
    assert np.abs(s.plane(1, int(sc[0][1]), dims)).max() > 0
    with pytest.raises(afx.AfxError):
        s.plane(1, 30, dims)
    s.close()


def _workload_pair(w, steps, threads=16):
    s = rooms.make_sim(w, afx.Sim)
    s.run(steps)
    got = s.receivers()
    field = s.pressure(0)
    s.close()
    o = orc.from_workload(w, compat=False, threads=threads)
    o.run(steps)
    want, wfield = o.receivers(), o.field(0)
    o.close()
    return got, want, field, wfield


def test_shoebox_small_full_run_bitexact():
    """64x48x32 shoebox, 3 IIR materials, 8-cell PML on x+, Ricker at (10,12,8),
    4 receivers, all 2000 steps: identical to the oracle."""
    _case("shoebox_small 64x48x32, 2000 steps, physical PML")
    w = rooms.shoebox_small(seed=1, steps=2000)
    got, want, f, wf = _workload_pair(w, 2000)
    assert np.abs(want).max() > 0
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


def test_furnished_room_bitexact():
    """The benchmark room itself (256x192x128, 12 boxes, 4 IIR materials,
    12-cell PML on x+, 16 receivers) for 800 steps: identical to the oracle.
    The wave travels 0.57 cells per step, so after 800 steps it has crossed
    the whole room (the farthest cell is 244 cells from the source): the
    x+ PML slab, every box's boundary cells and every receiver carry it."""
    _case("furnished 256x192x128, 800 steps")
    steps = 800
    w = rooms.furnished(steps=steps)
    got, want, f, wf = _workload_pair(w, steps)
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")
    nz, ny, nx = w.vox.shape
    assert np.all(np.abs(want).max(axis=1) > 0), "a receiver was not reached"
    assert np.abs(wf[:, :, nx - 1 - w.pml_width:nx - 1]).max() > 0, "the PML slab was not reached"
    pos, grp = orc.classify(w.vox)
    bvals = wf[pos[:, 2], pos[:, 1], pos[:, 0]]
    assert np.count_nonzero(bvals) > 0.99 * len(bvals), "boundary cells not reached"


def test_large_room_linearity():
    """640x480x320 (98 M cells, two PML faces): the scheme is linear, and a
    power-of-two source scale is exact in floating point, so doubling every
    source doubles every trace bit for bit."""
    _case("large 640x480x320, 40 steps, source x1 and x2")
    w = rooms.large(steps=40)
    # receivers the wave reaches within 40 steps, next to the source
    near = w.src_cells[0] + np.array([[2, 0, 0], [0, 3, 0], [0, 0, -4]], np.int32)
    w.mic_cells = np.concatenate([w.mic_cells, near])
    s1 = rooms.make_sim(w, afx.Sim)
    s1.run(40)
    t1 = s1.receivers()
    s1.close()
    w.src = 2.0 * w.src
    s2 = rooms.make_sim(w, afx.Sim)
    s2.run(40)
    t2 = s2.receivers()
    s2.close()
    assert np.abs(t1).max() > 0
    assert np.array_equal(t2, 2.0 * t1)


TILES = [(64, 4, 16), (64, 4, 32), (64, 8, 16), (64, 2, 32), (128, 2, 16), (64, 4, 8), (256, 1, 16)]
# shapes with 4 or 2 cells per lane (room widths must be multiples of that)
VEC_TILES = [(256, 4, 8), (256, 2, 8), (128, 4, 16), (128, 4, 8)]


def _pml_room():
    v, a, b, rng = _closed_room(31, dims=(70, 40, 36))
    v[:, :, -1] = -1   # open x+ face
    v[:, 0, :] = -1    # open y- face
    v = rooms.repair(v)
    inner = rooms.interior_cells(v, rng, 3)
    sc = np.stack([inner[0], np.array([66, 20, 15], np.int32)])   # one source in the x+ slab
    mc = np.concatenate([inner[1:], np.array([[20, 3, 12], [66, 20, 1]], np.int32)])
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    return v, a, b, sc, mc, prof


@pytest.mark.parametrize("tile", TILES)
def test_tile_shapes_bitexact(tile):
    _case(f"tile shape {tile}, physical PML on x+ and y-")
    v, a, b, sc, mc, prof = _pml_room()
    steps = 120
    dt = rooms.courant_dt(0.02)
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 800.0, amp=0.5)])
    s = afx.Sim(c=rooms.C_AIR, dx=0.02, dt=dt, res=0.02, compat=False)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_pml(6, 2 | 4, prof)
    s.set_tile(*tile)
    for cell, sig in zip(sc, src):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    s.run(steps)
    got, f = s.receivers(), s.pressure(0)
    assert (s.stats().tile_x, s.stats().tile_y, s.stats().tile_z) == tile
    s.close()
    o = orc.Oracle(v, a, b, rooms.C_AIR, 0.02, dt, False, sc, src, mc, "f32", 2, 6, 2 | 4, prof)
    o.run(steps)
    _assert_bitexact(got, o.receivers(), "traces")
    _assert_bitexact(f, o.field(0), "field")


def test_tile_change_mid_run_and_autotune_keep_state():
    """A tile change between runs and an autotune pass (which steps copies of
    the state) leave the simulation exactly where one straight run is, for a
    batched handle with the PML on every room."""
    _case("set_tile mid-run and autotune, batched rooms with PML")
    ws = rooms.batch_rooms(seed=9, nrooms=3, steps=160, lo=(26, 22, 20), hi=(40, 36, 30), nmic=3)

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
    s.run(160)
    want = s.receivers()
    s.close()
    s = make()
    s.run(50)
    s.set_tile(128, 2, 16)
    s.run(50)
    tile = s.autotune(10)
    assert tile in TILES + VEC_TILES
    s.run(60)
    got = s.receivers()
    s.close()
    _assert_bitexact(got, want, "traces")
