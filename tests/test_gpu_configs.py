"""Every workload configuration at its full size on the GPU.

  - 48 uneven rooms (64^3 .. 256x192x128, 88 M cells) in one handle: each
    room bit-identical to its own single-handle run, with the default and the
    autotuned tile / steps per pass, and three rooms against the oracle;
  - the 640x480x320 room with FP16 and BF16 storage: bit-identical to the
    oracle's storage mode, and linear in the source;
  - the surrogate workflow in the 256x192x128 room: batched impulse
    responses equal single-room runs, the network answers the pairs it was
    validated on and FDTD answers the rest;
  - the compat PML with the intended kappa profile over more than 8000
    steps, so the reference's periodic phi/psi reset (simulation.hpp:411-430)
    is exercised, in one run call and split across two;
  - compat classification restricted to the voxel volume when minsize grows
    the grid (lrs-data.hpp:101-105,154-158).
"""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")


def _case(name):
    print(f"\n[case] {name}", flush=True)


def _assert_bitexact(got, want, what):
    got = np.asarray(got)
    want = np.asarray(want).astype(got.dtype)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    if not np.array_equal(got, want):
        d = np.abs(got.astype(np.float64) - want.astype(np.float64))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} values differ, max |d| {d.max():.3e}")


# ------------------------------------------------------------------ 48 rooms
@pytest.fixture(scope="module")
def batch48():
    _case("48-room batch: generating rooms")
    return rooms.batch_rooms()


def test_batch_48_rooms_full_size(batch48):
    from anc_field_py import batch
    ws = batch48
    steps = 240
    cells = sum(w.cells for w in ws)
    assert len(ws) == 48 and cells > 80_000_000
    _case(f"48-room batch, {cells} cells, {steps} steps, default tile")
    s = batch.make_batch(ws)
    s.run(steps)
    got = s.receivers()
    s.close()
    _case("48-room batch, autotuned")
    s = batch.make_batch(ws)
    tile = s.autotune(20)
    sp = s.time_block()[0]
    s.run(steps)
    tuned = s.receivers()
    s.close()
    print(f"autotuned tile {tile}, steps per pass {sp}", flush=True)
    _assert_bitexact(tuned, got, "autotuned batch vs default batch")
    _case("48 rooms one handle each")
    seq, _ = batch.run_sequential(ws, steps)
    _assert_bitexact(got, seq, "batch vs sequential")
    _case("48 rooms in groups that fit the Infinity Cache")
    grouped, _, groups = batch.run_grouped(ws, steps)
    assert len(groups) > 1 and sorted(i for g in groups for i in g) == list(range(48))
    _assert_bitexact(grouped, seq, "cache groups vs sequential")
    # the impulse reaches every receiver of most rooms within the run
    reached = np.abs(got).max(axis=1) > 0
    assert reached.mean() > 0.5
    # three rooms against the oracle: the smallest, the largest and one between
    order = np.argsort([w.cells for w in ws])
    nmic = len(ws[0].mic_cells)
    for k in (order[0], order[len(order) // 2], order[-1]):
        w = ws[k]
        _case(f"batch room {k} {w.dims} against the oracle")
        o = orc.from_workload(w, compat=False, threads=16)
        o.run(steps)
        _assert_bitexact(got[k * nmic:(k + 1) * nmic], o.receivers(), f"room {k}")
        o.close()


# ------------------------------------------------------------------ large room
@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_large_room_reduced_storage_bitexact(storage):
    """640x480x320, 6 materials, 16-cell PML on x+ and y+ (98 M cells),
    40 steps: traces and the newest field equal the oracle's storage mode."""
    steps = 40
    _case(f"large 640x480x320, {storage}, {steps} steps against the oracle")
    w = rooms.large(steps=steps)
    near = w.src_cells[0] + np.array([[2, 0, 0], [0, 3, 0], [0, 0, -4], [-5, 1, 1]], np.int32)
    w.mic_cells = np.concatenate([w.mic_cells, near])
    s = rooms.make_sim(w, afx.Sim, precision=storage)
    s.run(steps)
    got = s.receivers()
    field = s.pressure(0)
    s.close()
    o = orc.Oracle(w.vox, w.a, w.b, w.c, w.dx, w.dt, False, w.src_cells, w.src, w.mic_cells, "f32", 2,
                   w.pml_width, w.pml_faces, w.pml_profile, threads=16, storage=storage)
    o.run(steps)
    want, wfield = o.receivers(), o.field(0)
    o.close()
    assert np.abs(want[-4:]).max() > 0
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(field, wfield, "field")


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_large_room_reduced_storage_linearity(storage):
    """Halving the source halves the power-of-two storage scale, so the
    stored values and every trace scale exactly (60 steps, full size)."""
    steps = 60
    _case(f"large 640x480x320, {storage}, {steps} steps, source x1 and x0.5")
    w = rooms.large(steps=steps)
    near = w.src_cells[0] + np.array([[2, 0, 0], [0, 3, 0], [0, 0, -4]], np.int32)
    w.mic_cells = np.concatenate([w.mic_cells, near])
    out = []
    for f in (1.0, 0.5):
        w2 = rooms.large(steps=steps)
        w2.mic_cells = w.mic_cells
        w2.src = f * w.src
        s = rooms.make_sim(w2, afx.Sim, precision=storage)
        s.run(steps)
        out.append(s.receivers())
        s.close()
    assert np.abs(out[0][-3:]).max() > 0
    assert np.array_equal(out[1], 0.5 * out[0])


# ------------------------------------------------------------------ surrogate
def test_surrogate_in_the_furnished_room():
    """The surrogate workflow in the 256x192x128 room with 1024 taps (4096
    steps): batched responses of 4 sources x 16 receivers equal single-room
    runs bit for bit; a network fitted to 3 of the sources answers those
    validated pairs, and the 4th source falls back to FDTD."""
    from anc_field_py import surrogate
    _case("surrogate, furnished 256x192x128, 4 sources x 16 receivers, 4096 steps")
    w = rooms.furnished(steps=16)
    rng = np.random.Generator(np.random.PCG64(17))
    src = surrogate.free_cells(w, rng, 4)
    rcv = [surrogate.free_cells(w, rng, 16) for _ in range(4)]
    sur = surrogate.Surrogate(w, taps=1024, decim=4, rooms_per_handle=4, n_freq=16, width=128, depth=4,
                              batch=32, seed=3, freq_min=0.25, freq_max=8.0)
    try:
        ir = sur.simulate(src, rcv)
        assert ir.shape == (64, 1024)
        _case("surrogate: single-room runs")
        for i in (0, 3):
            s = afx.Sim(c=w.c, dx=w.dx, dt=w.dt, res=w.dx, compat=False)
            s.set_voxels(w.vox)
            s.set_materials(w.a, w.b)
            s.set_pml(w.pml_width, w.pml_faces, w.pml_profile)
            s.add_source_cell(src[i], np.array([1.0], np.float32))
            for c in rcv[i]:
                s.add_receiver_cell(c)
            s.run(4096)
            want = surrogate.lowpass_decimate(s.receivers(), w.dt, 1024, 4, 2000.0)
            s.close()
            _assert_bitexact(ir[16 * i:16 * (i + 1)], want, f"source {i}")
            assert np.all(np.linalg.norm(want, axis=1) > 0)
        _case("surrogate: fit 48 pairs, validate, query")
        ps, pr = surrogate.pairs(src[:3], rcv[:3])
        sur.fit(sur.encode(ps, pr), ir[:48], 3000, 2e-3, 2e-5, seed=1)
        err = sur.validate(ps, pr, ir[:48])
        ok = err <= sur.bound
        print(f"fitted pairs: median rel L2 {np.median(err):.3e}, {ok.sum()} of 48 within the bound", flush=True)
        assert ok.sum() >= 12
        runs = sur.fdtd_runs
        good = np.flatnonzero(ok)
        qs = np.concatenate([ps[good], np.repeat(src[3:4], 16, axis=0)])
        qr = np.concatenate([pr[good], rcv[3]])
        truth = np.concatenate([ir[good], ir[48:]])
        out, fb, _ = sur.query(qs, qr)
        assert not fb[:len(good)].any() and fb[len(good):].all()
        assert sur.fdtd_runs - runs == 1
        assert np.array_equal(out[fb], truth[fb])
        assert np.all(surrogate.rel_l2(out, truth) <= sur.bound)
    finally:
        sur.close()


# ------------------------------------------------------------------ compat PML reset
def _shipped_profile():
    """ksi_profile.dat: 8000 (k/16 - sin(2 pi k/16) / 2 pi), k = 0..15."""
    k = np.arange(16)
    return (8000 * (k / 16 - np.sin(2 * np.pi * k / 16) / (2 * np.pi))).astype(np.float32)


@pytest.mark.parametrize("split", [None, (5000, 3300)])
def test_compat_pml_over_8000_steps(split):
    """pml.cl with the intended profile on a 40x36x32 room for 8300 steps:
    the reference zeroes phi and psi after step i % 8000 == 0 of each run
    call, so steps 0 and 8000 of one call (or 0 of each of two calls) reset
    them.  Traces and field equal the oracle bit for bit."""
    steps = 8300
    _case(f"compat PML, 40x36x32, {steps} steps, runs {split}")
    rng = np.random.Generator(np.random.PCG64(41))
    a, b = rooms.materials(rng, 2)
    v = rooms.shoebox(40, 36, 32, [0, 1, 0, 1, 0, 1])
    v = rooms.repair(rooms.add_boxes(v, rng, 2, 3, 6, 2))
    dt = 3.125e-05
    sig = (rooms.band_noise(rng, steps, dt, 4000.0, centre=0.5, width=0.35)).astype(np.float32)
    sc = rooms.interior_cells(v, rng, 1)
    mc = np.concatenate([rooms.interior_cells(v, rng, 3), np.array([[3, 3, 3], [20, 18, 2]], np.int32)])
    prof = _shipped_profile()
    s = afx.Sim(c=340.0, dx=0.02, dt=dt, res=0.02, compat=True)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_pml(len(prof), 0, prof)
    s.add_source_cell(sc[0], sig)
    for c in mc:
        s.add_receiver_cell(c)
    o = orc.Oracle(v, a, b, 340.0, 0.02, dt, True, sc, sig[None], mc, "f32", 1, len(prof), 0, prof)
    for k in (split or (steps,)):
        s.run(k)
        o.run(k)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    want, wfield = o.receivers(), o.field(0)
    o.close()
    assert np.all(np.isfinite(want)) and np.abs(want[:, 8000:]).max() > 0
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(field, wfield, "field")


# ------------------------------------------------------------------ minsize scan
def test_compat_scan_stays_in_the_voxel_volume():
    """A 40x36x30 voxel volume (walls on every face, furniture touching the
    last planes) in a grid that minsize grows to 64x64x32.  Compat mode
    classifies only 1 <= x <= vx-2 (likewise y, z), so cells on and beyond
    the volume's last planes run the plain leapfrog; physical mode scans the
    whole grid.  Both equal the oracle bit for bit."""
    rng = np.random.Generator(np.random.PCG64(43))
    a, b = rooms.materials(rng, 3)
    vx, vy, vz = 40, 36, 30
    v = rooms.shoebox(vx, vy, vz, [0, 1, 2, 0, 1, 2])
    v[5:12, 10:20, vx - 6:vx - 1] = 1    # box against the x+ wall
    v[3:9, vy - 5:vy - 1, 8:16] = 2      # box against the y+ wall
    v = rooms.repair(v)
    res = 0.02
    # minsize in metres: (int)(x/res) + 1 rounded up to 32 (simulation.hpp:279-284)
    dims = []
    for m in (1.0, 1.0, 0.5):
        n = int(np.float32(m) / np.float32(res)) + 1
        dims.append(-(-n // 32) * 32)
    nx, ny, nz = dims
    assert (nx, ny, nz) == (64, 64, 32)
    grid = np.full((nz, ny, nx), -1, np.int32)
    grid[:vz, :vy, :vx] = v
    steps = 300
    for compat in (True, False):
        _case(f"minsize-grown grid 64x64x32 around a 40x36x30 volume, compat={compat}")
        dt = 3.125e-05 if compat else rooms.courant_dt(res)
        c = 340.0 if compat else rooms.C_AIR
        sig = rooms.ricker(steps, dt, 1200.0)
        sc = np.array([[20, 18, 15]], np.int32)
        mc = np.array([[30, 20, 10], [38, 20, 15], [45, 20, 15], [20, 40, 12], [50, 50, 20]], np.int32)
        s = afx.Sim(c=c, dx=res, dt=dt, res=res, compat=compat)
        s.set_voxels(v)
        s.set_materials(a, b)
        s.set_min_size(1.0, 1.0, 0.5)
        s.add_source_cell(sc[0], sig)
        for cell in mc:
            s.add_receiver_cell(cell)
        s.run(steps)
        st = s.stats()
        got = s.receivers()
        s.close()
        assert (st.nx, st.ny, st.nz) == (nx, ny, nz)
        vol = (vx, vy, vz) if compat else None
        pos, grp = orc.classify(grid, vol=vol)
        assert np.array_equal(np.array(st.group_count), np.bincount(grp, minlength=26))
        o = orc.Oracle(grid, a, b, c, res, dt, compat, sc, sig[None], mc, vol=vol)
        o.run(steps)
        want = o.receivers()
        o.close()
        if compat:
            # the reference steps solid voxels, and the cells past the volume
            # read them as plain leapfrog cells: the wave leaks out
            assert np.abs(want[[2, 3, 4]]).max(axis=1).min() > 0
        else:
            # the walls' outer faces are boundary cells too: nothing crosses
            assert np.abs(want[[0, 1]]).max(axis=1).min() > 0 and not want[[2, 3, 4]].any()
        _assert_bitexact(got, want, "traces")
    # the two scans differ on this grid: the physical one finds the cells outside the volume
    assert len(orc.classify(grid)[1]) > len(orc.classify(grid, vol=(vx, vy, vz))[1])
