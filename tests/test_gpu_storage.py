"""FP16 and BF16 field storage against the CPU oracle, bit for bit.

The library computes in FP32 and stores pressure as binary16 or bfloat16 in
units of pressure / scale (scale: the power of two at the sources' peak).
The oracle's f32 build with the same storage mode rounds every pressure
write to the storage type (leapfrog, boundary cells before their IIR rings
advance, both PMLs, and source cells once after all their sources are
added), so the two must agree exactly: traces, and the newest field.
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


def _room(seed, compat_pml=False):
    """136x30x36 shoebox with three boxes; open x+ face (physical PML)
    unless the compat PML is used (closed walls, the PML over the interior)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b = rooms.materials(rng, 3)
    opened = () if compat_pml else ("x+",)
    v = rooms.shoebox(136, 30, 36, [int(rng.integers(0, 3)) for _ in range(6)], open_faces=opened)
    v = rooms.repair(rooms.add_boxes(v, rng, 3, 3, 7, 3))
    return v, a, b, rng


def _run(v, a, b, sc, sigs, mc, steps, storage, compat=False, pml=None, tblock=1, **kw):
    dt = 3.125e-05 if compat else rooms.courant_dt(DX)
    c = 340.0 if compat else rooms.C_AIR
    s = afx.Sim(c=c, dx=DX, dt=dt, res=DX, precision=storage, compat=compat, **kw)
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
    s.run(steps)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    src = np.stack(sigs)
    o = orc.Oracle(v, a, b, c, DX, dt, compat, np.asarray(sc, np.int32), src, np.asarray(mc, np.int32), "f32",
                   pml_mode, pml[0] if pml else 0, pml[1] if pml else 0, pml[2] if pml else None,
                   storage=storage)
    o.run(steps)
    return got, o.receivers(), field, o.field(0), o.scale


@pytest.mark.parametrize("storage,tblock,fused", [("fp16", 1, True), ("bf16", 1, True), ("fp16", 2, True),
                                                  ("bf16", 2, True), ("fp16", 1, False)])
def test_physical_pml_storage_matches_oracle(storage, tblock, fused):
    _case(f"{storage} storage, physical PML x+, {tblock} step(s) per pass, fused={fused}")
    v, a, b, rng = _room(81)
    inner = rooms.interior_cells(v, rng, 3)
    edge = np.array([135, 12, 20], np.int32)
    in_pml = np.array([132, 14, 17], np.int32)
    steps = 161
    dt = rooms.courant_dt(DX)
    r1, r2 = rooms.ricker(steps, dt, 900.0, amp=3.0), rooms.ricker(steps, dt, 1400.0, amp=1.7)
    sc = [inner[0], inner[0], edge, in_pml]  # two sources on one cell: added, then stored once
    sigs = [r1, r2, r2, r1]
    mc = [inner[0], inner[1], inner[2], edge, in_pml]
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    got, want, f, wf, scale = _run(v, a, b, sc, sigs, mc, steps, storage, pml=(6, rooms.FACE_BITS["x+"], prof),
                                   tblock=tblock, fused=fused)
    print(f"scale {scale}, peak {np.abs(got).max():.3e}")
    assert scale == 4.0
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")
    assert np.abs(got).max() > 0


@pytest.mark.parametrize("storage,ncoef,with_pml", [("fp16", 1, False), ("fp16", 3, False), ("bf16", 3, False),
                                                    ("fp16", 3, True)])
def test_compat_storage_matches_oracle(storage, ncoef, with_pml):
    _case(f"{storage} storage, compat mode, ncoef={ncoef}, compat PML {with_pml}")
    v, a, b, rng = _room(82 + ncoef, compat_pml=with_pml)
    if ncoef == 1:
        a = (rng.uniform(0.02, 0.3, (3, 1))).astype(np.float32)
        b = np.ones((3, 1), np.float32)
    inner = rooms.interior_cells(v, rng, 3)
    steps = 140
    dt = 3.125e-05
    sigs = [rooms.ricker(steps, dt, 1000.0, amp=0.3)]
    prof = (8000.0 * np.linspace(0.05, 1.0, 8)).astype(np.float32) if with_pml else None
    got, want, f, wf, scale = _run(v, a, b, [inner[0]], sigs, [inner[1], inner[2], inner[0]], steps, storage,
                                   compat=True, pml=(8, 0, prof) if with_pml else None)
    assert scale == 0.5
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")


def test_fp16_long_tail_matches_oracle():
    """600 steps: the late field decays into binary16's subnormal range
    (|p / scale| < 2^-14), where the rounding quantum is fixed at 2^-24."""
    _case("fp16 storage, 600 steps, subnormal tails")
    v, a, b, rng = _room(85)
    inner = rooms.interior_cells(v, rng, 4)
    steps = 600
    dt = rooms.courant_dt(DX)
    sigs = [rooms.ricker(steps, dt, 1200.0, amp=1.0)]
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    got, want, f, wf, scale = _run(v, a, b, [inner[0]], sigs, inner[1:], steps, "fp16",
                                   pml=(6, rooms.FACE_BITS["x+"], prof))
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(f, wf, "field")
    sub = np.count_nonzero((np.abs(wf / scale) < 2.0**-14) & (wf != 0))
    print(f"cells in the subnormal range: {sub}")
    assert sub > 0


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_source_and_receiver_added_mid_run(storage):
    """A louder source, and a receiver on its cell, added after the first run: the storage
    scale stays the one the first run fixed (fields, rings and PML state are
    held in its units), the new source enters at its own sample index and the
    new receiver records from then on.  The oracle holds both sources from
    the start (the late one silent until then)."""
    _case(f"{storage}: source + receiver added after 100 of 220 steps")
    v, a, b, rng = _room(86)
    inner = rooms.interior_cells(v, rng, 5)
    steps, t1 = 220, 100
    dt = rooms.courant_dt(DX)
    first = rooms.ricker(steps, dt, 1000.0, amp=0.9)
    late = rooms.ricker(steps, dt, 1300.0, amp=6.0, t0=t1 * dt + 1.2 / 1300.0)
    late[:t1] = 0.0
    prof = (3e4 * ((np.arange(6) + 0.5) / 6) ** 3).astype(np.float32)
    s = afx.Sim(c=rooms.C_AIR, dx=DX, dt=dt, res=DX, precision=storage, compat=False)
    s.set_voxels(v)
    s.set_materials(a, b)
    s.set_pml(6, rooms.FACE_BITS["x+"], prof)
    s.add_source_cell(inner[0], first)
    s.add_receiver_cell(inner[1])
    s.add_receiver_cell(inner[2])
    s.run(t1)
    s.add_source_cell(inner[3], late)
    s.add_receiver_cell(inner[3])
    s.run(steps - t1)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    o = orc.Oracle(v, a, b, rooms.C_AIR, DX, dt, False, np.stack([inner[0], inner[3]]), np.stack([first, late]),
                   inner[[1, 2, 3]], "f32", 2, 6, rooms.FACE_BITS["x+"], prof,
                   storage=None if storage == "fp32" else storage,
                   scale=None if storage == "fp32" else orc.storage_scale(first))
    o.run(steps)
    want = o.receivers()
    want[2, :t1] = 0.0  # the late receiver records from its first step on
    # the late receiver sits on the late source, louder than the first scale
    assert np.abs(want[2, t1:]).max() > 4.0
    _assert_bitexact(got, want, "traces")
    _assert_bitexact(field, o.field(0), "field")
    o.close()
