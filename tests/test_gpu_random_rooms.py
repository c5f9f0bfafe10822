"""Seeded random configurations against the CPU oracle, bit for bit.

Each case draws a room (size, furniture boxes, wall materials with 1 to 4 IIR
coefficients), physical or compat mode, a PML (physical: random faces and
width; compat: the reference's PML or none), FP32 or FP64 (and each draw
again with FP16 or BF16 storage), a tile shape and one or two steps per pass.  Sources and receivers include interior cells and
a cell on the grid's edge plane.  The draws are fixed by the seed, so a
failure names a reproducible configuration.
"""
import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")

TILES = [(64, 4, 16), (64, 4, 32), (64, 8, 16), (64, 2, 32), (128, 2, 16), (64, 4, 8), (256, 1, 16)]


def _materials(rng, nmat, ncoef):
    if ncoef == 3:
        return rooms.materials(rng, nmat)
    a = (rng.uniform(0.02, 0.3, (nmat, ncoef)) * np.array([1.0] + [0.3] * (ncoef - 1))).astype(np.float32)
    b = np.zeros((nmat, ncoef), np.float32)
    b[:, 0] = 1.0
    if ncoef > 1:
        b[:, 1:] = rng.uniform(-0.2, 0.2, (nmat, ncoef - 1))
    return a, b


def _draw(seed):
    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    nx = int(rng.choice([72, 136, 200]))
    ny = int(rng.integers(24, 49))
    nz = int(rng.integers(24, 57))
    ncoef = int(rng.choice([1, 3, 4]))
    nmat = int(rng.integers(2, 5))
    a, b = _materials(rng, nmat, ncoef)
    v = rooms.shoebox(nx, ny, nz, [int(rng.integers(0, nmat)) for _ in range(6)])
    v = rooms.add_boxes(v, rng, int(rng.integers(0, 5)), 3, 8, nmat)
    compat = bool(rng.integers(0, 2))
    if compat:
        pml = (8, 0, None) if rng.integers(0, 2) else None
    else:
        faces = 0
        for bit in (1, 2, 4, 8, 16, 32):
            if rng.random() < 0.3:
                faces |= bit
        width = int(rng.integers(3, 8))
        if faces:
            for f, sl in ((1, (slice(None), slice(None), 0)), (2, (slice(None), slice(None), -1)),
                          (4, (slice(None), 0, slice(None))), (8, (slice(None), -1, slice(None))),
                          (16, (0, slice(None), slice(None))), (32, (-1, slice(None), slice(None)))):
                if faces & f:
                    v[sl] = -1  # an open face under the PML
        prof = (3e4 * ((np.arange(width) + 0.5) / width) ** 3).astype(np.float32)
        pml = (width, faces, prof) if faces else None
    v = rooms.repair(v)
    inner = rooms.interior_cells(v, rng, 4)
    plane = np.array([0, int(rng.integers(1, ny - 1)), int(rng.integers(1, nz - 1))], np.int32)
    sc = np.stack([inner[0], plane])
    mc = np.stack([inner[1], inner[2], inner[3], plane])
    precision = "fp64" if rng.random() < 0.3 else "fp32"
    tile = TILES[int(rng.integers(0, len(TILES)))]
    tblock = int(rng.integers(1, 3))
    steps = int(rng.integers(90, 200))
    return dict(v=v, a=a, b=b, compat=compat, pml=pml, sc=sc, mc=mc, precision=precision, tile=tile,
                tblock=tblock, steps=steps, ncoef=ncoef)


@pytest.mark.parametrize("seed", range(8))
def test_random_configuration_matches_oracle(seed):
    _check(_draw(seed), seed)


@pytest.mark.parametrize("seed", range(8))
def test_random_configuration_reduced_storage_matches_oracle(seed):
    """The same draws with FP16 (even seeds) or BF16 (odd seeds) storage."""
    d = _draw(seed)
    d["precision"] = "fp16" if seed % 2 == 0 else "bf16"
    _check(d, seed)


def _check(d, seed):
    v = d["v"]
    nz, ny, nx = v.shape
    pml = d["pml"]
    print(f"\n[case] random room {seed}: {nx}x{ny}x{nz}, ncoef {d['ncoef']}, compat {d['compat']}, "
          f"pml {None if pml is None else pml[:2]}, {d['precision']}, tile {d['tile']}, "
          f"{d['tblock']} step(s) per pass, {d['steps']} steps", flush=True)
    compat, steps = d["compat"], d["steps"]
    dx = 0.02
    dt = 3.125e-05 if compat else rooms.courant_dt(dx)
    c = 340.0 if compat else rooms.C_AIR
    src = np.stack([rooms.ricker(steps, dt), rooms.ricker(steps, dt, 800.0, amp=0.5)])
    s = afx.Sim(c=c, dx=dx, dt=dt, res=dx, precision=d["precision"], compat=compat)
    s.set_voxels(v)
    s.set_materials(d["a"], d["b"])
    pml_mode = 0
    if pml is not None:
        s.set_pml(*pml)
        pml_mode = 1 if compat else 2
    s.set_tile(*d["tile"])
    s.set_time_block(d["tblock"])
    for cell, sig in zip(d["sc"], src):
        s.add_source_cell(cell, sig)
    for cell in d["mc"]:
        s.add_receiver_cell(cell)
    s.run(steps)
    got, field = s.receivers(), s.pressure(0)
    s.close()
    o = orc.Oracle(v, d["a"], d["b"], c, dx, dt, compat, d["sc"], src, d["mc"],
                   "f64" if d["precision"] == "fp64" else "f32", pml_mode,
                   pml[0] if pml else 0, pml[1] if pml else 0, pml[2] if pml else None,
                   storage=d["precision"] if d["precision"] in ("fp16", "bf16") else None)
    o.run(steps)
    want, wfield = o.receivers(), o.field(0)
    assert np.array_equal(got, np.asarray(want).astype(got.dtype)), "traces differ"
    assert np.array_equal(field, np.asarray(wfield).astype(field.dtype)), "fields differ"
    assert np.abs(got).max() > 0
