"""Pins of the CPU oracle that the mathematics fixes.

Each test is chosen so a plausible mistake in the restated update (a dropped
term, a wrong sign, a swapped neighbour or material, lambda vs lambda^2)
fails it:
  - boundary classification against brute-force pattern enumeration;
  - normal-incidence reflection of a locally reacting wall, R=(1-b)/(1+b);
  - free-field arrival time and 1/r spreading of a point source;
  - linearity (exact doubling) and superposition;
  - reciprocity between source and receiver;
  - the compat PML at kappa == 0 collapsing to the leapfrog;
  - the physical PML absorbing what a pressure-release face reflects;
  - FP32 against FP64.
"""
import itertools

import numpy as np
import pytest

from anc_field_py import _afx, rooms
import orc

C0, DX = 343.0, 0.02
DT = rooms.courant_dt(DX)


# ---------------------------------------------------------------- classifier
def test_classifier_all_patterns():
    """Every solid-neighbour pattern of one air cell maps to the group the
    reference's decimal pattern table gives, or is rejected."""
    face_names = ["X_P", "X_M", "Y_P", "Y_M", "Z_P", "Z_M"]
    groups = (["X_P", "X_M", "Y_P", "Y_M", "Z_P", "Z_M"]
              + ["X_P_Y_P", "X_P_Y_M", "X_P_Z_P", "X_P_Z_M", "X_M_Y_P", "X_M_Y_M", "X_M_Z_P", "X_M_Z_M",
                 "Y_P_Z_P", "Y_P_Z_M", "Y_M_Z_P", "Y_M_Z_M"]
              + ["X_P_Y_P_Z_P", "X_P_Y_P_Z_M", "X_P_Y_M_Z_P", "X_P_Y_M_Z_M", "X_M_Y_P_Z_P", "X_M_Y_P_Z_M",
                 "X_M_Y_M_Z_P", "X_M_Y_M_Z_M"])
    offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for mask in range(1, 64):
        v = np.full((5, 5, 5), -1, np.int32)
        for f in range(6):
            if mask >> f & 1:
                dx, dy, dz = offs[f]
                v[2 + dz, 2 + dy, 2 + dx] = f  # material id = face index
        opposite = any((mask >> (2 * a) & 1) and (mask >> (2 * a + 1) & 1) for a in range(3))
        name = "_".join(face_names[f] for f in range(6) if mask >> f & 1)
        if opposite:
            with pytest.raises(ValueError):
                orc.classify(v)
            with pytest.raises(_afx.AfxError):
                _afx.classify_counts(v)
            continue
        pos, grp = orc.classify(v)
        sel = (pos[:, 0] == 2) & (pos[:, 1] == 2) & (pos[:, 2] == 2)
        assert sel.sum() == 1
        assert groups[grp[sel][0]] == name
        mats = [m for m in pos[sel][0, 3:] if m >= 0]
        assert mats == [f for f in range(6) if mask >> f & 1]
        lib_counts = _afx.classify_counts(v)
        assert np.array_equal(lib_counts, np.bincount(grp, minlength=26))


def test_classifier_random_grids_library_matches_oracle():
    rng = np.random.Generator(np.random.PCG64(5))
    for k in range(20):
        v = rooms.shoebox(20, 18, 16, [int(rng.integers(0, 3)) for _ in range(6)])
        v = rooms.repair(rooms.add_boxes(v, rng, 6, 2, 6, 3))
        pos, grp = orc.classify(v)
        assert np.array_equal(_afx.classify_counts(v), np.bincount(grp, minlength=26))


# ---------------------------------------------------------------- boundaries
def _duct(beta, compat=False, n=420, f0=250.0):
    """Plane wave in a 4x4-cell duct with near-rigid side walls; wall of
    admittance beta at x = n-1.  Returns (incident, reflected) peak amplitude."""
    v = np.full((4, 4, n), -1, np.int32)
    v[0, :, :] = v[3, :, :] = 1
    v[:, 0, :] = v[:, 3, :] = 1
    v[:, :, n - 1] = 0
    a = np.array([[beta], [1e-7]], np.float32)
    b = np.array([[1.0], [1.0]], np.float32)
    steps = 1400
    sig = rooms.ricker(steps, DT, f0, t0=2.0 / f0)
    src = [[100, y, z] for y in (1, 2) for z in (1, 2)]
    o = orc.Oracle(v, a, b, C0, DX, DT, compat, src, np.stack([sig] * 4), [[250, 1, 1]], precision="f64")
    tr = o.run(steps)[0]
    # direct pulse passes x=250 around (150 cells)/(c dt/dx) steps, the echo from x=n-2 later
    speed = C0 * DT / DX
    t_dir = int(150 / speed + 2.0 / f0 / DT)
    t_ref = int((150 + 2 * (n - 2 - 250)) / speed + 2.0 / f0 / DT)
    w = int(1.2 / f0 / DT)
    inc = np.abs(tr[t_dir - w:t_dir + w]).max()
    ref = tr[t_ref - w:t_ref + w]
    return inc, ref[np.argmax(np.abs(ref))]


@pytest.mark.parametrize("beta", [0.1, 0.5, 1.0])
def test_normal_incidence_reflection(beta):
    inc, ref = _duct(beta)
    want = (1 - beta) / (1 + beta)
    assert abs(ref / inc - want) < 0.03, (ref / inc, want)


def test_compat_boundary_differs_from_physical():
    """Compat feeds lambda^2 where the boundary update expects lambda
    (simulation.hpp:368-369): the echo differs from the physical one."""
    inc_p, ref_p = _duct(0.5)
    inc_c, ref_c = _duct(0.5, compat=True)
    assert abs(ref_p / inc_p - ref_c / inc_c) > 0.05


# ---------------------------------------------------------------- free field
def test_free_field_arrival_and_spreading():
    n = 130
    steps = 190
    f0 = 800.0
    sig = rooms.ricker(steps, DT, f0)
    c0 = [40, 65, 65]
    mics = [[40 + r, 65, 65] for r in (15, 30, 45)]
    o = orc.Oracle(None, None, None, C0, DX, DT, False, [c0], sig[None], mics, dims=(n, n, n), precision="f64")
    tr = o.run(steps)
    peaks = np.argmax(np.abs(tr), axis=1)
    amps = np.abs(tr).max(axis=1)
    # arrival: peak delay grows by r/c (within 3 samples of dispersion)
    d = np.diff(peaks)
    assert np.all(np.abs(d - 15 * DX / (C0 * DT)) < 3), d
    # spreading: amplitude ~ 1/r
    ratio = amps[0] / amps[1:]
    assert np.allclose(ratio, [2.0, 3.0], rtol=0.08), ratio


def test_linearity_and_superposition():
    v, a, b = _small_room()
    sig = rooms.ricker(200, DT)
    sig2 = rooms.ricker(200, DT, 700.0)
    mics = [[6, 7, 8], [14, 12, 9]]

    def run(srcs, sigs, precision="f32"):
        o = orc.Oracle(v, a, b, C0, DX, DT, False, srcs, np.stack(sigs), mics, precision=precision)
        return o.run(200)

    one = run([[5, 5, 5]], [sig])
    two = run([[5, 5, 5]], [2 * sig])
    assert np.array_equal(two, 2 * one)  # scaling by 2 is exact in IEEE arithmetic
    one = run([[5, 5, 5]], [sig], "f64")
    other = run([[12, 10, 7]], [sig2], "f64")
    both = run([[5, 5, 5], [12, 10, 7]], [sig, sig2], "f64")
    assert np.allclose(both, one + other, rtol=0, atol=1e-12 * np.abs(both).max())


def test_reciprocity():
    v, a, b = _small_room()
    sig = rooms.ricker(250, DT)
    p, q = [5, 6, 7], [15, 12, 9]
    ab = orc.Oracle(v, a, b, C0, DX, DT, False, [p], sig[None], [q], precision="f64").run(250)
    ba = orc.Oracle(v, a, b, C0, DX, DT, False, [q], sig[None], [p], precision="f64").run(250)
    assert np.allclose(ab, ba, rtol=0, atol=1e-9 * np.abs(ab).max())


def _small_room():
    rng = np.random.Generator(np.random.PCG64(3))
    a, b = rooms.materials(rng, 2)
    v = rooms.shoebox(22, 20, 18, [0, 1, 0, 1, 0, 1])
    return v, a, b


# ---------------------------------------------------------------- PML
def test_compat_pml_zero_profile_is_leapfrog():
    """With kappa == 0 (the shipped state) pml.cl's pressure update is the
    leapfrog written over dt^2; it must agree with the plain leapfrog."""
    dims = (30, 26, 24)
    sig = rooms.ricker(120, 3.125e-05)
    src, mics = [[12, 11, 10]], [[5, 5, 5], [20, 15, 12]]
    plain = orc.Oracle(None, None, None, 340.0, 0.02, 3.125e-05, True, src, sig[None], mics, dims=dims)
    pml = orc.Oracle(None, None, None, 340.0, 0.02, 3.125e-05, True, src, sig[None], mics, dims=dims,
                     pml_mode=1, pml_width=8)
    x, y = plain.run(120), pml.run(120)
    assert np.abs(x).max() > 0
    assert np.allclose(x, y, rtol=0, atol=2e-5 * np.abs(x).max())
    assert not np.array_equal(x, y)  # computed by a different formula


def test_compat_pml_absorbs_with_shipped_profile():
    """pml.cl with the intended kappa profile (ksi_profile.dat, 8000 (k/16 -
    sin(2 pi k/16) / 2 pi), reversed at both ends by SetPml) in a 64^3 box:

    - the direct sound (steps 0-80) carries the same energy as without the
      layer, to 1e-3 (kappa and phi are zero away from the edges);
    - the echo from the layer's inner face (steps 80-160) stays below 1e-2 of
      the direct sound's energy: the profile rises from zero there.  Left
      unreversed, it starts at its maximum and the jump reflects 4.6e-2;
    - the late field (steps 450-700) holds less than 1e-4 of the energy of
      the same box with pressure-release edges.  A flipped sign of the
      kappa/(2 dt) pm2 term gives 1.4, kappa missing from the denominator
      blows up.

    The corner term (kappa_x kappa_y kappa_z psi) and the reference's phi
    update quirks are not distinguished by these properties."""
    n, steps = 64, 700
    dt = 3.125e-05
    sig = rooms.ricker(steps, dt, 1000.0)
    src, mics = [[32, 32, 32]], [[36, 32, 32], [32, 40, 30]]
    k = np.arange(16)
    shipped = (8000 * (k / 16 - np.sin(2 * np.pi * k / 16) / (2 * np.pi))).astype(np.float32)
    bare = orc.Oracle(None, None, None, 340.0, 0.02, dt, True, src, sig[None], mics, precision="f64",
                      dims=(n, n, n)).run(steps)
    lay = orc.Oracle(None, None, None, 340.0, 0.02, dt, True, src, sig[None], mics, precision="f64",
                     dims=(n, n, n), pml_mode=1, pml_width=16, pml_profile=shipped).run(steps)
    assert np.all(np.isfinite(lay))
    e = lambda t, lo, hi: float(np.sum(t[:, lo:hi] ** 2))
    direct = e(bare, 0, 80)
    assert abs(e(lay, 0, 80) / direct - 1.0) < 1e-3
    assert e(lay, 80, 160) < 1e-2 * direct, e(lay, 80, 160) / direct
    assert e(lay, 450, 700) < 1e-4 * e(bare, 450, 700), e(lay, 450, 700) / e(bare, 450, 700)


def test_physical_pml_absorbs():
    """Energy left in a box after a pulse hits a PML face is far below what
    the same face reflects as a pressure-release (p=0) edge."""
    n = 64
    steps = 400
    sig = rooms.ricker(steps, DT, 1200.0)
    src, mics = [[32, 32, 32]], [[37, 32, 32]]
    prof = rooms.pml_profile(10, DX, C0)
    bare = orc.Oracle(None, None, None, C0, DX, DT, False, src, sig[None], mics, precision="f64", dims=(n, n, n))
    absorbed = orc.Oracle(None, None, None, C0, DX, DT, False, src, sig[None], mics, precision="f64",
                          dims=(n, n, n), pml_mode=2, pml_width=10, pml_faces=63, pml_profile=prof)
    tb, ta = bare.run(steps)[0], absorbed.run(steps)[0]
    late = slice(150, steps)  # after the first echoes from the faces
    eb = np.sum(tb[late] ** 2)
    ea = np.sum(ta[late] ** 2)
    assert ea < 1e-3 * eb, (ea, eb)
    # the direct sound is untouched by the layer
    early = slice(0, 45)
    assert np.abs(tb[early]).max() > 0.1 * np.abs(tb).max()
    assert np.array_equal(tb[early], ta[early])


def test_fp32_tracks_fp64():
    v, a, b = _small_room()
    sig = rooms.ricker(400, DT)
    o32 = orc.Oracle(v, a, b, C0, DX, DT, False, [[5, 5, 5]], sig[None], [[14, 12, 9]])
    o64 = orc.Oracle(v, a, b, C0, DX, DT, False, [[5, 5, 5]], sig[None], [[14, 12, 9]], precision="f64")
    x, y = o32.run(400), o64.run(400)
    assert np.linalg.norm(x - y) / np.linalg.norm(y) < 1e-4


def test_classifier_scan_range_brute_force():
    """The scan of lrs-data.hpp:154-158 runs over 1 <= x < vx-1 of the voxel
    volume (lrs-data.hpp:101-105).  Brute force in numpy over a volume padded
    into a larger grid: the boundary cells are exactly the air cells inside
    that range with a solid neighbour, and the whole-grid scan adds the cells
    outside it."""
    rng = np.random.Generator(np.random.PCG64(8))
    vx, vy, vz = 18, 16, 14
    v = rooms.repair(rooms.add_boxes(rooms.shoebox(vx, vy, vz, [0, 1, 0, 1, 0, 1]), rng, 4, 2, 5, 2))
    grid = np.full((20, 22, 25), -1, np.int32)
    grid[:vz, :vy, :vx] = v

    def brute(lim):
        s = grid != -1
        cells = set()
        for z in range(1, lim[2] - 1):
            for y in range(1, lim[1] - 1):
                for x in range(1, lim[0] - 1):
                    if s[z, y, x]:
                        continue
                    if s[z, y, x + 1] or s[z, y, x - 1] or s[z, y + 1, x] or s[z, y - 1, x] or s[z + 1, y, x] \
                            or s[z - 1, y, x]:
                        cells.add((x, y, z))
        return cells

    for lim in ((vx, vy, vz), (25, 22, 20)):
        pos, grp = orc.classify(grid, vol=lim)
        assert {tuple(p[:3]) for p in pos} == brute(lim)
    assert len(orc.classify(grid)[1]) > len(orc.classify(grid, vol=(vx, vy, vz))[1])
