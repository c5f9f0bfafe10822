"""Host-side parts of the surrogate workflow (no GPU): the identification
post-processing (2 kHz low-pass and decimation, as identification.py:36
keeps every 4th sample), the query encoding, the pair bookkeeping and the
no-network baseline."""
import numpy as np

from anc_field_py import rooms, surrogate


def test_lowpass_decimate_band_limits_and_delays_nothing():
    dt = rooms.courant_dt(0.02)
    n = 4096
    t = np.arange(n) * dt
    lo = np.sin(2 * np.pi * 300 * t)
    hi = np.sin(2 * np.pi * 9000 * t)
    y = surrogate.lowpass_decimate(np.stack([lo, hi, lo + hi]), dt, 1024, 4, 2000.0)
    mid = slice(50, 970)
    assert y.shape == (3, 1024)
    assert np.allclose(y[0, mid], lo[::4][mid], atol=2e-3)  # zero phase: no delay in the pass band
    assert np.abs(y[1, mid]).max() < 2e-3
    assert np.allclose(y[2], y[0] + y[1], atol=1e-6)  # linear


def test_lowpass_decimate_rejects_short_traces():
    import pytest
    with pytest.raises(ValueError):
        surrogate.lowpass_decimate(np.zeros((1, 100)), 1e-4, 1024, 4, 2000.0)


def test_rel_l2_and_pairs():
    a = np.array([[3.0, 4.0], [1.0, 0.0]])
    assert np.allclose(surrogate.rel_l2(a * 1.1, a), [0.1, 0.1])
    src = np.array([[1, 2, 3], [4, 5, 6]], np.int32)
    rcv = [np.array([[7, 8, 9], [10, 11, 12]]), np.array([[13, 14, 15]])]
    s, r = surrogate.pairs(src, rcv)
    assert s.tolist() == [[1, 2, 3], [1, 2, 3], [4, 5, 6]]
    assert r.tolist() == [[7, 8, 9], [10, 11, 12], [13, 14, 15]]


def test_free_cells_avoid_pml_and_walls():
    w = rooms.furnished(steps=10)
    rng = np.random.Generator(np.random.PCG64(0))
    c = surrogate.free_cells(w, rng, 500)
    nz, ny, nx = w.vox.shape
    assert len({tuple(v) for v in c}) == 500
    assert (c[:, 0] < nx - 1 - w.pml_width - 2).all()  # outside the x+ PML slab and its margin
    assert (w.vox[c[:, 2], c[:, 1], c[:, 0]] == rooms.AIR).all()


def test_nearest_pair_baseline_picks_the_nearest_training_response():
    ts = np.array([[0, 0, 0], [10, 10, 10]], np.int32)
    tr = np.array([[1, 1, 1], [20, 20, 20]], np.int32)
    ir = np.array([[1.0, 0.0], [0.0, 1.0]])
    got = surrogate.nearest_pair_baseline(ts, tr, ir, np.array([[9, 9, 9], [1, 0, 0]]), np.array([[19, 20, 21], [1, 2, 1]]))
    assert np.array_equal(got, ir[[1, 0]])


def test_zone_lattice_inside_and_trilinear_interpolation():
    z = surrogate.SourceZone([10, 20, 30], 5, 2, np.zeros((2, 3), np.int32))
    lat = z.lattice()
    assert len(lat) == 27 and lat[0].tolist() == [10, 20, 30] and lat[1].tolist() == [12, 20, 30]
    assert lat[-1].tolist() == [14, 24, 34]
    assert z.inside(np.array([[14, 24, 34], [13, 21, 30]])).all()
    assert not z.inside(np.array([[15, 20, 30], [9, 20, 30]])).any()
    # responses affine in the source cell (per receiver) are interpolated exactly, and
    # lattice cells return their own response
    coef = np.array([[0.5, -1.0, 2.0], [1.5, 0.25, -0.75]])
    f = lambda c, r: np.array([coef[r] @ c + r, 3.0 * r - coef[r] @ c])
    z.train_ir = np.stack([np.stack([f(c, r) for r in range(2)]) for c in lat])
    q = np.array([[11, 23, 33], [13, 20, 31], [12, 22, 34], [14, 24, 34]])
    for r in range(2):
        got = z.interpolate(q, np.full(len(q), r))
        assert np.allclose(got, [f(c, r) for c in q], rtol=0, atol=1e-4)


def test_ensemble_spread_and_threshold():
    a = np.array([[3.0, 4.0], [1.0, 0.0]])
    b = np.array([[3.0, 4.0], [0.8, 0.0]])
    sp = surrogate.Surrogate.spread(np.stack([a, b]))
    # two members: |a - b| / sqrt(2) over the norm of their mean
    assert np.allclose(sp, [0.0, 0.2 / np.sqrt(2) / 0.9])
    err = np.array([0.01, 0.02, 0.06, 0.03, 0.02])
    spread = np.array([0.1, 0.3, 0.4, 0.5, 0.35])
    # the good answer at 0.5 lies above the bad one at 0.4, so 0.35 is the threshold
    assert surrogate.spread_threshold(spread, err, 0.05) == 0.35
    assert surrogate.spread_threshold(spread, err, 0.1) == 0.5
    assert surrogate.spread_threshold(spread[2:3], err[2:3], 0.05) == -1.0
