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
