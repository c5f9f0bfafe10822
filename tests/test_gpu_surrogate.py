"""Surrogate workflow on a small room: batched FDTD impulse responses ->
BF16 MFMA training -> validation -> queries with FDTD fallback.

Room: the 64x48x32 shoebox with an 8-cell PML (rooms.shoebox_small); 128
taps after a 2 kHz low-pass and decimation by 4 (512 steps); 16 sources x
16 receivers for training, 4 x 16 held out.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAPS, DECIM = 128, 4
NET = dict(n_freq=16, width=128, depth=4, batch=64, seed=3, freq_min=0.25, freq_max=8.0)


@pytest.fixture(scope="module")
def data():
    print("\n[case] surrogate data: shoebox_small, 20 sources x 16 receivers, 512 steps", flush=True)
    from anc_field_py import rooms, surrogate
    w = rooms.shoebox_small(seed=1, steps=TAPS * DECIM)
    rng = np.random.Generator(np.random.PCG64(11))
    src = surrogate.free_cells(w, rng, 20)
    rcv = [surrogate.free_cells(w, rng, 16) for _ in range(20)]
    sur = surrogate.Surrogate(w, taps=TAPS, decim=DECIM, rooms_per_handle=8, **NET)
    ir = sur.simulate(src, rcv)
    yield w, sur, src, rcv, ir
    sur.close()


def test_simulate_matches_single_room_run(data):
    from anc_field_py import _afx, surrogate
    w, sur, src, rcv, ir = data
    s = _afx.Sim(c=w.c, dx=w.dx, dt=w.dt, res=w.dx, compat=False)
    s.set_voxels(w.vox)
    s.set_materials(w.a, w.b)
    s.set_pml(w.pml_width, w.pml_faces, w.pml_profile)
    s.add_source_cell(src[3], np.array([1.0], np.float32))
    for c in rcv[3]:
        s.add_receiver_cell(c)
    s.run(TAPS * DECIM)
    want = surrogate.lowpass_decimate(s.receivers(), w.dt, TAPS, DECIM, 2000.0)
    s.close()
    assert np.array_equal(ir[3 * 16:4 * 16], want)
    assert np.all(np.linalg.norm(want, axis=1) > 0)


def test_lowpass_decimate_keeps_passband_and_drops_stopband():
    from anc_field_py import surrogate
    dt = 1.0 / 32000
    n = np.arange(4096)
    lo = np.sin(2 * np.pi * 500 * n * dt)
    hi = np.sin(2 * np.pi * 6000 * n * dt)
    y = surrogate.lowpass_decimate(np.stack([lo, hi]), dt, 1000, 4, 2000.0)
    mid = slice(100, 900)
    assert np.allclose(y[0, mid], lo[::4][mid], atol=5e-3)
    assert np.abs(y[1, mid]).max() < 1e-2


def test_training_validation_and_fallback(data):
    print("\n[case] surrogate training 3000 steps + queries", flush=True)
    from anc_field_py import surrogate
    w, sur, src, rcv, ir = data
    ps, pr = surrogate.pairs(src[:16], rcv[:16])
    hs, hr = surrogate.pairs(src[16:], rcv[16:])
    t_ir, h_ir = ir[:len(ps)], ir[len(ps):]
    loss = sur.fit(sur.encode(ps, pr), t_ir, 3000, 2e-3, 2e-5, seed=1)
    assert loss[-100:].mean() < 0.05 * loss[:20].mean()
    e_tr = sur.validate(ps, pr, t_ir)
    e_ho = surrogate.rel_l2(sur.predict(hs, hr), h_ir)
    ok = e_tr <= sur.bound
    print(f"training pairs: median rel L2 {np.median(e_tr):.3e}, within 5e-2: {ok.mean():.2f}; "
          f"held-out median {np.median(e_ho):.3e}")
    assert np.median(e_tr) < 0.1
    assert ok.sum() >= len(ok) // 4
    # queries: validated training pairs and unseen pairs, mixed
    pick = np.concatenate([np.flatnonzero(ok)[:24], np.flatnonzero(~ok)[:8]])
    qs = np.concatenate([ps[pick], hs[:16]])
    qr = np.concatenate([pr[pick], hr[:16]])
    truth = np.concatenate([t_ir[pick], h_ir[:16]])
    runs = sur.fdtd_runs
    out, fb, err = sur.query(qs, qr)
    n_ok = min(24, int(ok.sum()))
    # a source with any untrusted pair is simulated, and its run answers all of its pairs
    rerun = {tuple(c) for c in qs[n_ok:]}
    assert fb.tolist() == [tuple(c) in rerun for c in qs]
    assert not fb[:n_ok].all()
    # every answer is within the bound of the FDTD response; fallbacks are the FDTD response
    assert np.all(surrogate.rel_l2(out, truth) <= sur.bound)
    assert np.array_equal(out[fb], truth[fb])
    assert np.all(np.isfinite(err[fb])) and np.all(np.isnan(err[~fb]))
    assert sur.fdtd_runs - runs == len({tuple(s) for s in qs[fb]})
    # the gain: a validated pair costs one network row instead of an FDTD run
    import time
    t0 = time.perf_counter()
    sur.predict(ps, pr)
    t_net = (time.perf_counter() - t0) / len(ps)
    t0, n0 = sur.fdtd_seconds, sur.fdtd_runs
    sur.simulate(src[:8], rcv[:8])
    t_fdtd = (sur.fdtd_seconds - t0) / (8 * 16)
    print(f"per impulse response: network {t_net * 1e6:.1f} us, FDTD {t_fdtd * 1e6:.1f} us "
          f"({t_fdtd / t_net:.0f}x)")
    assert sur.fdtd_runs - n0 == 8
    assert t_net < t_fdtd


def test_source_zone_ensemble_gain_and_gate():
    print("\n[case] source zone: furnished 128x96x64, 9^3 box, lattice step 2, 8 receivers, 512 taps, "
          "2 x 6000 steps", flush=True)
    from anc_field_py import rooms, surrogate
    taps = 512
    w = rooms.furnished(seed=2, steps=taps * DECIM, dims=(128, 96, 64), nbox=6)
    rng = np.random.Generator(np.random.PCG64(13))
    size, step = 9, 2
    lo = surrogate.zone_box(w, rng, size)
    mics = surrogate.free_cells(w, rng, 8)
    zone = surrogate.SourceZone(lo, size, step, mics)
    assert len(zone.lattice()) == 125 and zone.inside(zone.lattice()).all()
    sur = surrogate.Surrogate(w, taps=taps, decim=DECIM, members=2, seed=1, freq_min=0.25, freq_max=16.0)
    try:
        surrogate.fit_zone(sur, zone, 6000, 1e-3, 1e-5, seed=3)
        cells = lo + np.stack(np.meshgrid(*[np.arange(size)] * 3, indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
        off = cells[np.any((cells - lo) % step, axis=1)]
        pick = off[rng.permutation(len(off))]
        val, test = pick[:16], pick[16:32]
        e_val, t_val, sp_val = surrogate.calibrate_zone(sur, zone, val)
        assert np.all(sp_val > 0)  # the members differ
        ts, tr_ = surrogate.pairs(test, [mics] * len(test))
        truth = sur.simulate(test, [mics] * len(test))
        e_net = surrogate.rel_l2(sur.predict(ts, tr_), truth)
        e_tri = surrogate.rel_l2(zone.interpolate(ts, np.tile(np.arange(len(mics)), len(test))), truth)
        runs = sur.fdtd_runs
        ans, fb, err = sur.query(ts, tr_)
        per_src = fb.reshape(len(test), -1)
        print(f"held-out zone sources: ensemble median rel L2 {np.median(e_net):.3e} "
              f"(within 5e-2: {np.mean(e_net <= 0.05):.2f}), trilinear {np.median(e_tri):.3e} "
              f"({np.mean(e_tri <= 0.05):.2f}); threshold {zone.delta:.3e}; "
              f"sources answered by the network {int((~per_src.any(1)).sum())} of {len(test)}")
        # the gain over the same lattice responses without a network
        assert np.median(e_net) < 0.8 * np.median(e_tri)
        assert np.mean(e_net <= sur.bound) >= np.mean(e_tri <= sur.bound)
        # per-source decisions; every answer within the bound; runs only for rejected sources
        assert np.all(per_src.all(1) | ~per_src.any(1))
        assert np.all(surrogate.rel_l2(ans, truth) <= sur.bound)
        assert np.array_equal(ans[fb], truth[fb])
        assert sur.fdtd_runs - runs == int(per_src.any(1).sum())
    finally:
        sur.close()
