"""afx_mlp_train_noalloc, the training entry point that allocates no device
memory, against afx_mlp_train through the C ABI: two handles built alike
train on the same data, one through each entry point, for two calls in a
row (the second smaller); losses and parameters must agree bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(rows, out_dim, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (rows, 6)).astype(np.float32)
    t = (0.05 * rng.standard_normal((rows, out_dim))).astype(np.float32)
    w = (1.0 / (256 * np.sum(t.astype(np.float64) ** 2, axis=1))).astype(np.float32)
    return x, t, w


def test_train_noalloc_matches_train():
    print("\n[case] afx_mlp_train_noalloc vs afx_mlp_train, default net, 40 then 7 steps", flush=True)
    from anc_field_py import _afx
    x, t, w = _data(1024, 1024, 3)
    with _afx.Mlp(seed=5) as a, _afx.Mlp(seed=5) as b:
        assert np.array_equal(a.params(), b.params())
        a.set_data(x, t, w)
        b.set_data(x, t, w)
        for steps, seed in ((40, 11), (7, 12)):
            la = a.train(steps, 2e-3, 1e-4, seed=seed)
            lb = b.train_noalloc(steps, 2e-3, 1e-4, seed=seed)
            assert np.all(np.isfinite(lb)) and lb[-1] < lb[0] * 1.5
            assert np.array_equal(la, lb), (steps, np.abs(la - lb).max())
            assert np.array_equal(a.params(), b.params()), steps
        assert b.stats().train_steps == 47
        # a call beyond the scratch array is refused before any work, and the
        # handle keeps its state
        before = b.params()
        with pytest.raises(_afx.AfxError) as ei:
            b.train_noalloc(100_000, 2e-3)
        assert ei.value.code == _afx.AFX_E_RANGE
        assert "scratch" in str(ei.value)
        assert np.array_equal(b.params(), before) and b.stats().train_steps == 47


def test_train_noalloc_small_net_two_handles_share_the_scratch():
    print("\n[case] afx_mlp_train_noalloc, two small nets in turn", flush=True)
    from anc_field_py import _afx
    kw = dict(n_freq=16, width=128, depth=4, batch=64, out_dim=256, freq_min=0.25, freq_max=8.0)
    x, t, w = _data(512, 256, 4)
    with _afx.Mlp(seed=9, **kw) as a, _afx.Mlp(seed=9, **kw) as b, _afx.Mlp(seed=9, **kw) as c:
        for m in (a, b, c):
            m.set_data(x, t, w)
        la = a.train(300, 1e-3, seed=2)
        lb = b.train_noalloc(300, 1e-3, seed=2)
        # another handle's call in between overwrites the shared scratch
        c.train_noalloc(50, 1e-3, seed=7)
        la2 = a.train(20, 1e-3, seed=3)
        lb2 = b.train_noalloc(20, 1e-3, seed=3)
        assert np.array_equal(la, lb) and np.array_equal(la2, lb2)
        assert np.array_equal(a.params(), b.params())
