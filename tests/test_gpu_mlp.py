"""Surrogate MLP kernels (BF16 MFMA) against the float64 CPU oracle.

The oracle (oracle/mlp64.py) is pinned by finite differences on the CPU.
Here the GPU forward pass and gradients are compared with it twice: with the
oracle rounding to BF16 where the kernels do (tight bound: only FP32
accumulation order differs) and in plain float64 (the BF16 error itself).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp64  # noqa: E402

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _oracle(m, bf16):
    c = m.cfg
    return mlp64.Mlp64(m.params(), m.frequencies(), c.in_dim, c.n_freq, c.width, c.depth, c.out_dim, bf16=bf16)


def test_mlp_forward_matches_oracle():
    print("\n[case] mlp forward 6-768-512x5-1024, 300 rows", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(0)
    with _afx.Mlp(seed=7) as m:
        x = rng.uniform(0, 1, (300, 6)).astype(np.float32)
        y = m.forward(x)
        # chunking: rows 0..255 and 256..299 are independent of each other
        assert np.array_equal(m.forward(x[:256]), y[:256])
        assert np.array_equal(m.forward(x[256:]), y[256:])
        e16 = _rel(y, _oracle(m, True).forward(x))
        e64 = _rel(y, _oracle(m, False).forward(x))
        print(f"forward rel L2: vs BF16-rounding oracle {e16:.2e}, vs float64 {e64:.2e}")
        assert e16 < 2e-3, e16
        assert e64 < 3e-2, e64


def test_mlp_gradients_match_oracle():
    print("\n[case] mlp gradients, 200 of 256 rows", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(1)
    with _afx.Mlp(seed=11) as m:
        rows = 200
        x = rng.uniform(0, 1, (rows, 6)).astype(np.float32)
        t = (0.05 * rng.standard_normal((rows, 1024))).astype(np.float32)
        w = (1.0 / (rows * np.sum(t.astype(np.float64) ** 2, axis=1))).astype(np.float32)
        g, loss, y = m.gradients(x, t, w)
        o16, o64 = _oracle(m, True), _oracle(m, False)
        g16, l16, y16 = o16.grad(x, t, w)
        g64, l64, _ = o64.grad(x, t, w)
        assert _rel(y, y16) < 2e-3
        assert abs(loss - l16) / l16 < 1e-4, (loss, l16)
        off = 0
        for l, (fi, fo) in enumerate(o64.dims):
            for name, n in (("W", fi * fo), ("b", fo)):
                e16 = _rel(g[off:off + n], g16[off:off + n])
                e64 = _rel(g[off:off + n], g64[off:off + n])
                print(f"layer {l} d{name}: rel L2 vs BF16-rounding oracle {e16:.2e}, vs float64 {e64:.2e}")
                assert e16 < 1e-2, (l, name, e16)
                assert e64 < 1e-1, (l, name, e64)
                off += n
        assert off == g.size
        assert _rel(g, g64) < 5e-2


def _toy(rows, rng):
    x = rng.uniform(0, 1, (rows, 2)).astype(np.float32)
    k = np.arange(64)[None, :]
    t = (np.cos(2 * np.pi * (x[:, :1] * k / 16.0)) * np.exp(-2.0 * x[:, 1:2]) + 0.5 * x[:, 1:2]).astype(np.float32)
    return x, t


def test_mlp_training_decreases_loss_and_is_deterministic():
    print("\n[case] mlp training toy 2-32-64x3-64, 600 steps", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(2)
    x, t = _toy(512, rng)
    kw = dict(in_dim=2, n_freq=8, width=64, depth=3, out_dim=64, batch=64, seed=5, freq_min=0.5, freq_max=4.0)
    w = (1.0 / (64 * np.sum(t.astype(np.float64) ** 2, axis=1))).astype(np.float32)
    runs = []
    for _ in range(2):
        with _afx.Mlp(**kw) as m:
            m.set_data(x, t, w)
            loss = m.train(600, 3e-3, 1e-4, seed=9)
            runs.append((m.params(), loss, m.stats().train_steps))
    (p0, l0, n0), (p1, l1, n1) = runs
    assert n0 == 600 and n1 == 600
    assert np.array_equal(p0, p1) and np.array_equal(l0, l1)
    first, last = float(np.mean(l0[:20])), float(np.mean(l0[-20:]))
    print(f"mean relative squared error: first 20 steps {first:.3e}, last 20 {last:.3e}")
    assert last < 0.05 * first, (first, last)


def test_mlp_first_adam_step():
    print("\n[case] mlp one Adam step vs gradient", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(4)
    x, t = _toy(64, rng)
    kw = dict(in_dim=2, n_freq=8, width=64, depth=3, out_dim=64, batch=64, seed=3, freq_min=0.5, freq_max=4.0)
    w = np.full(64, 1.0 / 64, np.float32)
    with _afx.Mlp(**kw) as m:
        p0 = m.params()
        g, _, _ = m.gradients(x, t, w)
        m.set_data(x, t, w)
        lr = 1e-3
        m.train(1, lr, lr, seed=0)
        p1 = m.params()
    # first Adam step: m_hat = g, v_hat = g^2, so the step is lr * g / (|g| + eps)
    big = np.abs(g) > 1e-4
    step = (p0 - p1).astype(np.float64)
    assert np.allclose(step[big], lr * np.sign(g[big]), rtol=1e-3, atol=0)
    assert np.all(np.abs(step[~big]) <= lr * 1.001)


def test_mlp_graph_replay_matches_eager_training():
    print("\n[case] mlp training 100 steps, captured graphs vs eager launches", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(6)
    x, t = _toy(256, rng)
    w = (1.0 / (64 * np.sum(t.astype(np.float64) ** 2, axis=1))).astype(np.float32)
    kw = dict(in_dim=2, n_freq=8, width=64, depth=3, out_dim=64, batch=64, seed=8, freq_min=0.5, freq_max=4.0)
    out = []
    for flags in (0, _afx.MLP_NO_GRAPH):
        with _afx.Mlp(flags=flags, **kw) as m:
            m.set_data(x, t, w)
            loss = m.train(100, 2e-3, 1e-4, seed=4)  # 3 graph replays of 32 steps + 4 eager steps
            loss2 = m.train(40, 1e-3, 1e-4, seed=5)  # continues: Adam step count and moments carry over
            out.append((m.params(), loss, loss2))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_mlp_forward_chunking_and_device_path():
    """Inference chunks of up to 32 768 rows: 40 000 rows (a full chunk on
    128x128 tiles, written in place on the device path, then a 7232-row tail
    on 64x64 tiles) give the same bits as 300-row pieces (split-K 32x32
    kernels), from host and from device memory."""
    print("\n[case] mlp forward 40 000 rows: host, device, 300-row pieces", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(9)
    n = 40000
    x = rng.uniform(0, 1, (n, 6)).astype(np.float32)
    with _afx.Mlp(seed=13) as m:
        y = m.forward(x)
        parts = np.concatenate([m.forward(x[i:i + 300]) for i in range(0, n, 300)])
        assert np.array_equal(y, parts)
        with _afx.DeviceBuffer(x.nbytes) as xd, _afx.DeviceBuffer(n * 1024 * 4) as yd:
            xd.upload(x)
            m.forward_device(xd, n, yd)
            yg = yd.download((n, 1024))
        assert np.array_equal(yg, y)
        st = m.stats()
        assert st.last_forward_rows == n and st.last_forward_gpu_ms > 0
        print(f"device forward of {n} rows: {st.last_forward_gpu_ms:.3f} ms")


def test_mlp_large_m_gemm_matches_training_kernels():
    """Forward passes of 1024 rows or more run on 128x128 tiles (row counts
    that are multiples of 128) or 64x64 tiles (multiples of 64) staged through
    LDS; their outputs equal, bit for bit, the outputs the training path's
    split-K 32x32 kernels give for the same rows."""
    print("\n[case] mlp 128x128 and 64x64 inference tiles vs split-K training kernels", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(10)
    x = rng.uniform(0, 1, (2048, 6)).astype(np.float32)
    t = (0.05 * rng.standard_normal((256, 1024))).astype(np.float32)
    w = np.full(256, 1.0 / 256, np.float32)
    with _afx.Mlp(seed=21) as m:
        _, _, y_train = m.gradients(x[:256], t, w)
        y_128 = m.forward(x)          # 2048 rows: 128x128 tiles
        y_64 = m.forward(x[:1088])    # 1088 rows: 64x64 tiles
    assert np.array_equal(y_128[:256], y_train)
    assert np.array_equal(y_64[:256], y_train)
    assert np.array_equal(y_64, y_128[:1088])


@pytest.mark.parametrize("in_dim,n_freq,width,out_dim", [(3, 64, 384, 128), (2, 32, 128, 256)])
def test_mlp_large_m_gemm_other_shapes_match_training_kernels(in_dim, n_freq, width, out_dim):
    """The 128x128 inference tiles on shapes off the surrogate's own: K of
    384 (a multiple of 128, not of 256) and of 128 (one k-step per slice),
    and 1152 rows, whose 9 x 3 (or 9 x 1, 9 x 2) tiles are not a multiple of
    8, so the tiles keep the row-major order instead of the XCD-grouped one.
    Outputs equal the training path's split-K kernels bit for bit."""
    print(f"\n[case] mlp inference tiles, in_dim {in_dim}, n_freq {n_freq}, width {width}, out {out_dim}", flush=True)
    from anc_field_py import _afx
    rng = np.random.default_rng(11)
    rows = 1152
    x = rng.uniform(0, 1, (rows, in_dim)).astype(np.float32)
    t = (0.05 * rng.standard_normal((256, out_dim))).astype(np.float32)
    w = np.full(256, 1.0 / 256, np.float32)
    with _afx.Mlp(seed=22, in_dim=in_dim, n_freq=n_freq, width=width, out_dim=out_dim, depth=4) as m:
        assert 2 * in_dim * n_freq in (384, 128)
        y_train = np.concatenate([m.gradients(x[k:k + 256], t, w)[2] for k in range(0, 1024, 256)])
        y = m.forward(x)  # 1152 rows: 128x128 tiles
    assert np.abs(y).max() > 0
    assert np.array_equal(y[:1024], y_train)
