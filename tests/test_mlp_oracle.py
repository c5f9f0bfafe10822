"""CPU checks of the surrogate's float64 oracle (oracle/mlp64.py) and of the
host-side argument checking of afx_mlp_create.

The oracle's backward pass is pinned by central finite differences of its
own loss; its BF16 rounding helper by exact bit patterns; its features by
the closed form of a one-frequency ladder.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp64  # noqa: E402


def test_bf16_round_bit_patterns():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 2.0 ** -130, 65504.0, 0.1],
                 np.float32)
    r = mlp64.bf16_round(v)
    # 1+2^-8 is the tie between 1 and 1+2^-7: even mantissa -> 1; 1+3*2^-8 ties up to 1+2^-6
    want = [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 2.0 ** -130, 65536.0, 0.10009765625]
    assert np.array_equal(r, np.array(want))
    assert mlp64.bf16_round(np.float32(3.0)) == 3.0


def test_features_closed_form():
    x = np.array([[0.0, 0.25], [0.5, 0.125]], np.float32)
    f = np.array([1.0, 2.0], np.float32)
    F = mlp64.features(x, f)
    assert F.shape == (2, 8)
    # coordinate j=1 of row 0 (x=0.25): sin block [sin(pi/2), sin(pi)], cos block [cos(pi/2), cos(pi)]
    assert np.allclose(F[0, 4:8], [1.0, 0.0, 0.0, -1.0], atol=1e-12)
    # coordinate j=0 of row 1 (x=0.5): sin(pi), sin(2pi), cos(pi), cos(2pi)
    assert np.allclose(F[1, 0:4], [0.0, 0.0, -1.0, 1.0], atol=1e-12)
    lad = mlp64.frequencies(5, 0.5, 8.0)
    assert np.allclose(lad, [0.5, 1.0, 2.0, 4.0, 8.0])


def _small(seed=0, bf16=False):
    rng = np.random.default_rng(seed)
    in_dim, nf, width, depth, out = 2, 3, 8, 3, 5
    dims = mlp64.layer_dims(in_dim, nf, width, depth, out)
    n = sum(fi * fo + fo for fi, fo in dims)
    p = rng.normal(0, 0.5, n)
    freq = mlp64.frequencies(nf, 0.5, 2.0)
    net = mlp64.Mlp64(p, freq, in_dim, nf, width, depth, out, bf16=bf16)
    x = rng.uniform(0, 1, (4, in_dim)).astype(np.float32)
    t = rng.normal(0, 1, (4, out))
    w = rng.uniform(0.5, 2.0, 4)
    return p, freq, (in_dim, nf, width, depth, out), x, t, w


def test_oracle_gradient_matches_finite_differences():
    p, freq, dims, x, t, w = _small()
    net = mlp64.Mlp64(p, freq, *dims)
    g, loss, _ = net.grad(x, t, w)
    assert np.isclose(loss, net.loss(x, t, w))
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        q = p.copy()
        q[i] += h
        lp = mlp64.Mlp64(q, freq, *dims).loss(x, t, w)
        q[i] -= 2 * h
        lm = mlp64.Mlp64(q, freq, *dims).loss(x, t, w)
        fd[i] = (lp - lm) / (2 * h)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert err < 1e-7, err


def test_oracle_depth2_zero_hidden_weights():
    in_dim, nf, width, depth, out = 1, 2, 3, 2, 2
    dims = mlp64.layer_dims(in_dim, nf, width, depth, out)
    W0 = np.zeros((width, 2 * in_dim * nf))
    b0 = np.array([0.5, -1.0, 2.0])
    W1 = np.array([[1.0, 2.0, 3.0], [-1.0, 0.0, 1.0]])
    b1 = np.array([0.25, -0.5])
    p = np.concatenate([W0.ravel(), b0, W1.ravel(), b1])
    net = mlp64.Mlp64(p, mlp64.frequencies(nf, 1.0, 2.0), in_dim, nf, width, depth, out)
    silu = b0 / (1 + np.exp(-b0))
    y = net.forward(np.array([[0.3], [0.7]], np.float32))
    assert np.allclose(y, W1 @ silu + b1, rtol=1e-14)
    assert dims == [(4, 3), (3, 2)]


def test_bf16_oracle_close_to_float64():
    p, freq, dims, x, t, w = _small(seed=3)
    g64, l64, _ = mlp64.Mlp64(p, freq, *dims).grad(x, t, w)
    g16, l16, _ = mlp64.Mlp64(p, freq, *dims, bf16=True).grad(x, t, w)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2
    assert abs(l16 - l64) / l64 < 5e-2


def test_mlp_create_rejects_bad_shapes():
    from anc_field_py import _afx
    for kw in ({"width": 100}, {"batch": 48}, {"depth": 1}, {"n_freq": 5}):
        with pytest.raises(_afx.AfxError) as ei:
            _afx.Mlp(**kw)
        assert ei.value.code == _afx.AFX_E_ARG, kw
