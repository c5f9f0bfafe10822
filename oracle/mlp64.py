"""CPU oracle of the surrogate MLP (test infrastructure only).

The surrogate has no counterpart in the reference, so this oracle restates
the network the C ABI documents (include/afx_mlp.h) from its definition:

  features  [sin(2 pi f_k x_j), cos(2 pi f_k x_j)], j-major, sin block then cos block
  hidden    z_l = a_{l-1} W_l^T + b_l,  a_l = silu(z_l) = z_l / (1 + exp(-z_l))
  output    y = a_{L-2} W_{L-1}^T + b_{L-1}
  loss      sum_r w_r |y_r - t_r|^2

and its gradient by reverse-mode differentiation, all in float64.

With ``bf16=True`` the values the GPU kernels round to BF16 (features,
hidden activations, back-propagated gradients, weight copies) are rounded
the same way (round to nearest even from float32), and everything else stays
float64.  That is the form the GPU gradients are compared with; the plain
float64 form is the ground truth.

Pins:
  - ``grad`` against central finite differences of ``loss`` in float64
    (tests/test_mlp_oracle.py), for every parameter of a small network;
  - ``forward`` of a depth-2 network with zero hidden weights reduces to the
    bias (closed form); features against numpy's sin/cos of the same ladder;
  - the BF16 rounding helper against the exact bit patterns of known values
    (1, 1 + 2^-8 ties to even, 1 + 3*2^-8, negative values, powers of two).
"""
import numpy as np


def bf16_round(a):
    """float32 -> nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def frequencies(n_freq, fmin, fmax):
    """Geometric ladder as the library computes it (double, then float32)."""
    r = np.arange(n_freq, dtype=np.float64) / max(n_freq - 1, 1)
    return (fmin * (float(fmax) / float(fmin)) ** r).astype(np.float32)


def features(x, freq, bf16=False):
    """x [rows][in_dim] float32, freq [n_freq] float32 -> [rows][2*in_dim*n_freq]."""
    x = np.asarray(x, np.float32)
    freq = np.asarray(freq, np.float32)
    rows, d = x.shape
    nf = freq.size
    cyc = (freq[None, None, :] * x[:, :, None]).astype(np.float32)  # FP32 product, as on the device
    arg = np.pi * (2.0 * cyc.astype(np.float64))
    s, c = np.sin(arg), np.cos(arg)
    out = np.concatenate([s, c], axis=2).reshape(rows, 2 * d * nf)
    if bf16:
        out = bf16_round(out.astype(np.float32))
    return out


def unpack(params, dims):
    """Flat parameter vector -> [(W [out][in], b [out])] per layer."""
    layers, o = [], 0
    p = np.asarray(params, np.float64)
    for fi, fo in dims:
        W = p[o:o + fi * fo].reshape(fo, fi)
        o += fi * fo
        b = p[o:o + fo]
        o += fo
        layers.append((W, b))
    assert o == p.size
    return layers


def layer_dims(in_dim, n_freq, width, depth, out_dim):
    f = 2 * in_dim * n_freq
    return [(f if l == 0 else width, out_dim if l == depth - 1 else width) for l in range(depth)]


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


class Mlp64:
    def __init__(self, params, freq, in_dim, n_freq, width, depth, out_dim, bf16=False):
        self.dims = layer_dims(in_dim, n_freq, width, depth, out_dim)
        self.layers = unpack(params, self.dims)
        self.freq = np.asarray(freq, np.float32)
        self.bf16 = bf16
        self.rnd = (lambda a: bf16_round(np.asarray(a, np.float32))) if bf16 else (lambda a: a)
        self.Wq = [self.rnd(W) for W, _ in self.layers]

    def _forward(self, x):
        a = features(x, self.freq, self.bf16)
        acts, zs = [a], []
        L = len(self.layers)
        for l, (W, b) in enumerate(self.layers):
            z = a @ self.Wq[l].T + b
            if l == L - 1:
                return z, acts, zs
            zs.append(z)
            a = self.rnd(_silu(z))
            acts.append(a)

    def forward(self, x):
        return self._forward(x)[0]

    def loss(self, x, t, w):
        y = self.forward(x)
        return float(np.sum(np.asarray(w, np.float64)[:, None] * (y - t) ** 2))

    def grad(self, x, t, w):
        """(flat gradient, loss, y) in the library's parameter layout."""
        y, acts, zs = self._forward(x)
        w = np.asarray(w, np.float64)[:, None]
        d = y - np.asarray(t, np.float64)
        loss = float(np.sum(w * d * d))
        g = self.rnd(2.0 * w * d)
        out = [None] * len(self.layers)
        for l in range(len(self.layers) - 1, -1, -1):
            dW = g.T @ acts[l]
            db = g.sum(axis=0)
            out[l] = (dW, db)
            if l > 0:
                g = self.rnd((g @ self.Wq[l]) * _silu_grad(zs[l - 1]))
        flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in out])
        return flat, loss, y
