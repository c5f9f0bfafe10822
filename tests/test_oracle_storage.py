"""The oracle's FP16/BF16 storage rounding (orc_round_store) against two
independent conversions: numpy's float32 -> float16 (IEEE binary16, round to
nearest even, subnormals kept, overflow to infinity) and torch's float32 ->
bfloat16 (round to nearest even).  Inputs cover both ranges of binary16
(normal and subnormal), exact ties at each, the overflow threshold, signed
zeros and random bit patterns over the whole float32 range.
"""
import numpy as np
import pytest

import orc


def _inputs():
    rng = np.random.Generator(np.random.PCG64(90))
    bits = rng.integers(0, 2**32, 200000, dtype=np.uint64).astype(np.uint32)
    rand = bits.view(np.float32)
    rand = rand[np.isfinite(rand)]
    special = np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e30, -1e30,
                        2.0**-14, 2.0**-24, 2.0**-25, 3 * 2.0**-25, 2.0**-26, 6.1e-5, 5.9e-8, 1e-45, 1e-40,
                        3.0e38, 3.4028235e38, -3.4028235e38], np.float32)
    # exact ties: halfway between adjacent binary16 values, normal and subnormal
    h = np.arange(1, 2**15 - 1024, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    ties16 = (h.astype(np.float64) + np.spacing(np.abs(h.astype(np.float16))).astype(np.float64) / 2).astype(np.float32)
    # and between adjacent bfloat16 values (16 low bits 0x8000)
    b = (rng.integers(0, 2**16, 5000, dtype=np.uint64).astype(np.uint32) << 16) | 0x8000
    ties_bf = b.view(np.float32)
    ties_bf = ties_bf[np.isfinite(ties_bf)]
    # near-zero and around 1 (the scaled fields' range), densely
    dense = np.concatenate([rng.uniform(-1, 1, 50000), rng.uniform(-1e-4, 1e-4, 50000),
                            rng.uniform(-1e-7, 1e-7, 20000)]).astype(np.float32)
    return np.concatenate([rand, special, ties16, -ties16, ties_bf, dense])


def _round(x, storage):
    lib = orc.lib
    mode = orc.STORE_MODES[storage]
    return np.fromiter((lib.orc_round_store(float(v), mode) for v in x), np.float32, len(x))


def test_binary16_rounding_matches_numpy():
    x = _inputs()
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float32)
    got = _round(x, "fp16")
    bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    assert not bad.any(), (x[bad][:5], got[bad][:5], want[bad][:5])
    assert np.array_equal(np.signbit(got), np.signbit(want))


def test_bfloat16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = _inputs()
    want = torch.from_numpy(x.copy()).to(torch.bfloat16).to(torch.float32).numpy()
    got = _round(x, "bf16")
    bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    assert not bad.any(), (x[bad][:5], got[bad][:5], want[bad][:5])
    assert np.array_equal(np.signbit(got), np.signbit(want))


def test_round_store_mode_zero_is_identity():
    x = _inputs()[:5000]
    assert np.array_equal(_round(x, None).view(np.uint32), x.view(np.uint32))
