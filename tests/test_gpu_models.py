"""A second reference room through the Kerkythea loader: the dsp model
(examples/models/dsp, 5 materials with 4 IIR coefficients each, a
256x128x384 grid), loaded by the library and voxelized for the oracle by the
host-only voxelizer, in compat mode with the reference's constants.  Sources
and receivers on interior cells and on cells next to walls; traces and the
newest field bit for bit against the oracle.
"""
import os

import numpy as np
import pytest

from anc_field_py import rooms
import orc

pytestmark = pytest.mark.gpu

afx = pytest.importorskip("anc_field_py._afx")

MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dsp", "model.xml")


@pytest.mark.parametrize("tblock", [1, 2])
def test_dsp_model_matches_oracle(tblock):
    print(f"\n[case] dsp model, compat, 150 steps, {tblock} step(s) per pass", flush=True)
    v, a, b = afx.voxelize_kerkythea(MODEL)
    assert a.shape[1] == 4
    rng = np.random.Generator(np.random.PCG64(91))
    inner = rooms.interior_cells(v, rng, 4)
    # air cells with a solid neighbour (boundary entries), picked from the voxels
    s_ = v != rooms.AIR
    near = np.zeros_like(s_)
    near[1:-1, 1:-1, 1:-1] = ~s_[1:-1, 1:-1, 1:-1] & (s_[1:-1, 1:-1, 2:] | s_[1:-1, 1:-1, :-2] | s_[1:-1, 2:, 1:-1]
                                                     | s_[1:-1, :-2, 1:-1] | s_[2:, 1:-1, 1:-1] | s_[:-2, 1:-1, 1:-1])
    idx = np.argwhere(near)
    walls = idx[rng.choice(len(idx), 3, replace=False)][:, ::-1].astype(np.int32)  # (x, y, z)
    steps = 150
    dt = 3.125e-05
    src = np.stack([rooms.ricker(steps, dt, 1000.0), rooms.ricker(steps, dt, 700.0, amp=0.5)])
    sc = np.stack([inner[0], walls[0]])
    mc = np.concatenate([inner[1:], walls[1:]])
    s = afx.Sim()
    s.load_kerkythea(MODEL)
    s.set_time_block(tblock)
    for cell, sig in zip(sc, src):
        s.add_source_cell(cell, sig)
    for cell in mc:
        s.add_receiver_cell(cell)
    s.run(steps)
    got, f = s.receivers(), s.pressure(0)
    st = s.stats()
    s.close()
    assert (st.nz, st.ny, st.nx) == v.shape
    o = orc.Oracle(v, a, b, 340.0, 0.02, dt, True, sc, src, mc, "f32")
    o.run(steps)
    assert np.array_equal(got, o.receivers().astype(got.dtype)), "traces differ"
    assert np.array_equal(f, o.field(0).astype(f.dtype)), "fields differ"
    assert np.abs(got).max() > 0


@pytest.mark.parametrize("precision", ["fp64", "fp16", "bf16"])
def test_ie224_precisions_match_oracle(precision):
    """The ie224 room (416x160x576), compat mode, the reference's actuator
    source and 7 receivers, 200 steps against the oracle bit for bit: FP64
    against its double build (traces come back as FP32 copies of FP64
    state), FP16/BF16 storage against its storage mode."""
    print(f"\n[case] ie224 model, {precision}, compat, 200 steps", flush=True)
    model = os.path.join(os.path.dirname(MODEL), "..", "ie224", "model.xml")
    v, a, b = afx.voxelize_kerkythea(model)
    steps = 200
    dt = 3.125e-05
    sig = rooms.ricker(steps, dt, 800.0)
    src_pos = [4, 1.6, 5.5]
    mic_pos = [[4, 1.6, 5.3], [4, 1.6, 4.5], [4, 1.6, 6.5], [3, 1.6, 5.5], [5, 1.6, 5.5], [4, 0.6, 5.5], [4, 2.6, 5.5]]
    s = afx.Sim(precision=precision)
    s.load_kerkythea(model)
    s.add_source(src_pos, sig)
    for m in mic_pos:
        s.add_receiver(m)
    s.run(steps)
    got = s.receivers()
    s.close()

    def cell(p):  # (cl_int)(pos / res_) in single precision, simulation.hpp:322,338
        return [int(np.float32(q) / np.float32(0.02)) for q in p]

    o = orc.Oracle(v, a, b, 340.0, 0.02, dt, True, [cell(src_pos)], sig[None], [cell(m) for m in mic_pos],
                   "f64" if precision == "fp64" else "f32", threads=16,
                   storage=None if precision == "fp64" else precision)
    want = o.run(steps)
    assert np.array_equal(got, want.astype(np.float32)), "traces differ"
    assert np.abs(got).max() > 0
