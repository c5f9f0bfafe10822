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
