"""Oracle against receiver traces the reference itself produced.

tests/golden/ie224_impulse_first8000.npz holds the first 8000 samples of the
14 traces in the reference's examples/ie224-impulse-results (a GPU run of
examples/ie224-impulse.py, written with 6 significant digits); the room is
tests/golden/ie224 (the reference's examples/models/ie224).  The oracle runs
in compat mode with the reference constants (c=340, dx=0.02, dt=3.125e-5)
and the same inputs: 5*impulse(2000, ., 32000) at [4,1.6,5.5] m (actuator),
receivers at the positions of examples/ie224-impulse.py:19-38.
"""
import os

import numpy as np
import pytest

from anc_field_py import _afx, ancutil
import orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MICS = [[4, 1.6, 5.3], [4, 1.6, 4.5], [4, 1.6, 6.5], [3, 1.6, 5.5], [5, 1.6, 5.5], [4, 0.6, 5.5], [4, 2.6, 5.5]]
SOURCES = {"noise": [4, 1.6, 1.0], "actuator": [4, 1.6, 5.5]}


def cell(p, res=np.float32(0.02)):
    """(cl_int)(pos / res_) in single precision, simulation.hpp:322,338."""
    return [int(np.float32(q) / res) for q in p]


def golden(source):
    g = np.load(os.path.join(GOLD, "ie224_impulse_first8000.npz"))
    names = list(g["names"])
    k = names.index(f"ie224-{source}-to-error")
    return g["traces"][k:k + 7]


def test_cell_quantisation():
    assert cell([4, 1.6, 5.5]) == [200, 80, 275]
    assert cell([4, 0.6, 5.5]) == [200, 30, 275]


@pytest.mark.slow
def test_oracle_reproduces_reference_actuator_traces():
    vox, a, b = _afx.voxelize_kerkythea(os.path.join(GOLD, "ie224", "model.xml"))
    assert vox.shape == (576, 160, 416)
    steps = 250
    src = (5 * ancutil.impulse(2000, steps, 32000)).astype(np.float32)
    o = orc.Oracle(vox, a, b, 340.0, 0.02, 3.125e-05, True, [cell(SOURCES["actuator"])], src[None],
                   [cell(m) for m in MICS], threads=min(8, os.cpu_count() or 1))
    got = o.run(steps)
    want = golden("actuator")[:, :steps]
    for i in range(7):
        rel = np.linalg.norm(got[i] - want[i]) / np.linalg.norm(want[i])
        # the fixture carries 6 significant digits: 1e-6 relative is its resolution
        assert rel < 5e-6, (i, rel)


def test_dsp_model_voxelizes_like_the_reference():
    """The reference's dsp room (tests/golden/dsp): the grid its loader
    builds, materials numbered in sorted-name order with the coefficients
    read from <name>_a.dat / _b.dat (kerkythea.hpp:63-196, util.hpp:101-112),
    and the same boundary groups from the library's classifier and the
    oracle's."""
    d = os.path.join(GOLD, "dsp")
    vox, a, b = _afx.voxelize_kerkythea(os.path.join(d, "model.xml"))
    assert vox.shape == (384, 128, 256)
    names = sorted(f[:-6] for f in os.listdir(d) if f.endswith("_a.dat"))
    assert names == ["absorber", "concrete", "door", "glass", "wood"]
    assert a.shape == b.shape == (5, 4)
    for k, n in enumerate(names):
        assert np.array_equal(a[k], np.loadtxt(os.path.join(d, f"{n}_a.dat"), dtype=np.float32))
        assert np.array_equal(b[k], np.loadtxt(os.path.join(d, f"{n}_b.dat"), dtype=np.float32))
    assert set(np.unique(vox).tolist()) <= set(range(-1, 5)) and (vox == -1).any()
    counts = _afx.classify_counts(vox, compat=True)
    _, grp = orc.classify(vox)
    assert counts.sum() > 0
    assert np.array_equal(counts, np.bincount(grp, minlength=26))
