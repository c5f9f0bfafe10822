"""HIP compat path against the reference's own ie224 receiver traces.

Drives the library the way examples/ie224-impulse.py drives the reference
(load the Kerkythea room, add the source and receivers in metres, run), for
the 8000 samples kept in tests/golden, and compares each trace with the
reference's output (6 significant digits).
"""
import os

import numpy as np
import pytest

from anc_field_py import ancutil
from test_oracle_golden import GOLD, MICS, SOURCES, golden

pytestmark = pytest.mark.gpu
afx = pytest.importorskip("anc_field_py._afx")


@pytest.mark.parametrize("source", ["actuator", "noise"])
def test_ie224_matches_reference_traces(source):
    print(f"\n[case] ie224 {source} 8000 steps", flush=True)
    steps = 8000
    s = afx.Sim()  # reference constants, compat mode, FP32
    s.load_kerkythea(os.path.join(GOLD, "ie224", "model.xml"))
    s.add_source(SOURCES[source], (5 * ancutil.impulse(2000, 6 * 32000, 32000)).astype(np.float32))
    for m in MICS:
        s.add_receiver(m)
    s.run(steps)
    got = s.receivers()
    st = s.stats()
    s.close()
    assert (st.nx, st.ny, st.nz) == (416, 160, 576)
    want = golden(source)[:, :steps]
    rel = [float(np.linalg.norm(got[i] - want[i]) / np.linalg.norm(want[i])) for i in range(7)]
    print("rel L2 per receiver:", ["%.2e" % r for r in rel])
    assert max(rel) < 1e-4, rel
