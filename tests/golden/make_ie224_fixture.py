"""Write the ie224 golden fixtures from data files held in the reference.

Inputs (reference root): examples/ie224-impulse-results/ie224-<src>-to-<mic>.dat,
14 receiver traces of 128 000 samples produced by the reference's GPU run of
examples/ie224-impulse.py, and the ie224 room (examples/models/ie224: model.xml
plus <material>_a.dat/_b.dat).

Outputs (this directory):
  ie224_impulse_first8000.npz   traces[14, 8000] float32, names[14]
  ie224/                        model.xml and material files, copied as data

Run once in a container that has the reference mounted:
  python tests/golden/make_ie224_fixture.py /root/reference
"""
import os
import shutil
import sys

import numpy as np

N = 8000
SOURCES = ["noise", "actuator"]
MICS = ["error", "reference_front", "reference_back", "reference_left", "reference_right", "reference_bottom",
        "reference_top"]


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    res = os.path.join(ref, "examples", "ie224-impulse-results")
    names, traces = [], []
    for s in SOURCES:
        for m in MICS:
            name = f"ie224-{s}-to-{m}"
            x = np.loadtxt(os.path.join(res, name + ".dat"), max_rows=N)
            names.append(name)
            traces.append(x.astype(np.float32))
    np.savez_compressed(os.path.join(here, f"ie224_impulse_first{N}.npz"), traces=np.stack(traces),
                        names=np.array(names))
    dst = os.path.join(here, "ie224")
    os.makedirs(dst, exist_ok=True)
    src_model = os.path.join(ref, "examples", "models", "ie224")
    for f in sorted(os.listdir(src_model)):
        if f == "model.xml" or f.endswith(".dat"):
            shutil.copyfile(os.path.join(src_model, f), os.path.join(dst, f))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
