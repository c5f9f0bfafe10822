"""Copy the reference's dsp room (examples/models/dsp: the Kerkythea model and
its <material>_a.dat / _b.dat coefficient files, 4 IIR coefficients per
material) into tests/golden/dsp as data.  The project file dsp.wings is not
needed.  Usage: python tests/golden/copy_dsp_model.py /root/reference
"""
import glob
import os
import shutil
import sys


def main(ref):
    src = os.path.join(ref, "examples", "models", "dsp")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dsp")
    os.makedirs(dst, exist_ok=True)
    files = [os.path.join(src, "model.xml")] + sorted(glob.glob(os.path.join(src, "*_[ab].dat")))
    for f in files:
        shutil.copyfile(f, os.path.join(dst, os.path.basename(f)))
    print(f"copied {len(files)} files to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
