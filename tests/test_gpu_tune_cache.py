"""afx_autotune results persist across processes (AFX_TUNE_CACHE file)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys, time
sys.path.insert(0, sys.argv[1])
from anc_field_py import _afx, rooms
w = rooms.shoebox_small(steps=40)
s = rooms.make_sim(w, _afx.Sim)
t0 = time.perf_counter()
tile = s.autotune(10)
dt = time.perf_counter() - t0
print(*tile, s.time_block()[0], s.tile_split(), f"{dt:.4f}")
s.close()
"""


def _child(env):
    out = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "anc-field_amd")], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    f = out.stdout.split()
    return tuple(int(v) for v in f[:5]), float(f[5])


def test_autotune_result_survives_the_process(tmp_path):
    print("\n[case] autotune cache file across two processes", flush=True)
    path = tmp_path / "tune.txt"
    env = dict(os.environ, AFX_TUNE_CACHE=str(path))
    first, t_first = _child(env)
    lines = path.read_text().splitlines()
    assert len(lines) == 1
    fields = lines[0].split()
    assert fields[0].startswith("gfx950x") and fields[5] == "1" and fields[6:9] == ["64", "48", "32"]
    second, t_second = _child(env)
    # the second process finds the choice in the file: same result, no new line, no search
    assert second == first
    assert path.read_text().splitlines() == lines
    print(f"search {t_first:.3f} s, cached {t_second:.4f} s")
    assert t_second < t_first
    # lines that do not parse or name an unknown tile are ignored: the third
    # process searches again and appends its result (key, then tile, steps
    # per pass, XCD rotation and split)
    bad = " ".join(fields[:-4] + ["99"] + fields[-3:])
    path.write_text("garbage line\n" + bad + "\n")
    _child(env)
    after = path.read_text().splitlines()
    assert len(after) == 3 and after[2].split()[:-4] == fields[:-4]
