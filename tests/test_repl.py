"""Command surface: the afx-field REPL and the AncField wrapper.

CPU: argument checking and error messages of commandline.hpp:112-196.
GPU: the ie224 room driven through the REPL text protocol and through
AncField gives the library's traces exactly and the reference's traces to
the fixture resolution over the first samples.
"""
import os
import subprocess

import numpy as np
import pytest

from test_oracle_golden import GOLD, MICS, SOURCES, golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPL = os.path.join(ROOT, "anc-field_amd", "bin", "afx-field")


def repl(commands, timeout=120):
    p = subprocess.run([REPL], input="\n".join(commands) + "\nquit\n", capture_output=True, text=True,
                       timeout=timeout)
    assert p.returncode == 0, p.stderr
    return p.stdout


def test_repl_argument_errors():
    out = repl(["bogus", "load", "load /nonexistent/model.xml", "pml x", "minsize [1,a,2]", "addmic [1,2,3]"])
    lines = [l for l in out.split(">> ") if l.strip()]
    assert lines[0].strip() == "Error: unrecognized command"
    assert lines[1].strip() == "Error: command load takes 1 argument(s)"
    assert lines[2].strip() == "Error: file /nonexistent/model.xml doesn't exist"
    assert lines[3].strip() == "Error: argument[0] should be a(n) int"
    assert lines[4].strip() == "Error: argument[0] should be a(n) float_list"
    assert lines[5].strip().startswith("Error:")  # no device selected yet


def _src_file(tmp_path, n):
    from anc_field_py import ancutil
    sig = 5 * ancutil.impulse(2000, n, 32000)
    path = tmp_path / "src.dat"
    np.savetxt(path, sig)  # what AncField.AddSource wrote (ancfield.py:75)
    return path, sig


@pytest.mark.gpu
def test_repl_ie224_run_and_save(tmp_path):
    print("\n[case] REPL ie224 run 0.03 s", flush=True)
    afx = pytest.importorskip("anc_field_py._afx")
    seconds = 0.03
    steps = int(np.float32(seconds) * np.float32(32000))
    src, sig = _src_file(tmp_path, steps + 10)
    model = os.path.join(GOLD, "ie224", "model.xml")
    cmds = ["sdt gpu", f"load {model}", f"addsource [{','.join(map(str, SOURCES['actuator']))}] {src}"]
    cmds += [f"addmic [{','.join(map(str, m))}]" for m in MICS]
    cmds += [f"run {seconds}", f"save {tmp_path / 'out.dat'}"]
    out = repl(cmds)
    assert "BEGIN" in out and f"{steps}/{steps}" in out
    saved = np.loadtxt(tmp_path / "out.dat")
    assert saved.shape == (7, steps)
    # the library run with the same inputs, read back exactly
    s = afx.Sim()
    s.load_kerkythea(model)
    s.add_source(SOURCES["actuator"], np.loadtxt(src).astype(np.float32))
    for m in MICS:
        s.add_receiver(m)
    s.run(steps)
    lib = s.receivers()
    s.close()
    # the text file carries 6 significant digits (default ostream formatting)
    assert np.allclose(saved, lib, rtol=1e-5, atol=1e-12)
    want = golden("actuator")[:, :steps]
    rel = np.linalg.norm(saved - want) / np.linalg.norm(want)
    assert rel < 1e-5, rel


@pytest.mark.gpu
def test_repl_device_and_setup_commands(tmp_path):
    """lsd, sd, cd, sdt and vis answer as the command table says
    (commandline.hpp:107-196); minsize and pml reach the simulation: a
    session with both gives the library's traces for the same calls."""
    print("\n[case] REPL device commands, minsize and pml", flush=True)
    afx = pytest.importorskip("anc_field_py._afx")
    model = os.path.join(GOLD, "ie224", "model.xml")
    out = repl(["lsd", "sd [0]", "cd", "sdt cpu", "sdt tpu", "vis 0.5"])
    parts = [l.strip() for l in out.split(">> ") if l.strip()]
    assert parts[0].startswith("Available compute devices") and "\n0. " in parts[0]
    assert parts[1].endswith("added to active devices")
    assert parts[2] == "Compute devices removed"
    assert parts[3].startswith("Error:")
    assert parts[4] == "Valid command arguments are 'cpu' or 'gpu'"
    assert parts[5].startswith("Error:")
    s = afx.Sim()
    s.load_kerkythea(model)
    s.run(0)
    st = s.stats()
    nx, ny, nz = st.nx, st.ny, st.nz
    s.close()
    size = [float(np.float32(f"{v:.4f}")) for v in ((nx + 32) * 0.02, ny * 0.02, nz * 0.02)]
    seconds = 0.01
    steps = int(np.float32(seconds) * np.float32(32000))
    src, sig = _src_file(tmp_path, steps + 10)
    cmds = ["sd [0]", f"load {model}", "minsize [%.4f,%.4f,%.4f]" % tuple(size), "pml 8",
            f"addsource [{','.join(map(str, SOURCES['actuator']))}] {src}"]
    cmds += [f"addmic [{','.join(map(str, m))}]" for m in MICS]
    cmds += [f"run {seconds}", f"save {tmp_path / 'out.dat'}"]
    out = repl(cmds)
    assert f"{steps}/{steps}" in out and "Error" not in out
    saved = np.loadtxt(tmp_path / "out.dat")
    runs = []
    for setup in (True, False):
        s = afx.Sim()
        s.load_kerkythea(model)
        if setup:
            s.set_min_size(*size)
            s.set_pml(8)
        s.add_source(SOURCES["actuator"], np.loadtxt(src).astype(np.float32))
        for m in MICS:
            s.add_receiver(m)
        s.run(steps)
        st = s.stats()
        runs.append((s.receivers(), (st.nx, st.ny, st.nz)))
        s.close()
    (lib, dims_set), (plain, dims_plain) = runs
    # SetMinSize: RoundTo((size_t)(x / res) + 1, 32) in single precision (simulation.hpp:279-284)
    want = tuple(int(-(-(int(np.float32(v) / np.float32(0.02)) + 1) // 32) * 32) for v in size)
    assert dims_set == want and dims_plain == (nx, ny, nz), (dims_set, want)
    assert want[0] > nx
    assert np.allclose(saved, lib, rtol=1e-5, atol=1e-12)
    assert not np.allclose(lib, plain, rtol=1e-3, atol=1e-9)


@pytest.mark.gpu
def test_ancfield_matches_library():
    print("\n[case] AncField ie224 0.02 s", flush=True)
    from anc_field_py import _afx, ancutil
    from anc_field_py.ancfield import AncField
    sig = 5 * ancutil.impulse(2000, 6 * 32000, 32000)
    anc = AncField("gpu", os.path.join(GOLD, "ie224"))
    for m in MICS:
        anc.AddMic(m)
    anc.AddSource(SOURCES["noise"], sig)
    x, y = anc.Run(0.02)
    steps = int(np.float32(0.02) * np.float32(32000))
    assert x.shape == (1, 640) and y.shape == (7, steps)
    assert np.array_equal(x[0], sig[:640])
    s = _afx.Sim()
    s.load_kerkythea(os.path.join(GOLD, "ie224", "model.xml"))
    s.add_source(SOURCES["noise"], sig.astype(np.float32))
    for m in MICS:
        s.add_receiver(m)
    s.run(steps)
    assert np.array_equal(y, s.receivers())
    s.close()


def test_ancfield_rejects_bad_arguments():
    from anc_field_py.ancfield import AncField
    with pytest.raises(Exception):
        AncField("tpu")
    with pytest.raises(Exception):
        AncField("gpu", "/nonexistent/model")
