"""The C-ABI library loads and exports every function include/*.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(afx_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    for must in ("afx_create", "afx_run", "afx_get_receivers", "afx_set_pml", "afx_add_source", "afx_add_receiver",
                 "afx_load_kerkythea", "afx_set_min_size", "afx_device_count", "afx_destroy"):
        assert must in names


def test_library_exports_all_declared():
    from anc_field_py import _afx
    lib = ctypes.CDLL(_afx.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_afx.EXPORTED) == set(_declared())


def test_config_default_matches_reference_constants():
    from anc_field_py import _afx
    cfg = _afx.default_config()
    assert cfg.c == 340.0 and abs(cfg.dx - 0.02) < 1e-9 and abs(cfg.dt - 3.125e-5) < 1e-11
    assert cfg.flags & _afx.COMPAT and cfg.precision == _afx.FP32


def test_create_without_device_fails_cleanly():
    from anc_field_py import _afx
    try:
        n = _afx.device_count()
    except _afx.AfxError:
        n = 0
    if n == 0:
        try:
            _afx.Sim()
        except _afx.AfxError as e:
            assert e.code in (_afx.AFX_E_HIP,)
        else:
            raise AssertionError("afx_create succeeded without a device")


def test_bad_config_rejected():
    from anc_field_py import _afx
    import pytest
    with pytest.raises(_afx.AfxError) as ei:
        _afx.Sim(c=340.0, dx=0.02, dt=1e-4)  # Courant number 1.7
    assert ei.value.code in (_afx.AFX_E_ARG, _afx.AFX_E_HIP)
