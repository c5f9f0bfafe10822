"""Pins of the oracle's compat PML (pml.cl:154-265) with a non-zero kappa.

In a region where the three per-axis kappas are constant and the pressure
is spatially uniform, the Laplacian and the divergence of phi vanish and
pml.cl's pressure update reduces to a discretisation of

    p'' + s1 p' + s2 p + s3 psi = 0,   psi' = p,
    s1 = kx + ky + kz,  s2 = kx ky + ky kz + kz kx,  s3 = kx ky kz,

whose characteristic polynomial factors as (s + kx)(s + ky)(s + kz).  The
scheme is second-order accurate, so the trace of a cell far from the grid
edges obeys a linear three-term recurrence whose roots converge to
exp(-k dt) at third order in k dt.  The corner term kx ky kz psi
(pml.cl:197) is the only place psi enters the pressure, so a dropped, mis-
signed or mis-averaged corner term, or a psi that does not accumulate
(pml.cl:185), moves the roots far outside that error.

The phi updates (pml.cl:231-263, with the reference's Dxp/Dyp/Dzp and
ksi_avg(ksi_x, y) quirks) do not enter a uniform field and stay pinned only
by reading against pml.cl; with the shipped kappa == 0 they keep phi at zero
(test_compat_pml_zero_profile_is_leapfrog).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc

DT = 3.125e-05
N = 64


def _uniform_trace(kappa, steps=26, lib=None):
    """Centre-cell trace of a 64^3 grid started at rest from p = 1 on every
    interior cell, with constant per-axis kappa (x, y, z) everywhere."""
    o = orc.Oracle(None, None, None, 340.0, 0.02, DT, True, [], None, [[32, 32, 32], [29, 30, 34]],
                   precision="f64", dims=(N, N, N), pml_mode=1, pml_width=N // 2,
                   compat_kappa=[np.full(N, k) for k in kappa], lib=lib)
    u = np.zeros((N, N, N))
    u[1:-1, 1:-1, 1:-1] = 1.0
    o.set_field(0, u)
    o.set_field(1, u)
    tr = o.run(steps)
    o.close()
    return tr


def _recurrence_roots(trace, lo=4):
    """Roots of the recurrence p[n+1] = c0 p[n] + c1 p[n-1] + c2 p[n-2]
    fitted to the trace (p^-1 = p^0 = 1 prepended), and the fit residual.
    The first rows are skipped: the reference zeroes phi and psi after the
    first step of each run call (simulation.hpp:411-430)."""
    p = np.concatenate([[1.0, 1.0], trace])
    rows = np.arange(lo, len(p) - 1)
    A = np.stack([p[rows], p[rows - 1], p[rows - 2]], 1)
    c, *_ = np.linalg.lstsq(A, p[rows + 1], rcond=None)
    resid = float(np.abs(A @ c - p[rows + 1]).max())
    return np.sort(np.roots([1.0, -c[0], -c[1], -c[2]])), resid


KAPPAS = [(800.0, 1200.0, 1600.0), (1600.0, 2400.0, 3200.0)]


def _root_errors(lib=None):
    errs = []
    for kap in KAPPAS:
        tr = _uniform_trace(kap, lib=lib)
        # the centre stays in the uniform region for the whole trace
        assert np.array_equal(tr[0], tr[1])
        roots, resid = _recurrence_roots(tr[0])
        assert resid < 1e-12 * np.abs(tr[0]).max()
        want = np.sort(np.exp(-np.array(kap) * DT))
        errs.append(float(np.abs(roots - want).max()))
    return errs


def _rates_hold(errs):
    """Roots within 3 (k dt)^3 of exp(-k dt) at both scales, and third-order
    convergence: halving kappa divides the error by about 8."""
    within = all(e < 3.0 * (max(kap) * DT) ** 3 for kap, e in zip(KAPPAS, errs))
    return within and errs[1] / errs[0] > 5.0


def test_uniform_field_decays_with_the_pml_rates():
    errs = _root_errors()
    assert _rates_hold(errs), errs


PLANTS = {
    "corner term dropped": ("REAL t3 = kx[x] * ky[y] * kz[z] * (REAL)0.5 * (psiph[i] + psimh[i]);",
                            "REAL t3 = 0;"),
    "corner term sign flipped": ("REAL num = t1 + t2 - t3 - t4 - t5 + t6;",
                                 "REAL num = t1 + t2 + t3 - t4 - t5 + t6;"),
    "corner term with psi^{n+1/2} only": ("REAL t3 = kx[x] * ky[y] * kz[z] * (REAL)0.5 * (psiph[i] + psimh[i]);",
                                          "REAL t3 = kx[x] * ky[y] * kz[z] * psiph[i];"),
    "psi not accumulated": ("psiph[i] = dt * pm1[i] + psimh[i];", "psiph[i] = dt * pm1[i];"),
    "kappa pair product wrong": ("REAL t5 = (kx[x] * ky[y] + ky[y] * kz[z] + kz[z] * kx[x]) * pm1[i];",
                                 "REAL t5 = (kx[x] * ky[y] + ky[y] * kz[z] + kz[z] * ky[y]) * pm1[i];"),
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    """Oracle libraries built from copies of oracle/ with one mistake each."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = os.path.join(orc.ROOT, "oracle")
    libs = {}
    for i, (name, (old, new)) in enumerate(PLANTS.items()):
        d = tmp_path_factory.mktemp(f"pml_plant{i}")
        for fn in ("orc.c", "orc.h", "orc_impl.h"):
            shutil.copy(os.path.join(src, fn), d / fn)
        text = (d / "orc_impl.h").read_text()
        assert text.count(old) == 1, name
        (d / "orc_impl.h").write_text(text.replace(old, new))
        so = d / "liborc.so"
        subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-std=c99",
                        "-o", str(so), str(d / "orc.c")], check=True)
        libs[name] = orc.load_lib(str(so))
    return libs


@pytest.mark.parametrize("name", list(PLANTS))
def test_planted_pml_mistake_moves_the_rates(planted, name):
    errs = _root_errors(lib=planted[name])
    assert not _rates_hold(errs), errs
