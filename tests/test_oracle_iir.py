"""Pins of the oracle's frequency-dependent boundary (ncoef >= 2).

The locally reacting surface keeps, per boundary cell and material, two
rings of the filter's input x and output y (lrs_1.cl:71-76,122-124,
lrs_2.cl:79-85, lrs_3.cl:70,83-92,176-186).  The ring index is decremented
once per step (simulation.hpp:365), so slot (ix_buf + i) % N holds step n-i,
and the update writes

    x[n] = a0/(l b0) (p+ - p-) - g/b0,   y[n] = (b0 x[n] + g)/a0,
    g    = sum_{i>=1} b_i x[n-i] - a_i y[n-i].

Two consequences fix the recursion without restating it:

  - y = scipy.signal.lfilter(b, a, x) for every cell and material, and
    y[n] = (p+ - p-)/l at the cell (l = lambda in physical mode);
  - at normal incidence in a rigid duct the wall reflects a plane wave with
    the scheme's exact discrete reflection coefficient
        r(w) = (sin th - beta(w) sin(w dt)/lambda) / (sin th + beta(w) sin(w dt)/lambda),
    beta = A(e^{iw dt})/B(e^{iw dt}), cos th = 1 - (1 - cos w dt)/lambda^2,
    which tends to (1 - beta)/(1 + beta) as dx -> 0.

Mistakes planted in a CPU copy of the oracle (ring read backwards, a and b
swapped for i >= 1, ring index incremented) must fail both.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.signal

from anc_field_py import rooms
import orc

C0, DX = 343.0, 0.02
DT = rooms.courant_dt(DX)
# lambda as the f64 oracle forms it from the single-precision c, dx, dt
LAM = float(np.float64(np.float32(DT)) * np.float64(np.float32(C0)) / np.float64(np.float32(DX)))


def _iir3(rng):
    """Third-order passive admittance (four coefficients): a second-order
    section from rooms.iir_material times a first-order one."""
    w = np.linspace(0, math.pi, 512)
    z1 = np.exp(-1j * w)
    for _ in range(1000):
        a2, b2 = rooms.iir_material(rng)
        q, s = rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6)
        a = np.convolve(a2.astype(np.float64), [1.0, -q])
        b = np.convolve(b2.astype(np.float64), [1.0, -s])
        beta = np.polyval(a[::-1], z1) / np.polyval(b[::-1], z1)
        if beta.real.min() >= 0:
            return a.astype(np.float32), b.astype(np.float32)
    raise RuntimeError("no passive filter drawn")


def _materials(ncoef, nmat, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    mats = [rooms.iir_material(rng) if ncoef == 3 else _iir3(rng) for _ in range(nmat)]
    return np.stack([m[0] for m in mats]), np.stack([m[1] for m in mats])


def _room(ncoef, seed=11):
    """A furnished 26x22x20 room: walls of three materials plus boxes, so
    faces, edges and corners of every kind carry boundary cells."""
    a, b = _materials(ncoef, 3, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    v = rooms.shoebox(26, 22, 20, [0, 1, 2, 0, 1, 2])
    v = rooms.repair(rooms.add_boxes(v, rng, 5, 2, 6, 3))
    return v, a, b


# ------------------------------------------------------------- ring histories
def _ring_histories(v, a, b, compat, steps, lib=None):
    """Step the f64 oracle one step at a time.  Returns the boundary list,
    per-step written slots and the histories x[e, k, t], y[e, k, t] read
    from the slot each step wrote, plus (p+ - p-) at every boundary cell."""
    sig = rooms.ricker(steps, DT, 1500.0)
    o = orc.Oracle(v, a, b, C0, DX, DT, compat, [[9, 8, 7]], sig[None], [], precision="f64", lib=lib)
    pos, grp = orc.classify(v)
    nb, N = len(grp), a.shape[1]
    assert o.nbound() == nb
    cells = (pos[:, 2], pos[:, 1], pos[:, 0])
    x = np.zeros((nb, 3, steps))
    y = np.zeros((nb, 3, steps))
    dp = np.zeros((nb, steps))
    slots = []
    prev_in, prev_out = o.rings()
    for t in range(steps):
        o.run(1)
        rin, rout = o.rings()
        w = o.ring_index()
        slots.append(w)
        # the step writes slot w of each material's ring and nothing else
        # (compat corners take the reference's ring stride ncoef/2, so they
        # are left out there)
        keep = np.ones(3 * N, bool)
        keep[[w + k * N for k in range(3)]] = False
        chk = grp < 18 if compat else np.ones(nb, bool)
        assert np.array_equal(rin[chk][:, keep], prev_in[chk][:, keep])
        assert np.array_equal(rout[chk][:, keep], prev_out[chk][:, keep])
        for k in range(3):
            x[:, k, t] = rin[:, w + k * N]
            y[:, k, t] = rout[:, w + k * N]
        dp[:, t] = o.field(0)[cells] - o.field(2)[cells]
        prev_in, prev_out = rin, rout
    o.close()
    return pos, grp, np.array(slots), x, y, dp


def _lfilter_residual(pos, grp, a, b, x, y):
    """max |y - lfilter(b, a, x)| / max |y| over every (cell, material)."""
    K = np.where(grp <= 5, 1, np.where(grp <= 17, 2, 3))
    worst, scale = 0.0, 0.0
    for e in range(len(grp)):
        for k in range(K[e]):
            m = pos[e, 3 + k]
            want = scipy.signal.lfilter(b[m].astype(np.float64), a[m].astype(np.float64), x[e, k])
            worst = max(worst, float(np.abs(y[e, k] - want).max()))
            scale = max(scale, float(np.abs(y[e, k]).max()))
    return worst / scale


@pytest.mark.parametrize("ncoef", [3, 4])
def test_ring_histories_follow_lfilter_physical(ncoef):
    v, a, b = _room(ncoef)
    steps = 90
    pos, grp, slots, x, y, dp = _ring_histories(v, a, b, False, steps)
    # every class of cell is present and has been reached by the wave
    for lo, hi, least in ((0, 5, 100), (6, 17, 40), (18, 25, 8)):
        sel = (grp >= lo) & (grp <= hi)
        assert sel.sum() >= least
        assert np.abs(x[sel]).max() > 0
    # the ring index is decremented once per step, from slot 0 (simulation.hpp:229, :365)
    assert np.array_equal(slots, (-(np.arange(steps) + 1)) % ncoef)
    assert _lfilter_residual(pos, grp, a, b, x, y) < 1e-12
    # y[n] = (p+ - p-)/lambda for every material of the cell
    K = np.where(grp <= 5, 1, np.where(grp <= 17, 2, 3))
    for k in range(3):
        sel = K > k
        assert np.allclose(y[sel, k], dp[sel] / LAM, rtol=0, atol=1e-12 * np.abs(dp).max() / LAM)


def test_ring_histories_follow_lfilter_compat_faces_and_edges():
    """Compat mode feeds lambda^2 as the boundary's l (simulation.hpp:368-369);
    faces and edges keep the same recursion, with y = (p+ - p-)/lambda^2."""
    v, a, b = _room(3, seed=12)
    steps = 80
    pos, grp, slots, x, y, dp = _ring_histories(v, a, b, True, steps)
    sel = grp < 18
    assert _lfilter_residual(pos[sel], grp[sel], a, b, x[sel], y[sel]) < 1e-12
    l2 = LAM * LAM
    f = grp <= 5
    assert np.allclose(y[f, 0], dp[f] / l2, rtol=1e-9, atol=1e-12 * np.abs(dp).max() / l2)


# ------------------------------------------------------------- duct spectrum
def _duct_run(a_end, b_end, n, steps, lib=None):
    """Plane wave in a 2x2-cell duct with rigid side walls (admittance
    1e-12), wall of the given filter on the x+ end (x = n-1).  Soft source
    over the cross section at x = 400, receiver at x = 500; f64."""
    N = len(a_end)
    v = np.full((4, 4, n), -1, np.int32)
    v[0, :, :] = v[3, :, :] = 1
    v[:, 0, :] = v[:, 3, :] = 1
    v[:, :, n - 1] = 0
    side_a = np.zeros(N, np.float32)
    side_a[0] = 1e-12
    side_b = np.zeros(N, np.float32)
    side_b[0] = 1.0
    a = np.stack([a_end, side_a]).astype(np.float32)
    b = np.stack([b_end, side_b]).astype(np.float32)
    sig = rooms.ricker(steps, DT, 1000.0, t0=1.5e-3)
    src = [[400, yy, zz] for yy in (1, 2) for zz in (1, 2)]
    o = orc.Oracle(v, a, b, C0, DX, DT, False, src, np.stack([sig] * 4), [[500, 1, 1]], precision="f64", lib=lib)
    tr = o.run(steps)[0]
    o.close()
    return tr


def _reflection_measured(a_end, b_end, lib=None):
    """(frequencies, measured incident->reflected transfer at the receiver).

    The reflection is the difference of the run with the wall at x = 799 and
    a run whose wall is too far to answer in time; the incident pulse is the
    far run before the echo of the x = 0 edge reaches the receiver."""
    steps = 2500
    near = _duct_run(a_end, b_end, 800, steps, lib)
    far = _duct_run(a_end, b_end, 2000, steps, lib)
    nfft = 8192
    inc = np.fft.rfft(far[:1400], nfft)
    ref = np.fft.rfft(near - far, nfft)
    f = np.fft.rfftfreq(nfft, DT)
    return f, ref / inc


def _reflection_theory(a_end, b_end, f, d):
    w = 2 * np.pi * f * DT
    z1 = np.exp(-1j * w)
    a = a_end.astype(np.float64)
    b = b_end.astype(np.float64)
    beta = np.polyval(a[::-1], z1) / np.polyval(b[::-1], z1)
    th = np.arccos(1.0 - (1.0 - np.cos(w)) / LAM ** 2)
    q = beta * np.sin(w) / (LAM * np.sin(th))
    r = (1 - q) / (1 + q)
    return r * np.exp(-2j * th * d), r, beta


def _reflection_error(a_end, b_end, lib=None):
    f, h = _reflection_measured(a_end, b_end, lib)
    band = (f >= 200) & (f <= 1500)
    want, r, beta = _reflection_theory(a_end, b_end, f[band], 798 - 500)
    return float(np.abs(h[band] - want).max()), r, beta


def _resonant_wall():
    """A passive second-order wall whose impedance resonance (pole angle
    0.2 rad, about 960 Hz at this dt) sits inside the measured band."""
    th, r, rho, R = 0.2, 0.9, 0.45, 0.8
    a = np.array([1.0, -2 * r * math.cos(th), r * r])
    b = np.array([1.0, -2 * rho * math.cos(th), rho * rho])
    a *= (1 - R) / (1 + R) * b.sum() / a.sum()
    w = np.linspace(0, math.pi, 2048)
    z1 = np.exp(-1j * w)
    assert (np.polyval(a[::-1], z1) / np.polyval(b[::-1], z1)).real.min() >= 0
    return a.astype(np.float32), b.astype(np.float32)


def _config_wall():
    """A wall filter as the seeded room generator draws them."""
    return rooms.iir_material(np.random.Generator(np.random.PCG64(21)))


@pytest.mark.parametrize("wall", [_config_wall, _resonant_wall], ids=["generator", "resonant"])
def test_duct_reflection_spectrum_second_order_wall(wall):
    a_end, b_end = wall()
    err, r, beta = _reflection_error(a_end, b_end)
    # the measured transfer equals the scheme's discrete reflection to
    # rounding (f64); the wall's frequency dependence over the band is four
    # orders of magnitude above that, so a frequency-blind or misrouted
    # filter cannot pass
    assert np.ptp(np.abs(r)) > 0.015
    assert np.ptp(np.angle(beta)) > 0.15
    assert err < 1e-6, err
    # in the continuum limit r -> (1 - beta)/(1 + beta): the scheme's own
    # reflection stays within its dispersion error of the textbook value
    textbook = (1 - beta) / (1 + beta)
    assert np.abs(r - textbook).max() < 0.02


# ------------------------------------------------------------- planted mistakes
PLANTS = {
    "ring read backwards": ("const unsigned slot = (ixb + ii) % sf + (unsigned)k * sf;",
                            "const unsigned slot = (ixb + sf - ii) % sf + (unsigned)k * sf;"),
    "a and b swapped for i >= 1": ("gn += bk[ii] * rin[slot] - ak[ii] * rout[slot];",
                                   "gn += ak[ii] * rin[slot] - bk[ii] * rout[slot];"),
    "ring index incremented": ("if (s->ix_buf == 0) s->ix_buf = N - 1; else s->ix_buf -= 1;",
                               "s->ix_buf = (s->ix_buf + 1) % N;"),
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    """Oracle libraries built from copies of oracle/ with one mistake each."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = os.path.join(orc.ROOT, "oracle")
    libs = {}
    for i, (name, (old, new)) in enumerate(PLANTS.items()):
        d = tmp_path_factory.mktemp(f"plant{i}")
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
def test_planted_mistake_fails_lfilter_identity(planted, name):
    v, a, b = _room(3)
    pos, grp, slots, x, y, dp = _ring_histories(v, a, b, False, 60, lib=planted[name])
    bad_slots = not np.array_equal(slots, (-(np.arange(60) + 1)) % 3)
    assert bad_slots or _lfilter_residual(pos, grp, a, b, x, y) > 1e-3


@pytest.mark.parametrize("name", list(PLANTS))
def test_planted_mistake_fails_duct_spectrum(planted, name):
    for wall in (_config_wall, _resonant_wall):
        a_end, b_end = wall()
        err, _, _ = _reflection_error(a_end, b_end, lib=planted[name])
        assert err > 1e-3, (wall.__name__, err)
