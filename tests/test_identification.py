"""Identification helpers (anc_field_py/ancutil.py of the reference:
GetLpFir :61-79, impulse :49-55, bandnoise :39-47, EstimateIr :24-37) and
the impulse method that replaces the LMS estimate
(anc_field_py/identification.py:36 fits EstimateIr to the input and output
decimated by 4)."""
import numpy as np
from scipy import signal

from anc_field_py import ancutil, surrogate

FS = 32000.0


def test_estimate_ir_closed_form_on_impulse_input():
    """With a unit impulse as input each tap sees its own output sample once
    per epoch: w_j <- w_j + 2 mu (y_j - w_j), so after E epochs
    w_j = y_j (1 - (1 - 2 mu)^E)."""
    y = np.array([0.7, -0.3, 0.25, 0.1])
    x = np.array([1.0, 0.0, 0.0, 0.0])
    mu = 0.1
    for epochs in (1, 16):
        w = ancutil.EstimateIr(x, y, 4, mu=mu, epochs=epochs)
        assert np.allclose(w, y * (1 - (1 - 2 * mu) ** epochs), rtol=0, atol=1e-15)


def test_estimate_ir_identifies_an_fir():
    rng = np.random.default_rng(3)
    h = np.array([0.5, -0.25, 0.125, 0.3, 0.0, -0.1, 0.05, 0.02])
    x = 2 * rng.random(4000) - 1
    y = signal.lfilter(h, 1, x)
    w = ancutil.EstimateIr(x, y, 8, mu=2e-3)
    assert np.abs(w - h).max() < 1e-3


def test_lowpass_fir_meets_its_kaiser_design():
    """GetLpFir: 80 dB stop band from cutoff + transition, unity pass band."""
    b = ancutil.GetLpFir(2000.0, 100.0, FS)
    f, H = signal.freqz(b, worN=1 << 16, fs=FS)
    assert np.abs(20 * np.log10(np.abs(H[f >= 2100.0]) + 1e-300)).min() > 78
    assert np.abs(np.abs(H[f <= 1900.0]) - 1).max() < 1e-3
    # impulse: that FIR, delayed by one sample
    x = ancutil.impulse(2000.0, len(b) + 10, FS)
    assert x[0] == 0 and np.allclose(x[1:len(b) + 1], b, rtol=0, atol=1e-15)


def test_bandnoise_stays_in_its_band():
    for bands in ([500.0, 1500.0], [0.0, 2000.0]):
        x = ancutil.bandnoise(bands, 1 << 17, FS, rng=np.random.default_rng(4))
        # past the filters' start-up transient (10 000 taps for the band pass)
        f, p = signal.welch(x[20000:], FS, nperseg=4096)
        inband = p[(f > bands[0] + 100) & (f < bands[1] - 100)]
        stop = p[f > bands[1] + 300]
        assert inband.min() > 1e5 * stop.max(), (bands, inband.min() / stop.max())
        assert np.ptp(10 * np.log10(inband)) < 6.0  # flat within the Welch estimate's own spread
        if bands[0] > 0:
            # (the lowest two bins hold Welch's per-segment mean removal)
            assert inband.min() > 1e5 * p[(f > 20) & (f < bands[0] - 300)].max()
    a = ancutil.bandnoise([0.0, 2000.0], 1000, FS, rng=np.random.default_rng(5))
    b = ancutil.bandnoise([0.0, 2000.0], 1000, FS, rng=np.random.default_rng(5))
    assert np.array_equal(a, b)


def test_impulse_method_matches_reference_lms_identification():
    """The reference identifies a secondary path by driving it with band
    noise and fitting EstimateIr to input and output decimated by 4.  The
    impulse method (one run with a unit impulse, low-pass, decimate) gives the
    same filter in the excited band: impulse_response to 0.5 %, and the
    surrogate's lowpass_decimate to the same after the factor 4."""
    rng = np.random.default_rng(1)
    n = 600
    h = rng.standard_normal(n) * np.exp(-np.arange(n) / 120.0)
    h = signal.lfilter(ancutil.GetLpFir(1500.0, 100.0, FS), 1, np.concatenate([h, np.zeros(1200)]))[:1600]
    x = ancutil.bandnoise([0.0, 2000.0], 64000, FS, rng=np.random.default_rng(2))
    y = signal.lfilter(h, 1, x)
    w = ancutil.EstimateIr(x[0::4], y[0::4], 400, mu=3e-3, epochs=4)
    lp = signal.firwin(101, 2000.0 / 4000.0)

    def band(v):
        return signal.filtfilt(lp, [1.0], v)

    ours = ancutil.impulse_response(h, FS, cutoff=2000.0, decimate=4)[:400]
    assert np.linalg.norm(band(w) - band(ours)) < 5e-3 * np.linalg.norm(band(w))
    dec = surrogate.lowpass_decimate(h[None], 1 / FS, 400, 4, 2000.0)[0]
    assert np.linalg.norm(band(w) - 4 * band(dec)) < 5e-3 * np.linalg.norm(band(w))
