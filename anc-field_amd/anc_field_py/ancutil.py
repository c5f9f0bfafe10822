"""Signal helpers of the identification workflow.

Mirrors anc_field_py/ancutil.py of the reference (paths relative to its
root): GetLpFir (:61-79), impulse (:49-55), bandnoise (:39-47), EstimateIr
(:24-37), plus the impulse-response post-processing that replaces the LMS
estimate when the excitation is a known impulse: low-pass + decimation
(anc_field_py/identification.py:36 takes every 4th sample).
"""
import numpy as np
from scipy import signal


def GetLpFir(cutoff, transition, fs):
    """Kaiser-window low-pass FIR: 80 dB stop band, `transition` Hz wide."""
    nyq_rate = fs / 2.0
    width = transition / nyq_rate
    n, beta = signal.kaiserord(80.0, width)
    return signal.firwin(n, cutoff / nyq_rate, window=("kaiser", beta))


def impulse(bw, length, fs):
    """Unit impulse at sample 1 through the band-limiting FIR of GetLpFir(bw, 100, fs)."""
    x = np.zeros(length)
    x[1] = 1
    return signal.lfilter(GetLpFir(bw, 100, fs), 1, x)


def bandnoise(bands, length, fs, rng=None):
    """Uniform white noise in [-1, 1) filtered to `bands` = [lo, hi] Hz.

    rng: numpy Generator for a reproducible draw (the reference draws from
    the global numpy state)."""
    rng = np.random.default_rng() if rng is None else rng
    z = 2 * rng.random(length) - 1
    if len(bands) > 1 and bands[0] != 0:
        b = signal.firwin(10000, [c / (fs / 2) for c in bands], pass_zero=False)
    else:
        b = GetLpFir(bands[1], 100, fs)
    return signal.lfilter(b, 1, z)


def EstimateIr(x, y, n, mu=1e-4, epochs=16):
    """LMS estimate of an n-tap FIR from input x to output y."""
    w = np.zeros(n)
    steps = min(np.size(x, 0), np.size(y, 0))
    for _ in range(epochs):
        buf = np.zeros(n)
        for k in range(steps):
            buf = np.concatenate((np.array([x[k]]), buf[:-1]))
            e = y[k] - np.dot(w, buf)
            w = w + 2 * mu * e * buf
    return w


def impulse_response(trace, fs, cutoff=2000.0, decimate=4, ntaps=255):
    """The secondary path at fs/decimate from the receiver trace of a unit
    impulse: band-limited at `cutoff` Hz with a zero-phase FIR, every
    `decimate`-th sample kept, times `decimate`.

    The factor makes it the filter the reference's identification estimates:
    EstimateIr fitted to the input and output decimated by 4
    (anc_field_py/identification.py:36) is the decimated-rate system, whose
    taps are `decimate` times the decimated impulse response of the full-rate
    one, for band-limited inputs."""
    h = signal.firwin(ntaps, cutoff / (fs / 2.0))
    y = signal.filtfilt(h, [1.0], np.asarray(trace, np.float64), axis=-1)
    return (decimate * y[..., ::decimate]).astype(np.float32)
