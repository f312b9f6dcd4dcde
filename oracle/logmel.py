"""CPU restatement of the log-mel front-end in float64 (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the product's cuopus_log_mel kernel.
Only tests/ and bench.py's CPU leg import it.

The reference feeds its resident f32 16 kHz tensor to a speech model's
feature front-end (ROADMAP.md:149-156) but ships none, so the front-end is
defined here and pinned independently (tests/test_logmel_oracle.py): the
filter bank against transformers.audio_utils.mel_filter_bank (Slaney scale
and norm), the whole pipeline against transformers.audio_utils.spectrogram,
and closed forms (Slaney mel scale at its knee, a windowed sinusoid at a bin
centre, Parseval on the zero-padded frame).

Definition: frame t covers x[160 t : 160 t + 400]; periodic Hann window of
length 400; zero-padded 512-point FFT; power |X_k|^2 for k = 0..256;
n_mels triangular filters on the Slaney mel scale between 0 and 8000 Hz with
area normalisation 2 / (f_hi - f_lo); out = ln(max(mel, 1e-10)).
"""
import numpy as np

N_FFT, WIN, HOP, SR = 512, 400, 160, 16000
FLOOR = 1e-10


def hz_to_mel(f):
    f = np.asarray(f, dtype=np.float64)
    f_sp, min_log_hz = 200.0 / 3.0, 1000.0
    min_log_mel = min_log_hz / f_sp
    logstep = np.log(6.4) / 27.0
    return np.where(f < min_log_hz, f / f_sp, min_log_mel + np.log(np.maximum(f, 1e-300) / min_log_hz) / logstep)


def mel_to_hz(m):
    m = np.asarray(m, dtype=np.float64)
    f_sp, min_log_hz = 200.0 / 3.0, 1000.0
    min_log_mel = min_log_hz / f_sp
    logstep = np.log(6.4) / 27.0
    return np.where(m < min_log_mel, f_sp * m, min_log_hz * np.exp(logstep * (m - min_log_mel)))


def mel_filters(n_mels=80, fmin=0.0, fmax=SR / 2):
    """(n_mels, 257) float64 filter weights."""
    edges = mel_to_hz(np.linspace(hz_to_mel(fmin), hz_to_mel(fmax), n_mels + 2))
    freqs = np.arange(N_FFT // 2 + 1) * SR / N_FFT
    lo, mid, hi = edges[:-2, None], edges[1:-1, None], edges[2:, None]
    tri = np.maximum(0.0, np.minimum((freqs - lo) / (mid - lo), (hi - freqs) / (hi - mid)))
    return tri * (2.0 / (hi - lo))


def window():
    j = np.arange(WIN)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * j / WIN)


def n_frames(n):
    return 0 if n < WIN else 1 + (n - WIN) // HOP


def power_spectrum(x):
    """(frames, 257) float64 power spectra of the framed, windowed signal."""
    x = np.asarray(x, dtype=np.float64)
    nf = n_frames(x.size)
    if nf == 0:
        return np.zeros((0, N_FFT // 2 + 1))
    idx = np.arange(nf)[:, None] * HOP + np.arange(WIN)[None, :]
    frames = x[idx] * window()[None, :]
    spec = np.fft.rfft(frames, n=N_FFT, axis=1)
    return spec.real ** 2 + spec.imag ** 2


def log_mel(x, n_mels=80):
    """(frames, n_mels) float64 log-mel features."""
    p = power_spectrum(x)
    mel = p @ mel_filters(n_mels).T
    return np.log(np.maximum(mel, FLOOR))
