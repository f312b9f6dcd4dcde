"""CPU: pins of the log-mel oracle (oracle/logmel.py) and of the product's
host-side table construction.

The reference ships no front-end (ROADMAP.md:149-156 names the consumer), so
the oracle is checked against an independent implementation
(transformers.audio_utils: mel_filter_bank with Slaney scale and norm,
spectrogram with the same framing) and against closed forms: the Slaney mel
scale's knee (15 mel at 1 kHz, linear below), a windowed sinusoid at a bin
centre (|X_k0| = A/2 * sum(window)), Parseval on the zero-padded frame, and
a planted mistake (a symmetric instead of periodic window) that the
independent check catches.
"""
import numpy as np
import pytest

from oracle import logmel as lm

tf_audio = pytest.importorskip("transformers.audio_utils")


def test_slaney_scale_closed_form():
    assert lm.hz_to_mel(1000.0) == pytest.approx(15.0, abs=1e-12)
    assert lm.hz_to_mel(200.0) == pytest.approx(3.0, abs=1e-12)  # linear below 1 kHz: f / (200/3)
    # one octave above 1 kHz adds 27 * ln 2 / ln 6.4 mel
    assert lm.hz_to_mel(2000.0) - lm.hz_to_mel(1000.0) == pytest.approx(27 * np.log(2) / np.log(6.4), abs=1e-12)
    f = np.linspace(0, 8000, 97)
    assert np.allclose(lm.mel_to_hz(lm.hz_to_mel(f)), f, atol=1e-9)


@pytest.mark.parametrize("n_mels", [40, 64, 80, 128])
def test_filters_match_transformers(n_mels):
    fb = tf_audio.mel_filter_bank(num_frequency_bins=257, num_mel_filters=n_mels, min_frequency=0.0,
                                  max_frequency=8000.0, sampling_rate=16000, norm="slaney", mel_scale="slaney")
    ours = lm.mel_filters(n_mels)
    assert ours.shape == (n_mels, 257)
    assert np.abs(fb.T - ours).max() < 1e-12
    # Slaney area normalisation: each triangle's area over frequency is 1
    edges = lm.mel_to_hz(np.linspace(lm.hz_to_mel(0), lm.hz_to_mel(8000), n_mels + 2))
    peak = 2.0 / (edges[2:] - edges[:-2])
    assert np.all(ours.max(axis=1) <= peak + 1e-12)


def test_pipeline_matches_transformers():
    x = np.random.default_rng(5).standard_normal(16000 * 3) * 0.05
    fb = tf_audio.mel_filter_bank(num_frequency_bins=257, num_mel_filters=80, min_frequency=0.0,
                                  max_frequency=8000.0, sampling_rate=16000, norm="slaney", mel_scale="slaney")
    w = tf_audio.window_function(400, "hann", periodic=True)
    ref = tf_audio.spectrogram(x, w, frame_length=400, hop_length=160, fft_length=512, power=2.0, center=False,
                               mel_filters=fb, mel_floor=1e-10, log_mel=None, dtype=np.float64)
    got = lm.log_mel(x)
    assert got.shape == (lm.n_frames(x.size), 80) == ref.T.shape
    # transformers frames the signal in float32: agreement to float32 round-off
    assert np.abs(np.log(np.maximum(ref.T, 1e-10)) - got).max() < 1e-6
    # planted mistake: a symmetric Hann window is caught by the same comparison
    wrong = np.hanning(400)
    idx = np.arange(got.shape[0])[:, None] * 160 + np.arange(400)[None, :]
    spec = np.fft.rfft(x[idx] * wrong, n=512, axis=1)
    bad = np.log(np.maximum((np.abs(spec) ** 2) @ lm.mel_filters(80).T, 1e-10))
    assert np.abs(np.log(np.maximum(ref.T, 1e-10)) - bad).max() > 1e-4


def test_sinusoid_at_bin_centre():
    k0, amp = 40, 0.3
    n = np.arange(4000)
    x = amp * np.cos(2 * np.pi * k0 * n / 512)
    p = lm.power_spectrum(x)
    expect = (amp / 2 * lm.window().sum()) ** 2  # the positive-frequency half of the cosine
    # the negative-frequency image leaks a little into bin k0
    assert np.allclose(p[:, k0], expect, rtol=2e-3)
    # outside the main lobe (2 bins of a 400-point transform = 2.56 bins of 512) the power is far lower
    assert p[:, k0].min() > 1e4 * np.delete(p, range(k0 - 3, k0 + 4), axis=1).max()


def test_parseval_and_frame_count():
    x = np.random.default_rng(9).standard_normal(1000)
    assert lm.n_frames(399) == 0 and lm.n_frames(400) == 1 and lm.n_frames(559) == 1 and lm.n_frames(560) == 2
    p = lm.power_spectrum(x)
    assert p.shape == (lm.n_frames(1000), 257)
    for t in range(p.shape[0]):
        y = x[160 * t:160 * t + 400] * lm.window()
        full = p[t, 0] + 2 * p[t, 1:256].sum() + p[t, 256]
        assert full == pytest.approx(512 * np.sum(y * y), rel=1e-10)


def test_silence_floor():
    out = lm.log_mel(np.zeros(2000))
    assert np.all(out == np.log(1e-10))


@pytest.mark.parametrize("n_mels", [40, 80, 128])
def test_product_filters_are_the_oracle_filters_in_float32(n_mels):
    """The table the kernel applies (built in C++ double, stored float32) is
    the oracle's filter bank rounded to float32."""
    from cuopus_amd import cuopus_native as cn
    got = cn.mel_filters(n_mels)
    want = lm.mel_filters(n_mels).astype(np.float32)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
