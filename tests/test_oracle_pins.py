"""CPU: pin the oracle before trusting it.

* libopus's own known-answer tests, built from the reference's vendored
  sources by oracle/Makefile (final-range accumulators of test_opus_decode,
  entropy coder, PVQ codeword enumeration, Laplace coder, fixed-point math,
  spreading rotation, LPC inverse prediction gain).
* The resampler restatement (oracle.c orc_taps/orc_resample) against closed
  forms evaluated independently in numpy float64: Kaiser window via
  numpy.kaiser, ideal sinc, unity DC gain, symmetry, and the float64
  convolution of the same input (tolerance 1e-6 abs).  A planted mistake
  (one tap index off) must fail the same check.
* The Python demuxer restatement against hand-built Ogg streams with known
  answers (lacing continuation across pages, end trim, granule mismatch).
"""
import os
import struct
import subprocess
import zlib  # noqa: F401  (CRC is not checked by the demuxer contract)

import numpy as np
import pytest

from tests import oracle_lib as ol
from oracle import demux as odemux

KATS = ["kat_opus_decode", "kat_entropy", "kat_cwrs32", "kat_laplace", "kat_mathops", "kat_rotation",
        "kat_lpc_inv_pred_gain"]


@pytest.mark.parametrize("kat", KATS)
def test_libopus_known_answers(kat):
    exe = os.path.join(ol.REF, kat)
    if not os.path.exists(exe):
        pytest.skip("oracle KAT not built: " + kat)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]


def _np_taps():
    n = np.arange(49) - 24.0
    fc = 1.0 / 3.0
    h = np.where(n == 0, fc, np.sin(np.pi * fc * n) / (np.pi * np.where(n == 0, 1, n)))
    h = h * np.kaiser(49, 8.6)
    return h / h.sum()


def test_taps_closed_form():
    t = ol.taps().astype(np.float64)
    ref = _np_taps()
    assert np.max(np.abs(t - ref)) < 1e-7
    assert abs(t.sum() - 1.0) < 1e-6
    assert np.allclose(t, t[::-1], atol=0)


def _np_resample(x, taps):
    xf = x.astype(np.float64) / 32768.0
    n16 = odemux.out16_len(len(x))
    pad = np.concatenate([np.zeros(24), xf, np.zeros(48)])
    out = np.array([np.dot(taps, pad[3 * i: 3 * i + 49]) for i in range(n16)])
    return out


def test_resampler_against_float64():
    rs = np.random.default_rng(3)
    x = (rs.standard_normal(20000) * 6000).clip(-32768, 32767).astype(np.int16)
    got = ol.resample(x)
    want = _np_resample(x, _np_taps())
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 1e-6
    # planted mistake: taps shifted by one position must fail the same check
    bad = np.roll(_np_taps(), 1)
    assert np.max(np.abs(got - _np_resample(x, bad))) > 1e-3


def test_resampler_length_contract():
    for n48 in [0, 10, 48, 49, 50, 51, 52, 3000]:
        assert len(ol.resample(np.zeros(n48, np.int16))) == max(0, (n48 - 48) // 3)


# ---------------------------------------------------------------- Ogg fixtures
def _page(payload_segments, granule, seq, flags=0, serial=7):
    lace = []
    body = b""
    for seg, complete in payload_segments:
        n = len(seg)
        while n >= 255:
            lace.append(255)
            n -= 255
        if complete:
            lace.append(n)
        body += seg
    hdr = b"OggS" + bytes([0, flags]) + struct.pack("<qII", granule, serial, seq) + b"\0\0\0\0" + bytes([len(lace)])
    return hdr + bytes(lace) + body


def _opus_head(preskip):
    return b"OpusHead" + bytes([1, 1]) + struct.pack("<HIhB", preskip, 48000, 0, 0)


def _pkt(config, nbytes=10):
    return bytes([config << 3]) + bytes(nbytes - 1)


def test_demux_known_answer_and_end_trim():
    # 20 ms SILK packets (config 1 -> 960 samples); 5 audio packets, end trim 100
    pre = 312
    pk = [_pkt(1) for _ in range(5)]
    s = _page([(_opus_head(pre), True)], 0, 0, 2)
    s += _page([(b"OpusTags" + bytes(8), True)], 0, 1)
    s += _page([(p, True) for p in pk[:3]], 3 * 960, 2)
    s += _page([(p, True) for p in pk[3:]], 5 * 960 - 100, 3, 4)
    d = odemux.demux(s)
    assert d.preskip == pre and d.channels == 1 and len(d.packets) == 5
    assert d.granule_checked == 2 and d.granule_match == 2
    assert d.pcm_samples_48k == 5 * 960 - 100 - pre
    assert d.sample_off == [0, 960, 1920, 2880, 3840]


def test_demux_lacing_continuation():
    # a 600-byte packet split across two pages (lacing 255,255 then 90)
    big = _pkt(1, 600)
    s = _page([(_opus_head(0), True)], 0, 0, 2)
    s += _page([(b"OpusTags" + bytes(8), True)], 0, 1)
    s += _page([(_pkt(1), True), (big[:510], False)], 960, 2)
    s += _page([(big[510:], True)], 1920, 3, 4)
    d = odemux.demux(s)
    assert [len(p) for p in d.packets] == [10, 600]
    assert d.packets[1] == big
    assert d.granule_match == d.granule_checked == 2


def test_demux_granule_mismatch_detected():
    s = _page([(_opus_head(0), True)], 0, 0, 2)
    s += _page([(b"OpusTags" + bytes(8), True)], 0, 1)
    s += _page([(_pkt(1), True), (_pkt(1), True)], 1000, 2)   # wrong: should be 1920
    s += _page([(_pkt(1), True)], 2880, 3, 4)
    d = odemux.demux(s)
    assert d.granule_match < d.granule_checked


def test_chains_cover_playable_exactly():
    blob = open(ol.small_corpus("silk_wb", ol.RECIPES["silk_wb"]), "rb").read()
    d, chains = ol.demux_chains(blob)
    assert sum(c[3] for c in chains) == d.pcm_samples_48k
    assert chains[0][2] == d.preskip
    for (first, npk, trim, keep) in chains:
        cap = sum(odemux.pkt_samples48(d.packets[first + k]) for k in range(npk))
        assert cap >= trim + keep


def test_threaded_resampler_matches_serial():
    """The CPU baseline's range-split resampler gives orc_resample's output."""
    rs = np.random.default_rng(21)
    for n in [0, 47, 48, 49, 1000, 123457]:
        x = rs.integers(-32768, 32767, size=n).astype(np.int16)
        a = ol.resample(x)
        b = ol.resample_threads(x, 4, chunk=1000)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["silk_wb_16k_60s", "silk_wb", "celt_fb", "hybrid_fec", "mixed", "celt_multi",
                                  "silk_dtx"])
def test_fast_cpu_build_bit_identical(name):
    """The CPU baseline's libopus build (-O3, SSE4.1/AVX2 intrinsics) decodes
    every chain bit-identically to the oracle build (intrinsics off), with and
    without loss, so its timing is of the same computation."""
    if not os.path.exists(os.path.join(ol.REF, "liboracle_fast.so")):
        pytest.skip("fast CPU build not present")
    if name in ol.BENCH_RECIPES:
        blob = open(ol.bench_corpus(name), "rb").read()
    else:
        blob = open(ol.small_corpus(name, ol.recipe(name)), "rb").read()
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    for lost in (None, ol.loss_mask(len(off), 0.05, seed=7)):
        a = ol.decode_chains(data, off, ln, chains, lost=lost)
        b = ol.decode_chains(data, off, ln, chains, lost=lost, fast=True)
        assert np.array_equal(a["ret"], b["ret"])
        assert np.array_equal(a["nsamp"], b["nsamp"]) and np.array_equal(a["rng"], b["rng"])
        assert np.array_equal(a["pcm"], b["pcm"])
