"""GPU: streams the reference's strict demuxer rejects or mis-reads decode
through the HIP path to the answer fixed by construction.

* chained (two files back to back): int16 48 kHz PCM is the two files' PCM
  back to back, bit for bit; the f32 16 kHz output is the oracle resampler
  over that concatenation (1e-6).
* non-zero start granule under CUOPUS_GRANULE=warn: identical tensor to the
  original file.
* a CRC-damaged page dropped (CUOPUS_CHECK_CRC=1, CUOPUS_GRANULE=warn): the
  remaining packets and chains, decoded through the per-packet gate, equal
  fixed-point libopus on the same packets and chains (every sample, sample
  count and final range).
"""
import numpy as np
import pytest

from tests import ogg_mux
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


def _blob(name):
    return open(ol.small_corpus(name, ol.RECIPES[name]), "rb").read()


@pytest.mark.parametrize("pair", [("silk_wb", "celt_fb"), ("hybrid_fb", "celt_hi")])
def test_chained_decode(cn, pair):
    import torch
    a, b = _blob(pair[0]), _blob(pair[1])
    print("case chained", pair, flush=True)
    sa = cn.decode_ogg_to_s16_tensor(a).cpu().numpy()
    sb = cn.decode_ogg_to_s16_tensor(b).cpu().numpy()
    sab = cn.decode_ogg_to_s16_tensor(a + b).cpu().numpy()
    assert np.array_equal(sab, np.concatenate([sa, sb]))
    out = cn.decode_ogg_to_cuda_tensor(a + b)
    torch.cuda.synchronize()
    want = ol.resample(sab)
    got = out.cpu().numpy()
    assert got.shape == want.shape
    assert float(np.abs(got - want).max()) <= 1e-6


def test_shifted_granules_lenient(cn, monkeypatch):
    import torch
    blob = _blob("celt_fb")
    sh = ogg_mux.remux(blob, granule_offset=3 * 48000)
    print("case lenient start offset", flush=True)
    ref = cn.decode_ogg_to_cuda_tensor(blob)
    monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    assert torch.equal(cn.decode_ogg_to_cuda_tensor(sh), ref)


def test_crc_dropped_page_gate(cn, monkeypatch):
    import torch
    blob = _blob("hybrid_fb")
    pgs = ogg_mux.pages(blob)
    pg = pgs[len(pgs) // 2]
    bad = bytearray(blob)
    bad[pg["off"] + pg["size"] - 5] ^= 0x5A
    bad = bytes(bad)
    monkeypatch.setenv("CUOPUS_CHECK_CRC", "1")
    monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    data, lens = cn.packets(bad)
    off = np.zeros(len(lens), np.uint32)
    off[1:] = np.cumsum(lens[:-1], dtype=np.uint64).astype(np.uint32)
    chains = cn.chain_table(bad)
    print("case crc-dropped page: packets", len(lens), "chains", len(chains), flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, lens, chains)
    ref = ol.decode_chains(data, off, lens, chains)
    assert rc == 0 and not stop.any() and (ref["ret"] >= 0).all()
    assert np.array_equal(nsamp, ref["nsamp"]) and np.array_equal(rng, ref["rng"])
    assert np.array_equal(pcm, ref["pcm"])
    out = cn.decode_ogg_to_cuda_tensor(bad)
    torch.cuda.synchronize()
    want = ol.resample(ref["pcm"])
    assert float(np.abs(out.cpu().numpy() - want).max()) <= 1e-6
