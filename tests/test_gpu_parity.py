"""GPU parity: the HIP decoder through the C ABI vs the fixed-point libopus
oracle, per packet (sample count, final range) and per sample.

Bar: bit-exact int16 PCM, sample counts and final ranges; f32 16 kHz output
within 1e-6 abs of the oracle's resampler restatement.
"""
import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

RS_TOL = 1e-6


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


def _load(name):
    path = ol.small_corpus(name, ol.recipe(name))
    blob = open(path, "rb").read()
    d, chains = ol.demux_chains(blob)
    return blob, d, chains


def _check_gate(cn, name, lost=None):
    blob, d, chains = _load(name)
    data, off, ln = d.blob()
    print("case", name, "packets", len(off), "chains", len(chains), "lost", 0 if lost is None else int(lost.sum()),
          flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, chains, lost=lost)
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    assert (ref["ret"] >= 0).all()
    assert rc == 0 and not stop.any(), (rc, np.nonzero(stop)[0][:5], stop[stop != 0][:5])
    assert np.array_equal(nsamp, ref["nsamp"]), np.nonzero(nsamp != ref["nsamp"])[0][:10]
    bad = np.nonzero(rng != ref["rng"])[0]
    assert bad.size == 0, ("final range mismatch at packets", bad[:10])
    assert pcm.shape == ref["pcm"].shape
    diff = np.nonzero(pcm != ref["pcm"])[0]
    assert diff.size == 0, ("pcm mismatch at", diff[:10])


@pytest.mark.parametrize("name", sorted(ol.RECIPES))
def test_gate_bitexact(cn, name):
    _check_gate(cn, name)


@pytest.mark.parametrize("name", ["hybrid_fec", "silk_wb", "hybrid_fb", "silk_nb", "celt_fb", "mixed", "celt_multi",
                                  "silk_40ms", "hybrid_10ms", "silk_dtx", "mixed_dtx"])
def test_gate_loss_bitexact(cn, name):
    _, d, _ = _load(name)
    lost = ol.loss_mask(len(d.packets), 0.05, seed=sum(name.encode()))
    _check_gate(cn, name, lost)


@pytest.mark.parametrize("name", sorted(ol.SEED_SWEEP))
def test_gate_seed_sweep_loss_bitexact(cn, name):
    _, d, _ = _load(name)
    lost = ol.loss_mask(len(d.packets), 0.05, seed=sum(name.encode()))
    _check_gate(cn, name, lost=lost)


@pytest.mark.parametrize("name", ["silk_wb", "celt_fb", "hybrid_fb", "mixed", "len_short", "len_1s", "len_1s_plus"])
def test_f16k_matches_oracle(cn, name):
    import torch
    blob, d, chains = _load(name)
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains)
    want = ol.resample(ref["pcm"])
    out = cn.decode_ogg_to_cuda_tensor(blob)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert got.shape == want.shape
    md = float(np.abs(got - want).max()) if got.size else 0.0
    print("resampler max abs diff", md)
    assert md <= RS_TOL


def test_s16_matches_oracle(cn):
    blob, d, chains = _load("hybrid_swb")
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains)
    got = cn.decode_ogg_to_s16_tensor(blob).cpu().numpy()
    assert np.array_equal(got, ref["pcm"])


def test_multi_equals_single(cn):
    import torch
    names = ["silk_wb", "celt_fb", "hybrid_fb", "silk_nb"]
    blobs = [_load(n)[0] for n in names]
    multi = cn.decode_ogg_multi(blobs)
    singles = [cn.decode_ogg_to_cuda_tensor(b) for b in blobs]
    torch.cuda.synchronize()
    for m, s in zip(multi, singles):
        assert torch.equal(m, s)


def test_lossy_api_matches_oracle(cn):
    import torch
    blob, d, chains = _load("hybrid_fec")
    data, off, ln = d.blob()
    lost = ol.loss_mask(len(off), 0.02, seed=5)
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    want = ol.resample(ref["pcm"])
    got = cn.decode_ogg_lossy(blob, lost)
    torch.cuda.synchronize()
    assert float(np.abs(got.cpu().numpy() - want).max()) <= RS_TOL


@pytest.mark.parametrize("name", sorted(ol.GAIN_RECIPES))
def test_output_gain_matches_libopus(cn, name):
    """OpusHead output gain applied as libopus OPUS_SET_GAIN: the int16 and
    f32 16 kHz outputs equal libopus opened with the same gain."""
    import torch
    blob, d, chains = _load(name)
    assert d.gain_q8 != 0
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, gain=d.gain_q8)
    got = cn.decode_ogg_to_s16_tensor(blob).cpu().numpy()
    assert np.array_equal(got, ref["pcm"])
    out = cn.decode_ogg_to_cuda_tensor(blob)
    torch.cuda.synchronize()
    assert float(np.abs(out.cpu().numpy() - ol.resample(ref["pcm"])).max()) <= RS_TOL


def test_output_gain_can_be_ignored(cn, monkeypatch):
    """CUOPUS_OUTPUT_GAIN=0 decodes as the reference's path does (no gain)."""
    blob, d, chains = _load("gain_up6")
    data, off, ln = d.blob()
    monkeypatch.setenv("CUOPUS_OUTPUT_GAIN", "0")
    got = cn.decode_ogg_to_s16_tensor(blob).cpu().numpy()
    assert np.array_equal(got, ol.decode_chains(data, off, ln, chains)["pcm"])


def test_lost_before_dtx_alone_equals_batched(cn):
    """A lost packet followed by a TOC-only (DTX) packet: the FEC decision
    reads no byte past the next packet, so a file decodes the same alone and
    merged with other files (whose bytes follow it in the merged buffer)."""
    import torch
    blob, d, chains = _load("silk_dtx")
    lens = np.array([len(p) for p in d.packets])
    dtx = np.nonzero(lens <= 1)[0]
    assert dtx.size > 0
    lost = np.zeros(len(d.packets), np.uint8)
    lost[dtx[dtx > 0] - 1] = 1
    lost[-1] = 1
    other, _, _ = _load("celt_hi")
    alone = cn.decode_ogg_lossy(blob, lost)
    b = cn.Batch([blob, other], lost_masks=[lost, None], mode=cn.MODE_PERSISTENT)
    b.run()
    b.check()
    torch.cuda.synchronize()
    assert torch.equal(b.outputs[0], alone)
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    assert float(np.abs(alone.cpu().numpy() - ol.resample(ref["pcm"])).max()) <= RS_TOL
    b.close()
