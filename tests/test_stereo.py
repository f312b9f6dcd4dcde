"""Stereo decode: two coded channels (CELT mid/side, intensity and dual
stereo, phase inversion; SILK mid/side prediction, mid-only frames, stereo /
mono switching; hybrid; FEC and concealment on two channels), decoded by
the stereo-capable kernel build against fixed-point libopus opened with the
same number of output channels.

Output channels 2: interleaved left/right, as opus_decode on a stereo
decoder.  Output channels 1: a stereo stream decoded by a mono decoder (SILK
renders the mid channel, CELT averages the two denormalised spectra and
turns phase inversion off), as opus_decode on a mono decoder; this is what
the f32 16 kHz entry points produce for a stereo file.

CPU: the decoder source compiled for the host (tests/hostgate, stereo build)
-- bit-exact return codes, per-packet sample counts, final ranges and PCM;
the stereo build on mono streams too (to one and to two output channels).
GPU: the same through cuopus_gate_chains_ch, the int16 entry point (stereo
interleaved, OpusHead gain applied), the f32 16 kHz entry point (mono
render, resampled) and a launch mixing mono and stereo files.
"""
import ctypes
import os
import time

import numpy as np
import pytest

from tests import oracle_lib as ol

HG_ST = os.path.join(ol.ROOT, "tests", "hostgate", "libhostgate_st.so")
P = ctypes.c_void_p
RS_TOL = 1e-6
ST = sorted(ol.ST_RECIPES)


def _load(name):
    path = ol.small_corpus(name, ol.recipe(name))
    blob = open(path, "rb").read()
    d, chains = ol.demux_chains(blob)
    return blob, d, chains


def _lost(name, n):
    return ol.loss_mask(n, 0.08, seed=sum(name.encode()))


@pytest.fixture(scope="module")
def hg():
    if not os.path.exists(HG_ST):
        pytest.skip("stereo host gate not built")
    L = ctypes.CDLL(HG_ST)
    L.hg_decode_chain_cfg.restype = ctypes.c_long
    L.hg_decode_chain_cfg.argtypes = [P, P, P, P, ctypes.c_int, P, ctypes.c_long, P, P, ctypes.c_int, ctypes.c_int]
    L.hg_max_channels.restype = ctypes.c_int
    return L


def _hg_chains(L, data, off, ln, chains, lost, channels):
    outs, rets = [], []
    nsamp, rng = np.zeros(len(off), np.int32), np.zeros(len(off), np.uint32)
    for (first, npk, trim, keep) in chains:
        cap = 5760 * (npk + 1)
        pcm = np.zeros(cap * channels, np.int16)
        o = np.ascontiguousarray(off[first:first + npk])
        n = np.ascontiguousarray(ln[first:first + npk])
        ns = np.zeros(npk, np.int32)
        rg = np.zeros(npk, np.uint32)
        lc = None if lost is None else np.ascontiguousarray(lost[first:first + npk])
        r = L.hg_decode_chain_cfg(data.ctypes.data, o.ctypes.data, n.ctypes.data,
                                  None if lc is None else lc.ctypes.data, npk, pcm.ctypes.data, cap,
                                  ns.ctypes.data, rg.ctypes.data, channels, 0)
        rets.append(r)
        nsamp[first:first + npk] = ns
        rng[first:first + npk] = rg
        outs.append(pcm[channels * trim:channels * (trim + keep)])
    return np.concatenate(outs), nsamp, rng, np.array(rets)


def _assert_same(got, ref, what):
    pcm, nsamp, rng = got
    assert np.array_equal(nsamp, ref["nsamp"]), (what, np.nonzero(nsamp != ref["nsamp"])[0][:10])
    bad = np.nonzero(rng != ref["rng"])[0]
    assert bad.size == 0, (what, "final range mismatch at packets", bad[:10])
    assert pcm.shape == ref["pcm"].shape, (what, pcm.shape, ref["pcm"].shape)
    diff = np.nonzero(pcm != ref["pcm"])[0]
    assert diff.size == 0, (what, "pcm mismatch at", diff[:10])


def test_corpora_are_stereo():
    """Every stereo recipe yields a 2-channel OpusHead; the mono-flip mixes
    hold mono packets between the stereo ones, and at 10 kbps the encoder
    codes the whole stereo stream as mono packets (mono to two outputs)."""
    for name in ST:
        _, d, _ = _load(name)
        assert d.channels == 2, name
        st = sum(1 for p in d.packets if p and p[0] & 4)
        if name == "st_silk_lo":
            assert st == 0, name
            continue
        assert st > 0, name
        if "flip" in name or name in ("st_mixed", "st_celt_lo"):
            assert st < len(d.packets), name


@pytest.mark.parametrize("name", ST)
@pytest.mark.parametrize("channels", [2, 1])
@pytest.mark.parametrize("lossy", [False, True])
def test_hostgate_stereo_bitexact(hg, name, channels, lossy):
    assert hg.hg_max_channels() == 2
    _, d, chains = _load(name)
    data, off, ln = d.blob()
    lost = _lost(name, len(off)) if lossy else None
    ref = ol.decode_chains(data, off, ln, chains, lost=lost, channels=channels)
    pcm, nsamp, rng, ret = _hg_chains(hg, data, off, ln, chains, lost, channels)
    assert np.array_equal(ret >= 0, ref["ret"] >= 0)
    _assert_same((pcm, nsamp, rng), ref, (name, channels, lossy))


@pytest.mark.parametrize("name", ["silk_wb", "celt_fb", "hybrid_fec", "mixed", "mixed_dtx", "celt_multi"])
@pytest.mark.parametrize("channels", [1, 2])
def test_hostgate_stereo_build_on_mono_streams(hg, name, channels):
    """The stereo build decodes mono streams exactly as libopus does, to one
    output channel and duplicated to two."""
    _, d, chains = _load(name)
    data, off, ln = d.blob()
    lost = _lost(name, len(off))
    ref = ol.decode_chains(data, off, ln, chains, lost=lost, channels=channels)
    pcm, nsamp, rng, ret = _hg_chains(hg, data, off, ln, chains, lost, channels)
    _assert_same((pcm, nsamp, rng), ref, (name, channels))


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


@pytest.mark.gpu
@pytest.mark.parametrize("name", ST)
@pytest.mark.parametrize("channels", [2, 1])
@pytest.mark.parametrize("lossy", [False, True])
def test_gpu_gate_stereo_bitexact(cn, name, channels, lossy):
    _, d, chains = _load(name)
    data, off, ln = d.blob()
    lost = _lost(name, len(off)) if lossy else None
    print("case stereo gate", name, "channels", channels, "lossy", lossy, "chains", len(chains), flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, chains, lost=lost, channels=channels,
                                               stereo_build=True)
    ref = ol.decode_chains(data, off, ln, chains, lost=lost, channels=channels)
    assert (ref["ret"] >= 0).all()
    assert rc == 0 and not stop.any(), (rc, np.nonzero(stop)[0][:5])
    _assert_same((pcm, nsamp, rng), ref, (name, channels, lossy))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["silk_wb", "hybrid_fec", "mixed_dtx"])
def test_gpu_stereo_build_on_mono_streams(cn, name):
    _, d, chains = _load(name)
    data, off, ln = d.blob()
    lost = _lost(name, len(off))
    print("case stereo build, mono stream", name, flush=True)
    for channels in (1, 2):
        pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, chains, lost=lost, channels=channels,
                                                   stereo_build=True)
        ref = ol.decode_chains(data, off, ln, chains, lost=lost, channels=channels)
        assert rc == 0
        _assert_same((pcm, nsamp, rng), ref, (name, channels))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["st_celt", "st_silk", "st_mixed", "st_gain"])
def test_gpu_s16_stereo_interleaved(cn, name):
    """int16 48 kHz entry point: (n, 2) interleaved L/R equal to libopus
    opened with 2 channels and OPUS_SET_GAIN of the OpusHead gain."""
    blob, d, chains = _load(name)
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, channels=2, gain=d.gain_q8)
    print("case s16 stereo", name, flush=True)
    got = cn.decode_ogg_to_s16_tensor(blob).cpu().numpy()
    assert got.ndim == 2 and got.shape[1] == 2
    assert np.array_equal(got.reshape(-1), ref["pcm"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["st_celt_lo", "st_hybrid", "st_silk_flip", "st_gain"])
def test_gpu_f16k_stereo_renders_mono(cn, name):
    """f32 16 kHz entry point on a stereo file: libopus's one-channel render
    of the stream (SILK mid, CELT downmix), resampled."""
    import torch
    blob, d, chains = _load(name)
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, channels=1, gain=d.gain_q8)
    print("case f16k stereo", name, flush=True)
    out = cn.decode_ogg_to_cuda_tensor(blob)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = ol.resample(ref["pcm"])
    assert got.shape == want.shape
    assert float(np.abs(got - want).max()) <= RS_TOL


@pytest.mark.gpu
def test_gpu_mixed_mono_stereo_launch(cn):
    """A launch holding mono and stereo files runs on the stereo build: every
    file equals its own decode."""
    import torch
    names = ["silk_wb", "st_celt", "hybrid_fb", "st_silk_flip"]
    blobs = [_load(n)[0] for n in names]
    print("case mixed mono/stereo launch", flush=True)
    multi = cn.decode_ogg_multi(blobs)
    torch.cuda.synchronize()
    for name, blob, m in zip(names, blobs, multi):
        _, d, chains = _load(name)
        data, off, ln = d.blob()
        ref = ol.decode_chains(data, off, ln, chains, channels=1, gain=d.gain_q8)
        want = ol.resample(ref["pcm"])
        got = m.cpu().numpy()
        assert got.shape == want.shape, name
        assert float(np.abs(got - want).max()) <= RS_TOL, name


def _oracle_render(name):
    _, d, chains = _load(name)
    data, off, ln = d.blob()
    return ol.resample(ol.decode_chains(data, off, ln, chains, channels=1, gain=d.gain_q8)["pcm"])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gpu_batch_with_stereo_files(cn, mode):
    """Resident batches holding stereo and mono files (the stereo kernel
    build decodes the launch): every output equals the oracle's one-channel
    render, in every batch mode, on repeated runs."""
    import torch
    names = ["st_celt", "silk_wb", "st_silk_flip", "st_fec"]
    blobs = [_load(n)[0] for n in names]
    wants = [_oracle_render(n) for n in names]
    print("case stereo batch mode", mode, flush=True)
    b = cn.Batch(blobs, mode=mode, n_slots=5 if mode == 2 else 0)
    for rep in range(2):
        for o in b.outputs:
            o.fill_(float("nan"))
        b.run()
        b.check()
        torch.cuda.synchronize()
        for name, o, want in zip(names, b.outputs, wants):
            got = o.cpu().numpy()
            assert got.shape == want.shape, name
            assert float(np.abs(got - want).max()) <= RS_TOL, (name, rep)
    b.close()


@pytest.mark.gpu
def test_gpu_server_with_stereo_requests(cn):
    """Stereo and mono requests through the continuous-batching server equal
    their single-file decodes."""
    import torch
    names = ["st_hybrid", "celt_fb", "st_gain", "st_mixed"]
    blobs = [_load(n)[0] for n in names]
    singles = [cn.decode_ogg_to_cuda_tensor(b) for b in blobs]
    srv = cn.Server(n_streams=2)
    ids = {srv.submit(b): i for i, b in enumerate(blobs)}
    got = {}
    for _ in range(20000):
        for rid, out, rc in srv.poll():
            assert rc == 0
            got[rid] = out
        if len(got) == len(ids):
            break
        time.sleep(0.001)
    torch.cuda.synchronize()
    srv.close()
    assert len(got) == len(ids)
    for rid, i in ids.items():
        assert torch.equal(got[rid], singles[i]), names[i]
