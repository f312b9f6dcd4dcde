"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from
fixed-point libopus 1.5.2): short seeded streams with their chain tables,
per-packet sample counts and final ranges and compacted int16 PCM.

CPU: the oracle build, the host demuxer and the same-source CPU gate all
reproduce the fixtures.  GPU: the HIP decoder through the C ABI reproduces
them bit for bit, and its f32 16 kHz output is within 1e-6 of the oracle's
resampler applied to the golden PCM.
"""
import ctypes
import os

import numpy as np
import pytest

from tests import oracle_lib as ol

GOLD = os.path.join(ol.ROOT, "tests", "golden")
NAMES = ["g_silk_wb", "g_celt_fb", "g_hybrid_fec", "g_mixed"]


def _load(name):
    blob = open(os.path.join(GOLD, name + ".opus"), "rb").read()
    z = np.load(os.path.join(GOLD, name + ".npz"))  # allow_pickle defaults to False
    lost = z["lost"] if z["lost"].size else None
    return blob, z, lost


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    blob, z, lost = _load(name)
    d, chains = ol.demux_chains(blob)
    assert np.array_equal(np.array(chains, dtype=np.uint32).reshape(-1, 4), z["chains"])
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    assert np.array_equal(ref["nsamp"], z["nsamp"])
    assert np.array_equal(ref["rng"], z["rng"])
    assert np.array_equal(ref["pcm"], z["pcm"])


@pytest.mark.parametrize("name", NAMES)
def test_host_demuxer_matches_golden_chains(name):
    from cuopus_amd import cuopus_native as cn
    blob, z, _ = _load(name)
    assert np.array_equal(cn.chain_table(blob), z["chains"])


@pytest.mark.parametrize("name", NAMES)
def test_hostgate_reproduces_golden(name):
    hg = os.path.join(ol.ROOT, "tests", "hostgate", "libhostgate.so")
    if not os.path.exists(hg):
        pytest.skip("host gate not built")
    from tests.test_hostgate import _hg_chains
    L = ctypes.CDLL(hg)
    P = ctypes.c_void_p
    L.hg_decode_chain.restype = ctypes.c_long
    L.hg_decode_chain.argtypes = [P, P, P, P, ctypes.c_int, P, ctypes.c_long, P, P]
    blob, z, lost = _load(name)
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    pcm, nsamp, rng, ret = _hg_chains(L, data, off, ln, chains, lost)
    assert (ret >= 0).all()
    assert np.array_equal(nsamp, z["nsamp"])
    assert np.array_equal(rng, z["rng"])
    assert np.array_equal(pcm, z["pcm"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(name):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native as cn
    blob, z, lost = _load(name)
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    print("case golden", name, "lost", 0 if lost is None else int(lost.sum()), flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, z["chains"], lost=lost)
    assert rc == 0 and not stop.any()
    assert np.array_equal(nsamp, z["nsamp"])
    assert np.array_equal(rng, z["rng"])
    assert np.array_equal(pcm, z["pcm"])
    out = cn.decode_ogg_lossy(blob, lost) if lost is not None else cn.decode_ogg_to_cuda_tensor(blob)
    torch.cuda.synchronize()
    want = ol.resample(z["pcm"])
    got = out.cpu().numpy()
    assert got.shape == want.shape
    md = float(np.abs(got - want).max()) if got.size else 0.0
    print("f16k max abs diff", md)
    assert md <= 1e-6
