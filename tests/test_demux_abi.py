"""CPU: the C ABI library loads without a GPU, exports every symbol declared
in include/cuopus.h, and its host-side demuxer / chain builder / probe agree
with the oracle's restatement of cuopus_demux.h on every corpus and on the
error paths the reference tests (cuopus_demux.h:154-181, libcuopus.cu:175-188).
"""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

from tests import oracle_lib as ol
from oracle import demux as odemux

HDR = os.path.join(ol.ROOT, "include", "cuopus.h")


@pytest.fixture(scope="module")
def cn():
    from cuopus_amd import cuopus_native
    cuopus_native._load()
    return cuopus_native


def test_exports_every_declared_symbol(cn):
    text = open(HDR).read()
    names = re.findall(r"^\s*(?:int|void|long long)\s+(cuopus_\w+)\s*\(", text, re.M)
    assert len(names) >= 14
    lib = ctypes.CDLL(cn.library_path())
    for n in names:
        assert hasattr(lib, n), n
    for n in ["cuopus_probe", "cuopus_decode_f16k", "cuopus_decode_f16k_on", "cuopus_decode_f16k_multi",
              "cuopus_warmup", "cuopus_last_arena_overflow", "cuopus_workspace_bytes",
              "cuopus_release_workspaces"]:
        assert n in names


@pytest.mark.parametrize("name", sorted(ol.RECIPES))
def test_demux_and_chains_match_oracle(cn, name):
    blob = open(ol.small_corpus(name, ol.RECIPES[name]), "rb").read()
    d, chains = ol.demux_chains(blob)
    assert cn.probe(blob) == odemux.out16_len(d.pcm_samples_48k)
    npk, nch = cn.packet_count(blob)
    assert npk == len(d.packets) and nch == len(chains)
    assert np.array_equal(cn.chain_table(blob), np.array(chains, dtype=np.uint32))


@pytest.mark.parametrize("name", ["silk_wb", "celt_fb"])
@pytest.mark.parametrize("seconds,preroll", [(1.0, 250.0), (0.5, 80.0), (2.0, 1000.0)])
def test_chain_params_env(cn, name, seconds, preroll, monkeypatch):
    monkeypatch.setenv("CUOPUS_SEG_S", str(seconds))
    monkeypatch.setenv("CUOPUS_PREROLL_MS", str(preroll))
    blob = open(ol.small_corpus(name, ol.RECIPES[name]), "rb").read()
    _, chains = ol.demux_chains(blob, seconds, preroll)
    assert np.array_equal(cn.chain_table(blob), np.array(chains, dtype=np.uint32))


def _rc(cn, blob):
    try:
        cn.probe(blob)
        return 0
    except cn.CuopusError as e:
        return e.rc


def test_error_paths(cn):
    blob = open(ol.small_corpus("silk_wb", ol.RECIPES["silk_wb"]), "rb").read()
    assert _rc(cn, b"") == 3                       # empty input -> caller error
    assert _rc(cn, b"not an ogg stream") == 1       # no pages
    assert _rc(cn, blob[:60]) == 1                  # truncated: fewer than 2 packets
    assert cn.probe(blob[:200]) == 0                # headers only: nothing playable
    # more than two channels (a multistream mapping) -> soft failure
    i = blob.find(b"OpusHead")
    st = bytearray(blob)
    st[i + 9] = 3
    assert _rc(cn, bytes(st)) == 1
    # corrupt a mid-stream granule -> cross-check failure
    pos = [m.start() for m in re.finditer(b"OggS", blob)]
    g = bytearray(blob)
    p = pos[len(pos) // 2]
    struct.pack_into("<q", g, p + 6, struct.unpack_from("<q", blob, p + 6)[0] + 480)
    assert _rc(cn, bytes(g)) == 1
    # the oracle restatement agrees on the mismatch
    d = odemux.demux(bytes(g))
    assert d.granule_match < d.granule_checked
