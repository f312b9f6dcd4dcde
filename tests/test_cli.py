"""The command-line front-end (cuopus_amd/cuopus_decode, the reference CLI's
--ogg mode, decode_cli.cu:680-734).

CPU: --dump-blob / --dump-chains write the CUOP / CUC2 layouts, equal to the
oracle demuxer's packets and chain table; usage errors exit 2.
GPU: the int16 48 kHz output and the --f16k output equal the oracle.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from tests import oracle_lib as ol

CLI = os.path.join(ol.ROOT, "cuopus_amd", "cuopus_decode")


def _cli(*args, check=True):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    r = subprocess.run([CLI] + list(args), capture_output=True, timeout=300)
    if check:
        assert r.returncode == 0, r.stderr.decode()
    return r


@pytest.mark.parametrize("name", ["silk_wb", "celt_hi", "mixed_dtx"])
def test_dumps_match_oracle(name, tmp_path):
    path = ol.small_corpus(name, ol.recipe(name))
    blob_f, chains_f = tmp_path / "p.cuop", tmp_path / "c.cuc2"
    _cli("--ogg", path, "--dump-blob", str(blob_f), "--dump-chains", str(chains_f))
    d, chains = ol.demux_chains(open(path, "rb").read())
    raw = blob_f.read_bytes()
    assert raw[:4] == b"CUOP"
    n = struct.unpack("<I", raw[4:8])[0]
    lens = np.frombuffer(raw[8:8 + 4 * n], "<u4")
    assert n == len(d.packets) and list(lens) == [len(p) for p in d.packets]
    assert raw[8 + 4 * n:] == b"".join(d.packets)
    raw = chains_f.read_bytes()
    assert raw[:4] == b"CUC2" and struct.unpack("<I", raw[4:8])[0] == len(chains)
    assert np.array_equal(np.frombuffer(raw[8:], "<u4").reshape(-1, 4), np.array(chains, dtype=np.uint32))


def test_seg_and_preroll_options(tmp_path):
    path = ol.small_corpus("celt_fb", ol.recipe("celt_fb"))
    out = tmp_path / "c.cuc2"
    _cli("--ogg", path, "--seg", "2", "--preroll", "80", "--dump-chains", str(out))
    _, chains = ol.demux_chains(open(path, "rb").read(), 2.0, 80.0)
    assert np.array_equal(np.frombuffer(out.read_bytes()[8:], "<u4").reshape(-1, 4), np.array(chains, np.uint32))


def test_usage_errors():
    assert _cli(check=False).returncode == 2
    assert _cli("--ogg", "x.opus", "--bogus", check=False).returncode == 2
    assert _cli("--ogg", "/nonexistent.opus", "-", check=False).returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["hybrid_fb", "celt_fb", "gain_up6"])
def test_decode_outputs_match_oracle(name, tmp_path):
    path = ol.small_corpus(name, ol.recipe(name))
    d, chains = ol.demux_chains(open(path, "rb").read())
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, gain=d.gain_q8)
    s16, f16 = tmp_path / "o.s16", tmp_path / "o.f32"
    print("case cli", name, flush=True)
    _cli("--ogg", path, str(s16), "--stride", "16")
    _cli("--ogg", path, str(f16), "--f16k")
    assert np.array_equal(np.fromfile(s16, "<i2"), ref["pcm"])
    got = np.fromfile(f16, "<f4")
    want = ol.resample(ref["pcm"])
    assert got.shape == want.shape and float(np.abs(got - want).max()) <= 1e-6
    r = _cli("--ogg", path, "-", "--f16k")
    assert r.stdout == f16.read_bytes()
