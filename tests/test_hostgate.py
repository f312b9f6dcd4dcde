"""CPU: the decoder source of the HIP kernel, compiled for the host, against
the fixed-point libopus oracle (the reference's same-source host gate,
harness_host.cpp).  Separates "decoder logic correct" from "GPU execution
correct"; the product never runs this build.

Bar: bit-exact per-packet sample counts, final ranges and every int16 sample.
"""
import ctypes
import os

import numpy as np
import pytest

from tests import oracle_lib as ol

HG = os.path.join(ol.ROOT, "tests", "hostgate", "libhostgate.so")
P = ctypes.c_void_p


@pytest.fixture(scope="module")
def hg():
    if not os.path.exists(HG):
        pytest.skip("host gate not built")
    L = ctypes.CDLL(HG)
    L.hg_decode_chain.restype = ctypes.c_long
    L.hg_decode_chain.argtypes = [P, P, P, P, ctypes.c_int, P, ctypes.c_long, P, P]
    L.hg_ec_schedule.restype = None
    L.hg_ec_schedule.argtypes = [P, ctypes.c_uint32, P]
    return L


def _hg_chains(L, data, off, ln, chains, lost=None):
    pcm_all, nsamp, rng = [], np.zeros(len(off), np.int32), np.zeros(len(off), np.uint32)
    rets = []
    for (first, npk, trim, keep) in chains:
        cap = 5760 * (npk + 1)
        pcm = np.zeros(cap, np.int16)
        o = np.ascontiguousarray(off[first:first + npk])
        n = np.ascontiguousarray(ln[first:first + npk])
        ns = np.zeros(npk, np.int32)
        rg = np.zeros(npk, np.uint32)
        lp = None
        if lost is not None:
            lc = np.ascontiguousarray(lost[first:first + npk])
            lp = lc.ctypes.data
        r = L.hg_decode_chain(data.ctypes.data, o.ctypes.data, n.ctypes.data, lp, npk, pcm.ctypes.data, cap,
                              ns.ctypes.data, rg.ctypes.data)
        rets.append(r)
        nsamp[first:first + npk] = ns
        rng[first:first + npk] = rg
        pcm_all.append(pcm[trim:trim + keep])
    return np.concatenate(pcm_all), nsamp, rng, np.array(rets)


@pytest.mark.parametrize("name", sorted(ol.RECIPES))
def test_hostgate_bitexact(hg, name):
    blob = open(ol.small_corpus(name, ol.RECIPES[name]), "rb").read()
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    pcm, nsamp, rng, ret = _hg_chains(hg, data, off, ln, chains)
    ref = ol.decode_chains(data, off, ln, chains)
    assert (ret >= 0).all()
    assert np.array_equal(nsamp, ref["nsamp"])
    assert np.array_equal(rng, ref["rng"])
    assert np.array_equal(pcm, ref["pcm"])


@pytest.mark.parametrize("name,p", [("hybrid_fec", 0.05), ("silk_wb", 0.05), ("silk_nb", 0.1),
                                    ("hybrid_swb", 0.05), ("silk_long", 0.05), ("celt_fb", 0.1),
                                    ("mixed", 0.05), ("celt_multi", 0.1), ("silk_40ms", 0.1),
                                    ("hybrid_10ms", 0.05), ("silk_dtx", 0.05), ("mixed_dtx", 0.05)])
def test_hostgate_loss_bitexact(hg, name, p):
    blob = open(ol.small_corpus(name, ol.RECIPES[name]), "rb").read()
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    lost = ol.loss_mask(len(off), p, seed=sum(name.encode()))
    pcm, nsamp, rng, ret = _hg_chains(hg, data, off, ln, chains, lost)
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    assert (ret >= 0).all()
    assert np.array_equal(nsamp, ref["nsamp"])
    assert np.array_equal(rng, ref["rng"])
    assert np.array_equal(pcm, ref["pcm"])


@pytest.mark.parametrize("name", sorted(ol.SEED_SWEEP))
def test_hostgate_seed_sweep_loss_bitexact(hg, name):
    blob = open(ol.small_corpus(name, ol.recipe(name)), "rb").read()
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    lost = ol.loss_mask(len(off), 0.05, seed=sum(name.encode()))
    pcm, nsamp, rng, ret = _hg_chains(hg, data, off, ln, chains, lost)
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    assert (ret >= 0).all()
    assert np.array_equal(nsamp, ref["nsamp"])
    assert np.array_equal(rng, ref["rng"])
    assert np.array_equal(pcm, ref["pcm"])


def test_range_decoder_schedule(hg):
    """The 144-symbol schedule (schedule.inc:19-37) through the product's
    range decoder vs libopus entdec, on random and real packets."""
    rs = np.random.default_rng(11)
    bufs = [rs.integers(0, 256, size=int(n), dtype=np.uint8) for n in rs.integers(1, 1276, size=300)]
    blob = open(ol.small_corpus("celt_fb", ol.RECIPES["celt_fb"]), "rb").read()
    d, _ = ol.demux_chains(blob)
    bufs += [np.frombuffer(p, np.uint8).copy() for p in d.packets[:300]]
    for b in bufs:
        a = np.zeros(147, np.uint32)
        c = np.zeros(147, np.uint32)
        ol.lib().orc_ec_schedule(b.ctypes.data, b.size, a.ctypes.data)
        hg.hg_ec_schedule(b.ctypes.data, b.size, c.ctypes.data)
        assert np.array_equal(a, c)


@pytest.mark.parametrize("name", sorted(ol.GAIN_RECIPES))
def test_hostgate_output_gain(hg, name):
    """OpusHead output gain through the decoder's OPUS_SET_GAIN path: every
    sample equal to libopus with the same gain (gain changes neither sample
    counts nor final ranges)."""
    hg.hg_decode_chain_gain.restype = ctypes.c_long
    hg.hg_decode_chain_gain.argtypes = [P, P, P, P, ctypes.c_int, P, ctypes.c_long, P, P, ctypes.c_int]
    blob = open(ol.small_corpus(name, ol.recipe(name)), "rb").read()
    d, chains = ol.demux_chains(blob)
    assert d.gain_q8 != 0
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, gain=d.gain_q8)
    plain = ol.decode_chains(data, off, ln, chains)
    assert not np.array_equal(ref["pcm"], plain["pcm"])
    pcm_all = []
    for (first, npk, trim, keep) in chains:
        cap = 5760 * (npk + 1)
        pcm = np.zeros(cap, np.int16)
        o = np.ascontiguousarray(off[first:first + npk])
        n = np.ascontiguousarray(ln[first:first + npk])
        r = hg.hg_decode_chain_gain(data.ctypes.data, o.ctypes.data, n.ctypes.data, None, npk, pcm.ctypes.data,
                                    cap, None, None, d.gain_q8)
        assert r >= 0
        pcm_all.append(pcm[trim:trim + keep])
    got = np.concatenate(pcm_all)
    assert np.array_equal(got, ref["pcm"])
    assert np.array_equal(ref["nsamp"], plain["nsamp"]) and np.array_equal(ref["rng"], plain["rng"])
