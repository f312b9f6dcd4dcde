"""CPU: the demuxer beyond the reference's strict one (the reference's known
limitation, README.md:232-235; CUOPUS_CHECK_CRC, BENCHMARKS.md:240).

Expected results are fixed by construction: the same packets re-laced into
other pages (spanning packets, continuation flags), two files back to back
(chained), foreign pages interleaved (multiplexed), every granule shifted
(non-zero start), bytes damaged or pages removed.  The original file's packet
list and chain table, from the reference restatement (oracle/demux.py), are
the answer each case must reproduce or provably reduce to.
"""
import re

import numpy as np
import pytest

from oracle import demux as odemux
from tests import oracle_lib as ol
from tests import ogg_mux


@pytest.fixture(scope="module")
def cn():
    from cuopus_amd import cuopus_native
    cuopus_native._load()
    return cuopus_native


def _blob(name):
    return open(ol.small_corpus(name, ol.RECIPES[name]), "rb").read()


def _oracle_packets(blob):
    d = odemux.demux(blob)
    return d, d.packets


def _packets(cn, blob):
    data, lens = cn.packets(blob)
    out, o = [], 0
    for n in lens:
        out.append(bytes(data[o:o + n]))
        o += int(n)
    return out


def _rc(cn, blob):
    try:
        cn.probe(blob)
        return 0
    except cn.CuopusError as e:
        return e.rc


def test_ogg_crc_known_answer():
    # CRC-32/CKSUM (poly 0x04C11DB7, init 0, no reflection) has check value
    # 0x765E7680 with its final xor 0xFFFFFFFF; the Ogg CRC omits that xor.
    assert odemux.ogg_crc(b"123456789") == 0x765E7680 ^ 0xFFFFFFFF
    assert odemux.ogg_crc(b"") == 0
    # linearity over GF(2): crc(a ^ b) = crc(a) ^ crc(b) for equal lengths
    a, b = bytes(range(40)), bytes(range(7, 47))
    assert odemux.ogg_crc(bytes(x ^ y for x, y in zip(a, b))) == odemux.ogg_crc(a) ^ odemux.ogg_crc(b)


@pytest.mark.parametrize("name", ["silk_wb", "celt_fb"])
def test_corpus_pages_carry_valid_crcs(cn, name, monkeypatch):
    blob = _blob(name)
    for pg in ogg_mux.pages(blob):
        raw = bytearray(blob[pg["off"]:pg["off"] + pg["size"]])
        want = int.from_bytes(raw[22:26], "little")
        raw[22:26] = b"\0\0\0\0"
        assert odemux.ogg_crc(bytes(raw)) == want
    monkeypatch.setenv("CUOPUS_CHECK_CRC", "1")
    st = cn.demux_stats(blob)
    assert st["crc_bad"] == 0 and st["pages"] == len(ogg_mux.pages(blob))


def test_dump_packets_match_oracle(cn, tmp_path):
    blob = _blob("mixed")
    _, want = _oracle_packets(blob)
    assert _packets(cn, blob) == want
    # CUOP / CUC2 dump files (cuopus_demux.h:387-411 layouts)
    cn.dump_blob(blob, str(tmp_path / "b.cuop"))
    raw = (tmp_path / "b.cuop").read_bytes()
    n = int.from_bytes(raw[4:8], "little")
    lens = np.frombuffer(raw[8:8 + 4 * n], "<u4")
    assert raw[:4] == b"CUOP" and n == len(want) and list(lens) == [len(p) for p in want]
    assert raw[8 + 4 * n:] == b"".join(want)
    cn.dump_chains(blob, str(tmp_path / "c.cuc2"))
    raw = (tmp_path / "c.cuc2").read_bytes()
    _, chains = ol.demux_chains(blob)
    assert raw[:4] == b"CUC2" and int.from_bytes(raw[4:8], "little") == len(chains)
    assert np.array_equal(np.frombuffer(raw[8:], "<u4").reshape(-1, 4), np.array(chains, np.uint32))


@pytest.mark.parametrize("name", ["celt_hi", "celt_fb", "hybrid_fec"])
@pytest.mark.parametrize("max_body", [97, 300, 1000])
def test_packets_spanning_pages(cn, name, max_body):
    """Re-laced into small pages (packets continue across pages): same packets,
    chain table and length, also through the reference restatement."""
    blob = _blob(name)
    d0, want = _oracle_packets(blob)
    rb = ogg_mux.remux(blob, max_body=max_body)
    assert len(ogg_mux.pages(rb)) > len(ogg_mux.pages(blob)) or max_body >= 1000
    assert _packets(cn, rb) == want
    assert _oracle_packets(rb)[1] == want
    assert np.array_equal(cn.chain_table(rb), cn.chain_table(blob))
    assert cn.probe(rb) == cn.probe(blob)
    st = cn.demux_stats(rb)
    assert st["dropped_packets"] == 0 and st["links"] == 1


@pytest.mark.parametrize("pair", [("silk_wb", "celt_fb"), ("hybrid_fb", "hybrid_fb")])
def test_chained_stream(cn, pair):
    """Two files back to back (RFC 7845 chaining; the second pair even shares
    its serial number): the links' packets in order, each link chained on its
    own timeline, playable lengths added."""
    a, b = _blob(pair[0]), _blob(pair[1])
    ab = a + b
    pa, pb = _packets(cn, a), _packets(cn, b)
    assert _packets(cn, ab) == pa + pb
    ca, cb = cn.chain_table(a), cn.chain_table(b)
    cb = cb.copy()
    cb[:, 0] += len(pa)
    assert np.array_equal(cn.chain_table(ab), np.concatenate([ca, cb]))
    st = cn.demux_stats(ab)
    assert st["links"] == 2
    assert st["play48"] == cn.demux_stats(a)["play48"] + cn.demux_stats(b)["play48"]
    assert cn.probe(ab) == odemux.out16_len(st["play48"])


def test_multiplexed_foreign_pages(cn):
    """Pages of another logical stream (a BOS page and data pages with their own
    serial) interleaved with the Opus pages are skipped."""
    blob = _blob("silk_wb")
    pgs = ogg_mux.pages(blob)
    foreign_serial = 0xF00D
    hdr = b"\x80theora-like header bytes"
    out = [ogg_mux.page_bytes(2, 0, foreign_serial, 0, [len(hdr)], hdr)]
    n_foreign = 1
    for i, pg in enumerate(pgs):
        out.append(blob[pg["off"]:pg["off"] + pg["size"]])
        if i % 2 == 1:
            body = bytes((i * 7 + j) & 0xFF for j in range(200))
            out.append(ogg_mux.page_bytes(0, i, foreign_serial, i, [200], body))
            n_foreign += 1
    mb = b"".join(out)
    assert _packets(cn, mb) == _packets(cn, blob)
    assert np.array_equal(cn.chain_table(mb), cn.chain_table(blob))
    assert cn.demux_stats(mb)["foreign_pages"] == n_foreign


def test_nonzero_start_granule(cn, monkeypatch):
    """Every granule shifted by 3 s (a stream cut from a live capture): the
    strict cross-check rejects it like the reference; the lenient policy
    re-bases once and decodes the same packets and chains."""
    blob = _blob("celt_fb")
    sh = ogg_mux.remux(blob, max_body=4096, granule_offset=3 * 48000)
    assert _rc(cn, sh) == 1
    monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    assert _packets(cn, sh) == _packets(cn, blob)
    assert np.array_equal(cn.chain_table(sh), cn.chain_table(blob))
    assert cn.probe(sh) == cn.probe(blob)
    assert cn.demux_stats(sh)["granule_warnings"] == 1
    # lenient on the unshifted file is the strict result
    assert cn.demux_stats(blob)["granule_warnings"] == 0
    assert np.array_equal(cn.chain_table(blob), np.array(ol.demux_chains(blob)[1], np.uint32))


def test_garbage_and_truncation(cn):
    """Junk between pages is skipped by the capture-pattern search and a torn
    final page is dropped, as in the reference restatement."""
    blob = _blob("hybrid_fb")
    pgs = ogg_mux.pages(blob)
    junk = b"\x00garbage\xffOgg" * 3
    parts = []
    for i, pg in enumerate(pgs):
        parts.append(blob[pg["off"]:pg["off"] + pg["size"]])
        if i == len(pgs) // 2:
            parts.append(junk)
    mangled = b"".join(parts)
    torn = mangled[:-10]
    for m in (mangled, torn):
        d, want = _oracle_packets(m)
        assert _packets(cn, m) == want
        assert np.array_equal(cn.chain_table(m), np.array(odemux.build_chains(d), np.uint32))
    st = cn.demux_stats(torn)
    assert st["skipped_bytes"] == len(junk) + pgs[-1]["size"] - 10


def _damage_page(blob, k):
    pgs = ogg_mux.pages(blob)
    pg = pgs[k]
    b = bytearray(blob)
    b[pg["off"] + pg["size"] - 5] ^= 0x5A   # a payload byte of the page's last packet
    return bytes(b), pg


def _packets_on_page(blob, k):
    """Indices (among audio packets) of the packets completing on page k."""
    idx, n = [], 0
    for i, pg in enumerate(ogg_mux.pages(blob)):
        for lac in pg["lacing"]:
            if lac < 255:
                if i == k:
                    idx.append(n - 2)
                n += 1
    return idx


def test_crc_damaged_page(cn, monkeypatch):
    blob = _blob("silk_wb")
    k = len(ogg_mux.pages(blob)) // 2
    bad, pg = _damage_page(blob, k)
    # without the check (the reference's default) the payload change passes
    assert _rc(cn, bad) == 0
    monkeypatch.setenv("CUOPUS_CHECK_CRC", "1")
    with pytest.raises(cn.CuopusError) as e:
        cn.probe(bad)
    assert e.value.rc == 1 and "CRC" in str(e.value)
    monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    gone = set(_packets_on_page(blob, k))
    want = [p for i, p in enumerate(_packets(cn, blob)) if i not in gone]
    assert _packets(cn, bad) == want
    st = cn.demux_stats(bad)
    assert st["crc_bad"] == 1 and st["granule_warnings"] >= 1
    dur = sum(odemux.pkt_samples48(_packets(cn, blob)[i]) for i in gone)
    assert st["play48"] == cn.demux_stats(blob)["play48"] - dur
    ch = cn.chain_table(bad)
    assert int(ch[:, 3].sum()) == st["play48"]


def _packet_pages(blob):
    """(first page, last page) of every audio packet."""
    rng, start, n = [], None, 0
    for i, pg in enumerate(ogg_mux.pages(blob)):
        for lac in pg["lacing"]:
            if start is None:
                start = i
            if lac < 255:
                if n >= 2:
                    rng.append((start, i))
                n += 1
                start = None
    return rng


def test_missing_page_drops_spanning_packet(cn, monkeypatch):
    """A page removed from a stream whose packets span pages: every packet with
    a piece on that page is dropped (never spliced), the rest are intact and in
    order, and the chains cover the remaining timeline."""
    blob = _blob("celt_hi")
    rb = ogg_mux.remux(blob, max_body=700)
    pgs = ogg_mux.pages(rb)
    k = next(i for i in range(len(pgs) // 2, len(pgs)) if pgs[i]["flags"] & 1 and pgs[i]["lacing"][-1] == 255)
    cut = rb[:pgs[k]["off"]] + rb[pgs[k]["off"] + pgs[k]["size"]:]
    assert _rc(cn, cut) == 1   # strict: the granules no longer match
    monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    full = _packets(cn, rb)
    want = [p for p, (a, b) in zip(full, _packet_pages(rb)) if not a <= k <= b]
    assert len(want) < len(full) - 1
    assert _packets(cn, cut) == want
    st = cn.demux_stats(cut)
    assert st["dropped_packets"] >= 1 and st["granule_warnings"] >= 1
    assert st["play48"] == cn.demux_stats(rb)["play48"] - sum(odemux.pkt_samples48(p) for p in full) + \
        sum(odemux.pkt_samples48(p) for p in want)
    assert int(cn.chain_table(cut)[:, 3].sum()) == st["play48"]


def test_error_paths_unchanged(cn, monkeypatch):
    """The lenient policy does not accept what is not an Opus stream."""
    monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    blob = _blob("silk_wb")
    assert _rc(cn, b"not an ogg stream") == 1
    assert _rc(cn, blob[:60]) == 1
    st = bytearray(blob)
    st[blob.find(b"OpusHead") + 9] = 3  # three channels: a multistream mapping, not supported
    assert _rc(cn, bytes(st)) == 1
    assert re.search(rb"OpusHead", blob)
