"""CPU: the demuxer's parallel path gives what its sequential path gives.

A well-formed single-stream file at least CUOPUS_DEMUX_PARALLEL_MIN bytes long
is demuxed in parallel page and packet ranges (host/demux.cpp scan_parallel
and the ranged packet table); any other stream falls back to the one-pass
scan.  Here every corpus and every framing case of the robust demuxer (pages
re-laced so packets span pages, chained links, foreign pages, junk between
pages, a torn final page, a damaged page, a shifted granule) is demuxed with
one host thread and with five threads forced onto the parallel path, under
each CRC and granule policy: packet bytes and lengths, chain table, demuxer
report, output length and error code and message must be identical.
"""
import numpy as np
import pytest

from tests import oracle_lib as ol
from tests import ogg_mux


@pytest.fixture(scope="module")
def cn():
    from cuopus_amd import cuopus_native
    cuopus_native._load()
    return cuopus_native


def _blob(name):
    return open(ol.small_corpus(name, ol.recipe(name)), "rb").read()


def _variants():
    out = {}
    for name in sorted(ol.RECIPES) + sorted(ol.ST_RECIPES) + sorted(ol.GAIN_RECIPES):
        out[name] = lambda name=name: _blob(name)
    out["spanning_700"] = lambda: ogg_mux.remux(_blob("celt_hi"), max_body=700)
    out["spanning_300"] = lambda: ogg_mux.remux(_blob("silk_wb"), max_body=300)
    out["chained"] = lambda: _blob("silk_wb") + _blob("celt_fb")
    out["chained_same_serial"] = lambda: _blob("hybrid_fb") + _blob("hybrid_fb")
    out["shifted_granule"] = lambda: ogg_mux.remux(_blob("celt_fb"), max_body=4096, granule_offset=3 * 48000)

    def foreign():
        blob = _blob("silk_wb")
        parts = [ogg_mux.page_bytes(2, 0, 0xF00D, 0, [9], b"\x80foreign")]
        for i, pg in enumerate(ogg_mux.pages(blob)):
            parts.append(blob[pg["off"]:pg["off"] + pg["size"]])
            if i % 3 == 2:
                parts.append(ogg_mux.page_bytes(0, i, 0xF00D, i, [50], bytes(range(50))))
        return b"".join(parts)
    out["foreign"] = foreign

    def junk():
        blob = _blob("hybrid_fb")
        pgs = ogg_mux.pages(blob)
        parts = []
        for i, pg in enumerate(pgs):
            parts.append(blob[pg["off"]:pg["off"] + pg["size"]])
            if i == len(pgs) // 2:
                parts.append(b"\x00garbage\xffOgg" * 3)
        return b"".join(parts)
    out["junk"] = junk
    out["torn"] = lambda: junk()[:-10]

    def damaged():
        blob = bytearray(_blob("silk_wb"))
        pg = ogg_mux.pages(bytes(blob))[len(ogg_mux.pages(bytes(blob))) // 2]
        blob[pg["off"] + pg["size"] - 5] ^= 0x5A
        return bytes(blob)
    out["damaged"] = damaged

    def dropped_page():
        rb = ogg_mux.remux(_blob("celt_hi"), max_body=700)
        pgs = ogg_mux.pages(rb)
        k = len(pgs) // 2
        return rb[:pgs[k]["off"]] + rb[pgs[k]["off"] + pgs[k]["size"]:]
    out["dropped_page"] = dropped_page
    out["eos_midway"] = lambda: _eos_midway(_blob("celt_fb"))
    return out


def _eos_midway(blob):
    """The EOS flag set on a middle page (a later page then follows an EOS)."""
    b = bytearray(blob)
    pgs = ogg_mux.pages(blob)
    pg = pgs[len(pgs) // 2]
    b[pg["off"] + 5] |= 4
    return bytes(b)


VARIANTS = _variants()


def _result(cn, blob):
    try:
        data, lens = cn.packets(blob)
        return ("ok", data.tobytes(), lens.tobytes(), cn.chain_table(blob).tobytes(),
                tuple(sorted(cn.demux_stats(blob).items())), cn.probe(blob))
    except cn.CuopusError as e:
        return ("error", e.rc, str(e))


@pytest.mark.parametrize("policy", ["default", "crc", "lenient", "crc_lenient"])
@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_parallel_demux_equals_sequential(cn, monkeypatch, name, policy):
    blob = VARIANTS[name]()
    if "crc" in policy:
        monkeypatch.setenv("CUOPUS_CHECK_CRC", "1")
    if "lenient" in policy:
        monkeypatch.setenv("CUOPUS_GRANULE", "warn")
    monkeypatch.setenv("CUOPUS_HOST_THREADS", "1")
    seq = _result(cn, blob)
    monkeypatch.setenv("CUOPUS_HOST_THREADS", "5")
    monkeypatch.setenv("CUOPUS_DEMUX_PARALLEL_MIN", "0")
    par = _result(cn, blob)
    assert seq[0] == par[0], (seq[:3], par[:3])
    for i, (a, b) in enumerate(zip(seq, par)):
        assert a == b, (name, policy, i)


def test_parallel_demux_of_many_files_matches_one_by_one(cn):
    """decode_ogg_multi opens its files side by side: each parsed handle
    equals the one opened alone."""
    names = ["silk_wb", "celt_fb", "hybrid_fb", "mixed", "st_celt"]
    blobs = [_blob(n) for n in names]
    together = cn._parse_all(blobs)
    try:
        for b, f in zip(blobs, together):
            with cn._Parsed(b) as g:
                assert [getattr(f, k) for k in cn._Parsed.FIELDS] == [getattr(g, k) for k in cn._Parsed.FIELDS]
    finally:
        for f in together:
            f.close()
    with pytest.raises(cn.CuopusError):
        cn._parse_all(blobs[:2] + [b"not an ogg stream"] + blobs[2:])
