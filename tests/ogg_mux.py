"""Test helper: split an Ogg/Opus file into its packets and write them back as
Ogg pages with chosen page sizes, serial numbers, granule offsets and flags,
so the demuxer can be fed packets that span pages, chained and multiplexed
streams, streams starting at a non-zero granule, damaged and missing pages.

Page layout follows RFC 3533 (header fields, lacing, continuation / BOS / EOS
flags, CRC from oracle.demux.ogg_crc); granules follow RFC 7845 (48 kHz
samples after the last packet completed on the page, -1 when none completes,
end-trim on the final page).
"""
import struct

from oracle import demux as odemux


def pages(blob):
    """List of page dicts (off, size, flags, granule, serial, seq, lacing, body)."""
    out = []
    pos = 0
    while True:
        i = blob.find(b"OggS", pos)
        if i < 0 or i + 27 > len(blob):
            return out
        flags = blob[i + 5]
        granule, serial, seq = struct.unpack_from("<qII", blob, i + 6)
        nseg = blob[i + 26]
        lacing = list(blob[i + 27:i + 27 + nseg])
        b0 = i + 27 + nseg
        body = blob[b0:b0 + sum(lacing)]
        size = 27 + nseg + len(body)
        out.append(dict(off=i, size=size, flags=flags, granule=granule, serial=serial, seq=seq,
                        lacing=lacing, body=body))
        pos = i + size


def split(blob):
    """(head, tags, audio packets, final granule) of a single-link file."""
    pk, acc = [], b""
    for pg in pages(blob):
        o = 0
        for lac in pg["lacing"]:
            acc += pg["body"][o:o + lac]
            o += lac
            if lac < 255:
                pk.append(acc)
                acc = b""
    last = max(pg["granule"] for pg in pages(blob))
    return pk[0], pk[1], pk[2:], last


def page_bytes(flags, granule, serial, seq, lacing, body):
    hdr = bytearray(b"OggS" + bytes([0, flags]) + struct.pack("<qIII", granule, serial, seq, 0) +
                    bytes([len(lacing)]) + bytes(lacing))
    crc = odemux.ogg_crc(bytes(hdr) + body)
    struct.pack_into("<I", hdr, 22, crc)
    return bytes(hdr) + body


def mux(head, tags, audio, final_granule, serial=0x5EED, max_body=4096, granule_offset=0, start_seq=0):
    """Pages for one link: head (BOS) and tags pages, then the audio packets
    laced into pages of at most max_body bytes (packets span pages where they
    do not fit; the continued flag marks such pages); the last page has EOS.
    Returns a list of page byte strings."""
    out = []
    seq = start_seq
    out.append(page_bytes(2, 0, serial, seq, [len(head)], head))
    seq += 1
    lac = [255] * (len(tags) // 255) + [len(tags) % 255]
    out.append(page_bytes(0, 0, serial, seq, lac, tags))
    seq += 1
    toc_total = sum(odemux.pkt_samples48(p) for p in audio)
    end_trim = toc_total - final_granule
    segs = []  # (lacing value, bytes, completes packet k or -1)
    cum = 0
    for k, p in enumerate(audio):
        cum += odemux.pkt_samples48(p)
        n = len(p)
        o = 0
        while n - o >= 255:
            segs.append((255, p[o:o + 255], -1))
            o += 255
        segs.append((n - o, p[o:], cum))
    cur_lac, cur_body, cur_gran = [], b"", -1
    continued = False
    for si, (lv, data, g) in enumerate(segs):
        if cur_lac and (len(cur_lac) == 255 or len(cur_body) + lv > max_body):
            nxt_cont = cur_lac[-1] == 255
            out.append((continued, cur_gran, cur_lac, cur_body))
            continued = nxt_cont
            cur_lac, cur_body, cur_gran = [], b"", -1
        cur_lac.append(lv)
        cur_body += data
        if g >= 0:
            cur_gran = g
    out.append((continued, cur_gran, cur_lac, cur_body))
    res = out[:2]
    tail = out[2:]
    for j, (cont, gran, lac, body) in enumerate(tail):
        flags = (1 if cont else 0) | (4 if j == len(tail) - 1 else 0)
        if gran >= 0:
            if j == len(tail) - 1:
                gran -= end_trim
            gran += granule_offset
        res.append(page_bytes(flags, gran, serial, seq, lac, body))
        seq += 1
    return res


def remux(blob, **kw):
    head, tags, audio, last = split(blob)
    return b"".join(mux(head, tags, audio, last, **kw))
