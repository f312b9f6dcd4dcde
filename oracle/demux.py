"""CPU restatement of the reference's Ogg/Opus demuxer and chain builder.

TEST INFRASTRUCTURE ONLY: the oracle for the product's C++ demuxer.  Follows
cuopus_demux.h (itself the C port of ogg_index.py):

* ``demux``          -- cu_demux_mem (cuopus_demux.h:79-269): page walk by
  "OggS" search, lacing-run reassembly (a run ending in a 255 lacing value
  continues on the next page; a trailing partial packet is dropped), OpusHead
  (channels, pre-skip), per-packet TOC durations (cu_dmx_pkt_samples48,
  :52-60), per-page granule cross-check with the RFC 7845 end-trim allowance
  on the final data page, playable = last granule - pre-skip.
* ``build_chains``   -- cu_build_chains2 (:293-378): 1 s playable windows,
  250 ms decoder preroll, bisect for first/last packet, trim/keep.
* ``pkt_samples48``  -- TOC duration in 48 kHz samples.
* ``out16_len``      -- libcuopus.cu:195-199, (play48 - 48) // 3 clamped at 0.
* ``ogg_crc``        -- the page CRC of ogg_index.py:22-35 (polynomial
  0x04C11DB7, initial value 0, MSB first, no final xor; RFC 3533 section 6).
"""
import bisect
import struct

import numpy as np

_QMS = [40, 80, 160, 240, 40, 80, 160, 240, 40, 80, 160, 240, 40, 80, 40, 80,
        10, 20, 40, 80, 10, 20, 40, 80, 10, 20, 40, 80, 10, 20, 40, 80]


def pkt_samples48(p):
    if len(p) == 0:
        return 0
    config, code = p[0] >> 3, p[0] & 3
    if code == 0:
        nf = 1
    elif code < 3:
        nf = 2
    else:
        if len(p) < 2:
            return 0
        nf = p[1] & 0x3F
    return _QMS[config] * 12 * nf


def _crc_table():
    t = []
    for i in range(256):
        r = i << 24
        for _ in range(8):
            r = ((r << 1) ^ 0x04C11DB7) & 0xFFFFFFFF if r & 0x80000000 else (r << 1) & 0xFFFFFFFF
        t.append(r)
    return t


_CRC = _crc_table()


def ogg_crc(data, crc=0):
    """Ogg page CRC-32 of a byte string (ogg_index.py:22-35)."""
    for b in bytes(data):
        crc = ((crc << 8) & 0xFFFFFFFF) ^ _CRC[((crc >> 24) ^ b) & 0xFF]
    return crc


def out16_len(play48):
    return max(0, (play48 - 48) // 3)


class DemuxError(ValueError):
    pass


class Demuxed:
    def __init__(self):
        self.packets = []        # audio packets (bytes)
        self.pkt_page = []
        self.sample_off = []
        self.channels = 0
        self.preskip = 0
        self.gain_q8 = 0
        self.last_granule = 0
        self.pcm_samples_48k = 0
        self.granule_checked = 0
        self.granule_match = 0
        self.n_pages = 0

    def blob(self):
        """(data uint8, off uint32, len uint32) of the audio packets."""
        lens = np.array([len(p) for p in self.packets], dtype=np.uint32)
        off = np.zeros(len(lens), dtype=np.uint32)
        if len(lens) > 1:
            off[1:] = np.cumsum(lens[:-1], dtype=np.uint64).astype(np.uint32)
        data = np.frombuffer(b"".join(self.packets), dtype=np.uint8)
        return data, off, lens


def demux(buf):
    """Parse an Ogg/Opus byte string (cuopus_demux.h:79-269)."""
    buf = bytes(buf)
    n = len(buf)
    if n == 0:
        raise DemuxError("empty stream")
    spans = []           # (offset, length, done, page)
    page_granule = []
    pos = 0
    while pos < n:
        idx = buf.find(b"OggS", pos)
        if idx < 0 or idx + 27 > n:
            break
        granule = struct.unpack_from("<q", buf, idx + 6)[0]
        nsegs = buf[idx + 26]
        seg = buf[idx + 27: idx + 27 + nsegs]
        body = idx + 27 + nsegs
        body_len = sum(seg)
        if body + body_len > n:
            break
        pg = len(page_granule)
        page_granule.append(granule)
        run_start, run_len, last = body, 0, 0
        for lac in seg:
            run_len += lac
            last = lac
            if lac < 255:
                spans.append((run_start, run_len, True, pg))
                run_start += run_len
                run_len = 0
        if run_len > 0 or (nsegs > 0 and last == 255):
            spans.append((run_start, run_len, False, pg))
        pos = body + body_len
    if not page_granule:
        raise DemuxError("no OGG pages found")
    packets, ppage, acc = [], [], []
    for (o, ln, done, pg) in spans:
        acc.append(buf[o:o + ln])
        if done:
            packets.append(b"".join(acc))
            ppage.append(pg)
            acc = []
    if len(packets) < 2:
        raise DemuxError("not an OGG Opus stream (fewer than 2 packets)")
    head = packets[0]
    if len(head) < 19 or head[:8] != b"OpusHead":
        raise DemuxError("not an OGG Opus stream (no OpusHead)")
    d = Demuxed()
    d.channels = head[9]
    d.preskip = head[10] | (head[11] << 8)
    d.gain_q8 = int.from_bytes(head[16:18], "little", signed=True)  # output gain, Q7.8 dB (RFC 7845 5.1)
    d.packets = packets[2:]
    d.pkt_page = ppage[2:]
    d.n_pages = len(page_granule)
    cum = 0
    for p in d.packets:
        d.sample_off.append(cum)
        cum += pkt_samples48(p)
    page_end = [-1] * len(page_granule)
    final_data_page = -1
    for k, p in enumerate(d.packets):
        pg = d.pkt_page[k]
        page_end[pg] = d.sample_off[k] + pkt_samples48(p)
        final_data_page = max(final_data_page, pg)
    last_dur = pkt_samples48(d.packets[-1]) if d.packets else 0
    checked = match = 0
    last_granule = 0
    for pg, g in enumerate(page_granule):
        if g != -1 and g > last_granule:
            last_granule = g
        if g == -1 or page_end[pg] < 0:
            continue
        checked += 1
        if g == page_end[pg]:
            match += 1
        elif pg == final_data_page and 0 <= page_end[pg] - g < last_dur:
            match += 1
    d.granule_checked, d.granule_match = checked, match
    d.last_granule = last_granule
    d.pcm_samples_48k = last_granule - d.preskip
    return d


def build_chains(d, seconds=1.0, preroll_ms=250.0):
    """cu_build_chains2 (cuopus_demux.h:293-378) -> list of (first, npkts, trim, keep)."""
    step = int(seconds * 48000)
    preroll = int(preroll_ms * 48)
    total_play = d.pcm_samples_48k
    if total_play <= 0:
        raise DemuxError("no playable samples")
    so = d.sample_off
    n_pk = len(so)
    cum = [0]
    for p in d.packets:
        cum.append(cum[-1] + pkt_samples48(p))
    chains = []
    k = 0
    while k * step < total_play:
        play_start = k * step
        play_end = min((k + 1) * step, total_play)
        toc_start = play_start + d.preskip
        toc_end = play_end + d.preskip
        start_target = max(0, toc_start - preroll)
        first = bisect.bisect_right(so, start_target) - 1
        first = max(first, 0)
        if cum[first + 1] <= start_target:
            first += 1
        last = bisect.bisect_left(so, toc_end) - 1
        if last < first:
            last = first
        trim = toc_start - so[first]
        keep = play_end - play_start
        if trim < 0:
            raise DemuxError("chain %d: negative trim" % k)
        if cum[last + 1] - cum[first] < trim + keep:
            raise DemuxError("chain %d: capacity < trim + keep" % k)
        chains.append((first, last - first + 1, trim, keep))
        k += 1
    if sum(c[3] for c in chains) != total_play:
        raise DemuxError("sum(keep) != playable")
    _ = n_pk
    return chains
