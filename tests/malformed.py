"""Malformed-input corpora for the decoder (test data, seeded).

Follows the structure of libopus's random-packet decode test
(vendor/opus-1.5.2/tests/test_opus_decode.c:295-367) for a mono 48 kHz
decoder, plus damaged real packets:

* ``random_toc``   every mono TOC configuration (0..31, code 0) with random
                   payloads of lengths from 1 byte to 1275, one chain per
                   configuration (a decoder carried across lengths);
* ``mode_pairs``   a de Bruijn sequence over the 32 configurations, so every
                   ordered pair of configurations follows each other, random
                   payloads of one length, in chains of 48 packets;
* ``framing``      codes 1, 2 and 3 with random framing bytes (frame counts,
                   padding, VBR sizes), most of them invalid;
* ``truncated``    real packets of the test corpora cut to a random shorter
                   length;
* ``bitflip``      real packets with 1-8 random bits flipped;
* ``lossy``        the damaged real chains again with 10% of packets lost
                   (FEC / concealment on top of damaged data).

Each set is (data uint8, off uint32, len uint32, lost uint8 or None, chains
(n, 4) of first/npkts/trim/keep with trim = keep = 0: callers fill keep from
the oracle).
"""
import numpy as np

from tests import oracle_lib as ol

# The mono sets keep the TOC's stereo flag clear (the mono kernel build
# decodes one coded channel); the stereo sets below set it freely and are
# decoded by the stereo build to one output channel (libopus's mono decoder:
# SILK mid channel, CELT downmix) and to two.
STEREO_PACKETS = False

REAL = ["silk_wb", "silk_nb", "celt_fb", "hybrid_fec", "mixed", "celt_multi", "silk_40ms", "hybrid_10ms",
        "silk_dtx"]


def _pack(packets, chain_sizes):
    data = np.frombuffer(b"".join(packets), np.uint8) if packets else np.zeros(0, np.uint8)
    lens = np.array([len(p) for p in packets], np.uint32)
    off = np.zeros(len(packets), np.uint32)
    if len(packets) > 1:
        off[1:] = np.cumsum(lens[:-1])
    chains, first = [], 0
    for n in chain_sizes:
        chains.append((first, n, 0, 0))
        first += n
    assert first == len(packets)
    return data.copy(), off, lens, np.array(chains, np.int64).reshape(-1, 4)


def _debruijn(k):
    """Order-2 de Bruijn sequence over k symbols (every ordered pair once)."""
    a = [0] * (2 * k)
    seq = []

    def db(t, p):
        if t > 2:
            if 2 % p == 0:
                seq.extend(a[1:p + 1])
        else:
            a[t] = a[t - p]
            db(t + 1, p)
            for j in range(a[t - p] + 1, k):
                a[t] = j
                db(t + 1, t)
    db(1, 1)
    return seq + seq[:1]


def random_toc(seed=1, stereo=False):
    """Every TOC configuration (mono, or stereo when `stereo`), code 0."""
    rs = np.random.default_rng(seed)
    packets, sizes = [], []
    for cfg in range(32):
        lens = list(range(1 + int(rs.integers(0, 7)), 1276, 37)) + [1275]
        for n in lens:
            packets.append(bytes([cfg << 3 | (4 if stereo else 0)]) +
                           rs.integers(0, 256, n - 1, dtype=np.uint8).tobytes())
        sizes.append(len(lens))
    return _pack(packets, sizes)


def mode_pairs(seed=2, plen=None, chain=48, stereo=False):
    """De Bruijn order over the 32 configurations, or over the 64
    configuration x channel-count symbols when `stereo`."""
    rs = np.random.default_rng(seed)
    plen = plen or int(rs.integers(3, 21)) * 8 + 3
    seq = _debruijn(64 if stereo else 32)
    packets = [bytes([(sym >> 1) << 3 | (sym & 1) << 2 if stereo else sym << 3]) +
               (rs.integers(0, 256, plen, dtype=np.uint8) |
                rs.integers(0, 256, plen, dtype=np.uint8)).astype(np.uint8).tobytes()
               for sym in seq]
    sizes = [chain] * (len(packets) // chain) + ([len(packets) % chain] if len(packets) % chain else [])
    return _pack(packets, sizes)


def framing(seed=3, n=1536, chain=1, stereo=False):
    rs = np.random.default_rng(seed)
    packets = []
    for _ in range(n):
        cfg = int(rs.integers(0, 32)) << 1 | (int(rs.integers(0, 2)) if stereo else 0)
        code = int(rs.integers(1, 4))
        length = int(rs.integers(1, 400))
        body = rs.integers(0, 256, length, dtype=np.uint8)
        if code == 3 and length > 1 and rs.random() < 0.7:
            body[0] = (body[0] & 0xC0) | int(rs.integers(1, 8))  # plausible frame counts
        packets.append(bytes([(cfg >> 1) << 3 | (cfg & 1) << 2 | code]) + body.tobytes())
    return _pack(packets, [chain] * (n // chain))  # one packet per chain: each from a fresh decoder


REAL_ST = ["st_celt", "st_celt_lo", "st_silk", "st_hybrid", "st_mixed", "st_fec", "st_silk_flip", "mixed",
           "hybrid_fec"]


def _real(names=REAL):
    out = []
    for name in names:
        blob = open(ol.small_corpus(name, ol.recipe(name)), "rb").read()
        d, _ = ol.demux_chains(blob)
        out.append(d.packets)
    return out


def damaged(kind, seed=4, chain=40, per_file=6, stereo=False):
    """Damaged real packets: mono corpora with the stereo flag kept clear, or
    (stereo) stereo and mono corpora with every bit open to flipping."""
    rs = np.random.default_rng(seed + (0 if kind == "truncated" else 100))
    packets, sizes = [], []
    for pk in _real(REAL_ST if stereo else REAL):
        for _ in range(per_file):
            s = int(rs.integers(0, max(1, len(pk) - chain)))
            seg = [bytearray(p) for p in pk[s:s + chain]]
            for p in seg:
                if rs.random() < 0.35 and len(p) > 1:
                    if kind == "truncated":
                        del p[int(rs.integers(1, len(p))):]
                    else:
                        for _b in range(int(rs.integers(1, 9))):
                            bit = int(rs.integers(0, 8 * len(p)))
                            p[bit >> 3] ^= 1 << (bit & 7)
                        if not (stereo or STEREO_PACKETS):
                            p[0] &= 0xFB  # keep the TOC's stereo flag clear (see STEREO_PACKETS)
            packets += [bytes(p) for p in seg]
            sizes.append(len(seg))
    return _pack(packets, sizes)


SETS = {
    "random_toc": lambda: random_toc(),
    "mode_pairs": lambda: mode_pairs(),
    "framing": lambda: framing(),
    "truncated": lambda: damaged("truncated"),
    "bitflip": lambda: damaged("bitflip"),
}

# Sets with stereo packets, for the stereo build
STEREO_SETS = {
    "st_random_toc": lambda: random_toc(seed=11, stereo=True),
    "st_mode_pairs": lambda: mode_pairs(seed=12, stereo=True),
    "st_framing": lambda: framing(seed=13, n=1024, stereo=True),
    "st_truncated": lambda: damaged("truncated", seed=14, stereo=True),
    "st_bitflip": lambda: damaged("bitflip", seed=15, stereo=True),
}


def loss_for(off, seed=5, p=0.1):
    return ol.loss_mask(len(off), p, seed)


def oracle_keep(data, off, ln, chains, lost=None, channels=1):
    """Decode with the oracle; fill each chain's keep with the samples its
    packets produced before the first error (the part both decoders emit),
    and return (chains, oracle result)."""
    ref = ol.decode_chains(data, off, ln, chains, lost=lost, channels=channels)
    ch = np.array(chains, np.int64).copy()
    ch[:, 3] = [len(seg) // channels for seg in ref["full"]]
    return ch, ol.decode_chains(data, off, ln, ch, lost=lost, channels=channels)


def write_set(path, data, off, ln, lost, chains):
    """Binary layout read by tests/hostgate/fuzz_gate.cpp: 'CUFZ', u32 packets,
    u32 chains, u32 bytes, u32 has_lost; off[], len[], lost[] (if any),
    first[], npkts[], keep[] (u32 each), packet bytes."""
    ch = np.asarray(chains, np.int64).reshape(-1, 4)
    with open(path, "wb") as f:
        f.write(b"CUFZ")
        f.write(np.array([len(off), len(ch), data.size, lost is not None], "<u4").tobytes())
        f.write(np.asarray(off, "<u4").tobytes() + np.asarray(ln, "<u4").tobytes())
        if lost is not None:
            f.write(np.asarray(lost, np.uint8).tobytes())
        for col in (0, 1, 3):
            f.write(ch[:, col].astype("<u4").tobytes())
        f.write(np.asarray(data, np.uint8).tobytes())
