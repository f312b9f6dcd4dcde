"""Test-side access to the CPU oracle (oracle/): corpus generation, per-chain
fixed-point libopus decode, the resampler restatement and the Python demuxer.

Only tests, the smoke check and bench.py's CPU-baseline leg use this module.
"""
import ctypes
import os
import shlex
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import demux as odemux  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
MKCORPUS = os.path.join(REF, "mkcorpus")
CACHE = os.path.join(ROOT, "corpus", "cache")

_libs = {}
P = ctypes.c_void_p


def lib(fast=False):
    """The oracle library; fast=True: the same code over the -O3 intrinsics
    build of libopus (the timed CPU baseline; bit-identity is a CPU test)."""
    if fast not in _libs:
        path = os.path.join(REF, "liboracle_fast.so" if fast else "liboracle.so")
        if not os.path.exists(path):
            raise OSError("oracle not built: %s (run __graft_entry__.build())" % path)
        L = ctypes.CDLL(path)
        L.orc_decode_chain.restype = ctypes.c_long
        L.orc_decode_chain.argtypes = [P, P, P, P, ctypes.c_int, P, ctypes.c_long, P, P]
        L.orc_decode_chains.restype = ctypes.c_int
        L.orc_decode_chains.argtypes = [P, P, P, P, P, P, ctypes.c_int, P, P, P, P, P, P, ctypes.c_int]
        L.orc_decode_chains_cfg.restype = ctypes.c_int
        L.orc_decode_chains_cfg.argtypes = [P, P, P, P, P, P, ctypes.c_int, P, P, P, P, P, P, ctypes.c_int,
                                            ctypes.c_int, P]
        L.orc_resample.restype = None
        L.orc_resample.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64]
        L.orc_resample_range.restype = None
        L.orc_resample_range.argtypes = [P, ctypes.c_int64, P, ctypes.c_int64, ctypes.c_int64]
        L.orc_taps.restype = None
        L.orc_taps.argtypes = [P]
        L.orc_ec_schedule.restype = None
        L.orc_ec_schedule.argtypes = [P, ctypes.c_uint32, P]
        L.orc_packet_has_lbrr.restype = ctypes.c_int
        L.orc_packet_has_lbrr.argtypes = [P, ctypes.c_int]
        L.orc_packet_nb_samples.restype = ctypes.c_int
        L.orc_packet_nb_samples.argtypes = [P, ctypes.c_int]
        _libs[fast] = L
    return _libs[fast]


def small_corpus(name, args):
    """Generate (once) a corpus file with oracle/_ref/mkcorpus and return its path."""
    os.makedirs(CACHE, exist_ok=True)
    path = os.path.join(CACHE, name + ".opus")
    if not os.path.exists(path):
        tmp = path + ".tmp%d" % os.getpid()
        subprocess.run([MKCORPUS, "--out", tmp] + shlex.split(args), check=True)
        os.replace(tmp, path)
    return path


def demux_chains(blob, seconds=1.0, preroll_ms=250.0):
    d = odemux.demux(blob)
    return d, odemux.build_chains(d, seconds, preroll_ms)


def decode_chains(data, off, ln, chains, lost=None, threads=1, caps=None, channels=1, gain=0, fast=False):
    """Decode every chain with a fresh fixed-point libopus decoder of
    `channels` output channels and OPUS_SET_GAIN(gain) (Q8 dB; an int for
    every chain or one per chain).

    caps: PCM capacity per chain in samples per channel (default 5760 per
    packet + 5760).  Returns dict: pcm (compacted keep windows, int16,
    interleaved for 2 channels), nsamp/rng per packet, ret per chain, full
    (list of per-chain full PCM arrays)."""
    L = lib(fast)
    ch = np.asarray(chains, dtype=np.int64).reshape(-1, 4)
    n = ch.shape[0]
    first = np.ascontiguousarray(ch[:, 0], dtype=np.uint32)
    npk = np.ascontiguousarray(ch[:, 1], dtype=np.uint32)
    if caps is None:
        caps = (npk.astype(np.uint64) + 1) * 5760
    caps = np.ascontiguousarray(caps, dtype=np.uint64)
    pcm_off = np.zeros(n, np.uint64)
    if n > 1:
        pcm_off[1:] = np.cumsum(caps[:-1])
    pcm = np.zeros(int(caps.sum()) * channels, np.int16)
    nsamp = np.zeros(len(off), np.int32)
    rng = np.zeros(len(off), np.uint32)
    ret = np.zeros(n, np.int64)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint32)
    ln = np.ascontiguousarray(ln, dtype=np.uint32)
    lp = None
    if lost is not None:
        lost = np.ascontiguousarray(lost, dtype=np.uint8)
        lp = lost.ctypes.data
    gains = np.ascontiguousarray(np.broadcast_to(np.asarray(gain, dtype=np.int32), (max(n, 1),)))
    L.orc_decode_chains_cfg(data.ctypes.data, off.ctypes.data, ln.ctypes.data, lp, first.ctypes.data,
                            npk.ctypes.data, n, pcm_off.ctypes.data, caps.ctypes.data, pcm.ctypes.data,
                            nsamp.ctypes.data, rng.ctypes.data, ret.ctypes.data, threads, channels,
                            gains.ctypes.data)
    kept = []
    full = []
    C = channels
    for i in range(n):
        r = int(ret[i])
        if r < 0:
            # a failed chain keeps what its packets produced before the error (a
            # decoded packet always yields >= 120 samples; records stop there)
            f0, np_ = int(first[i]), int(npk[i])
            r = 0
            for k in range(np_):
                if nsamp[f0 + k] <= 0:
                    break
                r += int(nsamp[f0 + k])
        seg = pcm[C * int(pcm_off[i]): C * (int(pcm_off[i]) + r)]
        full.append(seg)
        trim, keep = int(ch[i, 2]), int(ch[i, 3])
        kept.append(seg[C * trim: C * (trim + keep)])
    return {"pcm": np.concatenate(kept) if kept else np.zeros(0, np.int16), "nsamp": nsamp, "rng": rng,
            "ret": ret, "full": full}


def resample(pcm48):
    """The reference resampler restated in C (float32, tap order)."""
    pcm48 = np.ascontiguousarray(pcm48, dtype=np.int16)
    n16 = odemux.out16_len(len(pcm48))
    out = np.zeros(n16, np.float32)
    if n16:
        lib().orc_resample(pcm48.ctypes.data, len(pcm48), out.ctypes.data, n16)
    return out


def resample_threads(pcm48, threads, chunk=1 << 20):
    """orc_resample split over threads by output ranges (identical output)."""
    from concurrent.futures import ThreadPoolExecutor
    pcm48 = np.ascontiguousarray(pcm48, dtype=np.int16)
    n16 = odemux.out16_len(len(pcm48))
    out = np.zeros(n16, np.float32)
    L = lib()
    ranges = [(o, min(n16, o + chunk)) for o in range(0, n16, chunk)]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda r: L.orc_resample_range(pcm48.ctypes.data, len(pcm48), out.ctypes.data, r[0], r[1]),
                    ranges))
    return out


def taps():
    t = np.zeros(49, np.float32)
    lib().orc_taps(t.ctypes.data)
    return t


def loss_mask(n, p, seed):
    """Seeded Bernoulli(p) loss mask over n packets."""
    rs = np.random.default_rng(seed)
    return (rs.random(n) < p).astype(np.uint8)


# Corpus recipes shared by the tests (small, seconds to generate)
RECIPES = {
    "silk_wb": "--seed 11 --dur 12 --signal speech --app voip --bw wb --force silk --vbr 0 --bitrate 16000",
    "silk_nb": "--seed 12 --dur 12 --signal speech --app voip --bw nb --force silk --bitrate 12000",
    "silk_mb": "--seed 13 --dur 12 --signal speech --app voip --bw mb --force silk --bitrate 14000",
    "celt_fb": "--seed 14 --dur 12 --signal music --app lowdelay --bw fb --bitrate 64000 --frameset 120,240,480,960",
    "hybrid_fb": "--seed 15 --dur 12 --signal speech --app voip --bw fb --force hybrid --bitrate 32000",
    "hybrid_swb": "--seed 16 --dur 12 --signal speech --app voip --bw swb --force hybrid --bitrate 24000",
    "hybrid_fec": "--seed 17 --dur 12 --signal speech --app voip --bw fb --force hybrid --bitrate 32000 "
                  "--fec 1 --loss 2",
    "mixed": "--seed 18 --dur 20 --signal mixed --app audio --bw fb --brange 12000 64000",
    "silk_long": "--seed 19 --dur 12 --signal speech --app voip --bw wb --force silk --bitrate 20000 --frame 2880",
    # high-rate CELT: packets above 255 bytes (multi-segment lacing, pages split inside packets)
    "celt_hi": "--seed 20 --dur 8 --signal music --app lowdelay --bw fb --bitrate 256000 --frameset 120,240,480,960",
    # CELT-only narrowband and wideband (TOC configs 16-23)
    "celt_nb": "--seed 21 --dur 8 --signal speech --app lowdelay --bw nb --bitrate 24000 --frameset 120,240,480,960",
    "celt_wb": "--seed 22 --dur 8 --signal mixed --app lowdelay --bw wb --bitrate 32000 --frame 480",
    # multi-frame packets: CELT 60 and 120 ms (code 3, several 20 ms frames per packet), SILK 40 ms,
    # hybrid 10 ms frames
    "celt_multi": "--seed 23 --dur 8 --signal music --app lowdelay --bw fb --bitrate 48000 --frameset 2880,5760",
    "silk_40ms": "--seed 24 --dur 8 --signal speech --app voip --bw wb --force silk --bitrate 18000 --frame 1920",
    "hybrid_10ms": "--seed 25 --dur 8 --signal speech --app voip --bw swb --force hybrid --bitrate 28000 --frame 480",
    # discontinuous transmission over silent stretches: TOC-only packets (0-byte frames) decoded as
    # concealment / comfort noise
    "silk_dtx": "--seed 26 --dur 10 --signal speech --app voip --bw wb --force silk --bitrate 16000 --dtx 1 --gaps 0.5",
    "mixed_dtx": "--seed 27 --dur 10 --signal speech --app voip --bw fb --brange 12000 40000 --dtx 1 --gaps 0.5",
    # rate extremes: CELT at 6 kbps (most bands skipped or folded), CELT at 510 kbps (packets up to
    # 1275 bytes), SILK WB at 40 kbps CBR (large pulse counts, LSB escapes)
    "celt_lo": "--seed 28 --dur 8 --signal music --app lowdelay --bw fb --bitrate 6000 --frameset 120,240,480,960",
    "celt_max": "--seed 29 --dur 4 --signal music --app lowdelay --bw fb --bitrate 510000 --vbr 0 "
                "--frameset 120,240,480,960",
    "silk_hi": "--seed 30 --dur 8 --signal speech --app voip --bw wb --force silk --vbr 0 --bitrate 40000",
    # lengths at the chain protocol's edges: shorter than one chain, exactly one chain, one sample past it
    "len_short": "--seed 37 --dur 0.1 --signal speech --app audio --bw fb --bitrate 32000",
    "len_1s": "--seed 38 --dur 1 --signal mixed --app audio --bw fb --bitrate 32000",
    "len_1s_plus": "--seed 39 --dur 1.0000208333 --signal music --app lowdelay --bw fb --bitrate 48000",
}


# OpusHead output gain (Q7.8 dB): applied as libopus OPUS_SET_GAIN, with saturation
GAIN_RECIPES = {
    "gain_up6": "--seed 40 --dur 6 --signal speech --app voip --bw wb --force silk --bitrate 20000 --gain 1536",
    "gain_down10": "--seed 41 --dur 6 --signal music --app lowdelay --bw fb --bitrate 64000 "
                   "--frameset 120,240,480,960 --gain -2560",
    "gain_mixed": "--seed 42 --dur 8 --signal mixed --app audio --bw fb --brange 12000 64000 --gain 700",
    # a large boost: many samples saturate at +-32767
    "gain_clip": "--seed 43 --dur 4 --signal music --app audio --bw fb --bitrate 48000 --gain 6000",
}


# Stereo streams (OpusHead channels 2): CELT at high rate (dual stereo) and low rate (intensity
# stereo, phase inversion, some mono packets), SILK mid/side with mid-only frames, SILK at 10 kbps
# (coded as mono packets throughout), hybrid, mode switching with mono packets inside the stereo
# stream, SILK switching between stereo and mono packets, FEC + loss, and the output gain
ST_RECIPES = {
    "st_celt": "--seed 50 --dur 6 --signal music --app lowdelay --bw fb --bitrate 96000 "
               "--frameset 120,240,480,960 --channels 2",
    "st_celt_lo": "--seed 51 --dur 6 --signal music --app lowdelay --bw fb --bitrate 24000 "
                  "--frameset 120,240,480,960 --channels 2",
    "st_silk": "--seed 52 --dur 6 --signal speech --app voip --bw wb --force silk --bitrate 24000 --channels 2",
    "st_silk_lo": "--seed 53 --dur 6 --signal speech --app voip --bw nb --force silk --bitrate 10000 --channels 2",
    "st_hybrid": "--seed 54 --dur 6 --signal speech --app voip --bw fb --force hybrid --bitrate 40000 --channels 2",
    "st_mixed": "--seed 55 --dur 10 --signal mixed --app audio --bw fb --brange 12000 96000 --channels 2 "
                "--monoflip 0.3",
    "st_fec": "--seed 56 --dur 8 --signal speech --app voip --bw fb --force hybrid --bitrate 32000 --fec 1 "
              "--loss 5 --channels 2",
    "st_silk_flip": "--seed 57 --dur 8 --signal speech --app voip --bw wb --force silk --bitrate 20000 "
                    "--channels 2 --monoflip 0.3",
    "st_gain": "--seed 58 --dur 6 --signal mixed --app audio --bw fb --brange 16000 80000 --channels 2 "
               "--gain -900",
}


# More seeds of the mode-switching mixes (SILK/hybrid/CELT transitions, redundancy, FEC): decoded with
# seeded loss by one CPU-gate test and one GPU test each
SEED_SWEEP = {
    "sweep_%d" % sd: "--seed %d --dur 10 --signal mixed --app %s --bw %s --brange 10000 72000%s"
                     % (sd, app, bw, " --fec 1 --loss 4" if sd % 2 else "")
    for sd, app, bw in ((31, "audio", "fb"), (32, "voip", "fb"), (33, "audio", "swb"), (34, "voip", "wb"),
                        (35, "audio", "fb"), (36, "voip", "swb"))
}


def recipe(name):
    for table in (RECIPES, SEED_SWEEP, GAIN_RECIPES, ST_RECIPES):
        if name in table:
            return table[name]
    raise KeyError(name)


# Benchmark inputs (minutes to generate; written once under corpus/bench)
BENCH = os.path.join(ROOT, "corpus", "bench")
BENCH_RECIPES = {
    "celt_fb_64k_2h": "--seed 0x0C0FFEE2 --dur 7200 --signal music --app lowdelay --bw fb --vbr 1 --bitrate 64000 "
                      "--frameset 120,240,480,960",
    # a stereo stream for the stereo kernel build's timings: CELT FB 64 kbps, 1 h (3600 chains)
    "st_celt_fb_64k_1h": "--seed 0x0C0FFEE6 --dur 3600 --signal music --app lowdelay --bw fb --vbr 1 "
                         "--bitrate 64000 --frameset 120,240,480,960 --channels 2",
    # the reference's CPU-runnable plain case: 60 s SILK-only WB speech, 16 kbps CBR, 20 ms
    "silk_wb_16k_60s": "--seed 0x0C0FFEE1 --dur 60 --signal speech --app voip --bw wb --force silk --vbr 0 "
                       "--bitrate 16000 --frame 960",
}


def bench_corpus(name):
    """Generate (once) a benchmark input with oracle/_ref/mkcorpus and return its path."""
    os.makedirs(BENCH, exist_ok=True)
    path = os.path.join(BENCH, name + ".opus")
    if not os.path.exists(path):
        tmp = path + ".tmp%d" % os.getpid()
        subprocess.run([MKCORPUS, "--out", tmp] + shlex.split(BENCH_RECIPES[name]), check=True)
        os.replace(tmp, path)
    return path


# Multi-file benchmark sets: many streams of uneven length (log-uniform
# duration), each with a mode drawn per file; bitrate redrawn every second in
# [12, 64] kbps and the mode not forced, so the encoder switches SILK / hybrid /
# CELT and emits redundancy at the switches.  20 ms frames, speech-like signal.
BENCH_SET_MODES = {
    "silk_nb": "--app voip --bw nb",
    "silk_wb": "--app voip --bw wb",
    "hybrid_swb": "--app voip --bw swb",
    "hybrid_fb": "--app voip --bw fb",
    "celt_fb": "--app lowdelay --bw fb",
}
BENCH_SETS = {
    # the batching config at its stated size: 1000 files, durations exp(U(ln 5, ln 1800)) s
    "uneven_1000": dict(n=1000, seed=0x0C0FFEE3, dmin=5.0, dmax=1800.0),
    # its first 250 files (identical files: same per-file seeds)
    "uneven_250": dict(n=250, seed=0x0C0FFEE3, dmin=5.0, dmax=1800.0),
    # the first 50 files of uneven_250 (identical files: same per-file seeds),
    # about 12.6k chains: small enough to ship with every GPU run
    "uneven_50": dict(n=50, seed=0x0C0FFEE3, dmin=5.0, dmax=1800.0),
    # serving request pool: durations lognormal (median 45 s, sigma 1.0)
    # clipped to [2 s, 7200 s]
    "serve_pool": dict(n=60, seed=0x0C0FFEE4, dmin=2.0, dmax=7200.0, median=45.0, sigma=1.0),
    # loss workload: 200 files x 1000 s hybrid FB 20 ms 32 kbps with in-band FEC (200k chains)
    "hybrid_fec_200k": dict(n=200, seed=0x0C0FFEE5, fixed_dur=1000.0,
                            args="--signal speech --app voip --bw fb --force hybrid --bitrate 32000 --frame 960 "
                                 "--fec 1 --loss 2"),
}


def bench_set_plan(name):
    cfg = BENCH_SETS[name]
    plan = []
    modes = sorted(BENCH_SET_MODES)
    for i in range(cfg["n"]):
        rs = np.random.default_rng(cfg["seed"] + i)
        if "fixed_dur" in cfg:
            plan.append(("f%04d" % i, cfg["fixed_dur"], "hybrid_fb_fec",
                         "--seed %d --dur %.3f %s" % (cfg["seed"] + i, cfg["fixed_dur"], cfg["args"])))
            continue
        if "median" in cfg:
            dur = float(np.clip(cfg["median"] * np.exp(cfg["sigma"] * rs.standard_normal()), cfg["dmin"],
                                cfg["dmax"]))
        else:
            dur = float(np.exp(rs.uniform(np.log(cfg["dmin"]), np.log(cfg["dmax"]))))
        mode = modes[int(rs.integers(0, len(modes)))]
        args = "--seed %d --dur %.3f --signal speech %s --brange 12000 64000 --frame 960" % (
            cfg["seed"] + i, dur, BENCH_SET_MODES[mode])
        plan.append(("f%04d" % i, round(dur, 3), mode, args))
    return plan


def bench_set(name, jobs=8, limit=None, progress=False):
    """Generate (once) a multi-file benchmark set under corpus/bench/<name>/
    with a manifest.json; returns the list of file paths.  limit: only the
    first `limit` files of the set (the same files the full set has)."""
    import json
    from concurrent.futures import ThreadPoolExecutor
    d = os.path.join(BENCH, name)
    os.makedirs(d, exist_ok=True)
    plan = bench_set_plan(name)
    if limit is not None:
        plan = plan[:limit]
    done = [0]

    def one(item):
        fname, _, _, args = item
        path = os.path.join(d, fname + ".opus")
        if not os.path.exists(path):
            tmp = path + ".tmp%d" % os.getpid()
            subprocess.run([MKCORPUS, "--out", tmp] + shlex.split(args), check=True)
            os.replace(tmp, path)
        done[0] += 1
        if progress and done[0] % 25 == 0:
            print("[corpus] %s: %d/%d files" % (name, done[0], len(plan)), flush=True)
        return path

    # longest first so the pool finishes together
    order = sorted(plan, key=lambda it: -it[1])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(one, order))
    with open(os.path.join(d, "manifest.json"), "w") as f:
        json.dump([{"file": p[0] + ".opus", "dur": p[1], "mode": p[2], "args": p[3]} for p in plan], f, indent=0)
    return [os.path.join(d, p[0] + ".opus") for p in plan]
