#!/usr/bin/env python3
"""Throughput benchmark of the batched Opus decode path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--mode persistent|one|plain] [--slots N] [--no-cpu]

A step is one pass of the decode path over the whole workload with its packet
bytes already resident in HBM: every chain decoded (fixed-point Opus, int16
48 kHz, preroll dropped) and resampled to float32 16 kHz into per-file device
buffers.  The metric is playable audio-seconds decoded per second (sum of
keep / 48000 per step over the step's wall time, hipDeviceSynchronize on both
sides).  Prints one JSON line.

Default workload: one 2 h CELT-only fullband mono stream, 64 kbps VBR, frame
size drawn per 1 s segment from {2.5, 5, 10, 20} ms; 7200 chains of 1 s with
250 ms preroll, all in one launch.

roofline: the decode kernel is the dominant kernel.  It is bound by
instruction issue and latency (integer/scalar serial work, no MFMA), so
"bound" is "issue"; achieved/peak/frac are still its algorithmic HBM bytes
per launch (the packet bytes each chain reads, preroll packets once per
chain, 8 B of offset/length per packet read, 2 B per playable int16 sample
written, 40 B of chain descriptor + result per chain) over the kernel's
average duration measured with HIP events on its launch stream, against
8 TB/s.  traffic comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
profiles/ for the same workload, and "issue" from the same summary's SQ
counters: per-SIMD VALU and SALU busy fractions, VALU lane utilisation, the
fractions of a wavefront's cycles issuing or waiting, wavefronts per SIMD,
LDS bank conflicts, MFMA instructions (0) and the compiler's per-lane
scratch bytes.  The resampler's own roofline (10 B per 16 kHz output) and the
log-mel front-end's (4 B per input sample + 4 B per feature) are reported
beside it.

full_calls: the drop-in entry points timed as a caller sees them (demux on
the host, upload, decode, resample, synchronise), median of 5:
decode_ogg_to_cuda_tensor on the workload's first file and decode_ogg_multi
on the many-file set.

cpu_baseline: fixed-point libopus built from the reference's vendored 1.5.2
sources the way a CPU deployment would (-O3, SSE4.1/AVX2 intrinsics; a CPU
test proves it bit-identical to the oracle build) decoding the identical
chain list with a fresh decoder per chain, plus the float32 resampler
restated in C: on the box's CPU share (16 threads for one GPU) over the
workload, and on one thread over a bounded sample; best and median of 5;
beside it (stream_1thread) the same libopus decoding the start of the first
file as one stream without chains or preroll on one thread, the per-core rate
of a CPU pool that decodes whole files.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BOX_CPU_SHARE = 16

WORKLOADS = {
    "celt_fb_64k_2h": {
        "files": ["corpus/bench/celt_fb_64k_2h.opus"],
        "desc": "CELT-only FB mono 64 kbps VBR, frame size per 1 s from {2.5,5,10,20} ms, one 2 h file, "
                "1 s chains + 250 ms preroll, one launch, f32 16 kHz out",
    },
    "uneven_50": {
        "glob": "corpus/bench/uneven_50/f*.opus",
        "desc": "the first 50 files of the uneven set (3.5 h, about 12.6k chains): mono speech, durations "
                "log-uniform 5 s-30 min, mode per file from SILK-NB/WB, hybrid-SWB/FB, CELT-FB, bitrate "
                "redrawn every 1 s in 12-64 kbps, 20 ms frames; all chains in one persistent kernel vs one "
                "launch per file",
    },
    "hybrid_fec_200k": {
        "glob": "corpus/bench/hybrid_fec_200k/f*.opus",
        "loss": (0.02, 0x0C0FFEE5),
        "desc": "200 files x 1000 s (200k chains, 55.6 h) hybrid FB 20 ms 32 kbps with in-band FEC; 2% of "
                "packets lost by a seeded mask per file, decoded from the next packet's LBRR when present, "
                "concealed otherwise; all chains in one persistent kernel, f32 16 kHz out",
    },
    "uneven_1000": {
        "glob": "corpus/bench/uneven_1000/f*.opus",
        "desc": "1000 mono speech files, durations log-uniform 5 s-30 min, mode per file from SILK-NB/WB, "
                "hybrid-SWB/FB, CELT-FB, bitrate redrawn every 1 s in 12-64 kbps (mode switches, redundancy), "
                "20 ms frames; all chains in one persistent kernel vs one launch per file",
    },
    "st_celt_fb_64k_1h": {
        "files": ["corpus/bench/st_celt_fb_64k_1h.opus"] * 2,
        "desc": "CELT-only FB stereo 64 kbps VBR (mid/side, intensity and dual stereo), frame size per 1 s from "
                "{2.5,5,10,20} ms, two copies of one 1 h file in one batch (7200 chains, the headline's chain "
                "count), 1 s chains + 250 ms preroll, one launch of the stereo kernel build, f32 16 kHz out (the "
                "one-channel render)",
    },
    "uneven_250": {
        "glob": "corpus/bench/uneven_250/f*.opus",
        "desc": "250 mono speech files, durations log-uniform 5 s-30 min (18.1 h), mode per file from "
                "SILK-NB/WB, hybrid-SWB/FB, CELT-FB, bitrate redrawn every 1 s in 12-64 kbps (mode "
                "switches, redundancy), 20 ms frames; all chains in one persistent kernel vs one launch "
                "per file",
    },
}


def _args():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="celt_fb_64k_2h", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="persistent", choices=["persistent", "one", "plain"])
    ap.add_argument("--slots", type=int, default=0, help="persistent wavefronts (0: tuned default)")
    ap.add_argument("--graph", action="store_true", help="replay the launch sequence as a HIP graph")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner on this workload first and use its best launch shape")
    ap.add_argument("--calibrate", action="store_true",
                    help="fit the work-order cost model's per-frame costs from per-chain times first")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-file comparison")
    ap.add_argument("--no-batching", action="store_true",
                    help="skip the many-file batching comparison (persistent vs one launch per file on the "
                         "uneven_50 set, when present)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-audio-s", type=float, default=7200.0,
                    help="playable audio-seconds of the workload the multi-thread CPU baseline decodes")
    ap.add_argument("--cpu-one-thread-s", type=float, default=400.0,
                    help="playable audio-seconds of the one-thread CPU sample")
    ap.add_argument("--no-logmel", action="store_true", help="skip the log-mel front-end timing")
    ap.add_argument("--no-full-calls", action="store_true", help="skip the full drop-in call timings")
    ap.add_argument("--no-stereo", action="store_true",
                    help="skip the stereo timings (when the stereo benchmark file is present)")
    ap.add_argument("--serve-qps", default="50",
                    help="comma-separated request rates for the serving run ('' to skip)")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="arrival window per rate")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (calibrates HBM byte counters)")
    return ap.parse_args()


def _log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _read_traffic(workload):
    """HBM bytes per decode launch from the committed rocprofv3 counter summary
    (a line 'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(ROOT, "profiles", "decode_%s_pmc.txt" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return int(float(line.split(":", 1)[1]))
            except ValueError:
                return None
    return None


def _time_batch(b, steps, warmup, torch):
    for _ in range(warmup):
        b.run()
        b.check()
    dec_ms, rs_ms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
        torch.cuda.synchronize()
        if b.profiled:
            d, r = b.kernel_ms()
            dec_ms.append(d)
            rs_ms.append(r)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.check()
    return wall, dec_ms, rs_ms


def _host_cpu():
    """CPU model and core counts of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def _chain_list(blobs, max_audio_s):
    """The identical chain list of the first files of the workload, up to
    max_audio_s playable seconds: merged packet table + chains + keeps."""
    import numpy as np
    from tests import oracle_lib as ol
    datas, offs, lens, chains, keeps = [], [], [], [], []
    base_b = base_p = 0
    audio = 0.0
    for blob in blobs:
        d, ch = ol.demux_chains(blob)
        data, off, ln = d.blob()
        for c in ch:
            if audio >= max_audio_s:
                break
            chains.append((c[0] + base_p, c[1], c[2], c[3]))
            audio += c[3] / 48000.0
        keeps.append(sum(c[3] for c in ch))
        datas.append(data)
        offs.append(off.astype(np.uint64) + base_b)
        lens.append(ln)
        base_b += data.size
        base_p += off.size
        if audio >= max_audio_s:
            break
    return (np.concatenate(datas), np.concatenate(offs).astype(np.uint32), np.concatenate(lens), chains,
            keeps)


def _cpu_rate(data, off, ln, chains, threads, reps=5):
    """Decode + resample the chain list on `threads` threads, `reps` times:
    (best, median) audio-s/s and the decode/resample split of the best run."""
    import numpy as np
    from tests import oracle_lib as ol
    runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ref = ol.decode_chains(data, off, ln, chains, threads=threads, fast=True)
        t1 = time.perf_counter()
        assert (ref["ret"] >= 0).all()
        ol.resample_threads(ref["pcm"], threads)
        t2 = time.perf_counter()
        runs.append((t2 - t0, t1 - t0, t2 - t1))
        secs = ref["pcm"].size / 48000.0
    walls = sorted(r[0] for r in runs)
    best = min(runs)
    return {"best": round(secs / walls[0], 1), "median": round(secs / float(np.median(walls)), 1),
            "audio_s": round(secs, 1), "decode_s": round(best[1], 3), "resample_s": round(best[2], 3)}


def _cpu_baseline(blobs, threads, max_audio_s=7200.0, one_thread_audio_s=400.0):
    """libopus on the host over the identical chain list: `threads` threads
    over up to max_audio_s of the workload, and one thread over a bounded
    sample (one_thread_audio_s); fresh decoder per chain, then the float32
    resampler.  Each best and median of 5."""
    data, off, ln, chains, _ = _chain_list(blobs, max_audio_s)
    many = _cpu_rate(data, off, ln, chains, threads)
    d1, o1, l1, c1, _ = _chain_list(blobs, one_thread_audio_s)
    one = _cpu_rate(d1, o1, l1, c1, 1)
    host = _host_cpu()
    return {
        "value": many["best"],
        "unit": "audio-s/s",
        "cores": threads,
        "kind": "reference",
        "sample": "%d chains, %.0f audio-s of the workload: fixed-point libopus 1.5.2 (-O3, SSE4.1/AVX2) decode "
                  "(%.2f s) + float32 resampler (%.2f s) on %d threads, best of 5 (median %.1f audio-s/s)"
                  % (len(chains), many["audio_s"], many["decode_s"], many["resample_s"], threads, many["median"]),
        "median": many["median"],
        "one_thread": {"value": one["best"], "median": one["median"], "unit": "audio-s/s", "cores": 1,
                       "sample": "%d chains, %.0f audio-s, best of 5" % (len(c1), one["audio_s"])},
        "host": dict(host, threads_used=threads,
                     why=("one GPU's share of the box's host CPUs is 16; threads are capped there so the "
                          "baseline is the CPU a one-GPU deployment gets")),
    }


def _cpu_stream(blob, max_audio_s=1800.0):
    """libopus decoding the start of the first file as ONE stream (one decoder,
    no chains, no preroll: what a CPU pool does per file), one thread."""
    import numpy as np
    from tests import oracle_lib as ol
    L = ol.lib()
    d, _ = ol.demux_chains(blob)
    data, off, ln = d.blob()
    total = 0
    n = 0
    limit = int(max_audio_s * 48000) + d.preskip
    while n < off.size and total < limit:
        total += L.orc_packet_nb_samples(data[int(off[n]):].ctypes.data, int(ln[n]))
        n += 1
    keep = max(total - d.preskip, 0)
    t0 = time.perf_counter()
    ref = ol.decode_chains(data, off, ln, [(0, n, d.preskip, keep)], threads=1, caps=[total + 5760], fast=True)
    t1 = time.perf_counter()
    assert int(ref["ret"][0]) == total
    secs = keep / 48000.0
    return {"value": round(secs / (t1 - t0), 1), "unit": "audio-s/s", "cores": 1, "kind": "reference",
            "sample": "first %d packets (%.0f audio-s) of the first file as one stream, no preroll, 1 thread: "
                      "%.2f s" % (n, secs, t1 - t0)}


def _read_issue(workload):
    """Issue-side counters of the decode kernel from the committed rocprofv3
    summary (lines 'issue_fractions: k=v ...' and 'issue_counters: k=v ...')."""
    path = os.path.join(ROOT, "profiles", "decode_%s_pmc.txt" % workload)
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        for tag in ("issue_fractions:", "issue_counters:"):
            if line.startswith(tag):
                for kv in line.split(":", 1)[1].split():
                    k, v = kv.split("=")
                    try:
                        out[k] = float(v)
                    except ValueError:
                        out[k] = v
    return out or None


def _median_ms(fn, torch, reps=5):
    import numpy as np
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(1000.0 * (time.perf_counter() - t0))
    return float(np.median(ts)), min(ts)


def _full_calls(cn, torch, blobs, multi_name="uneven_50"):
    """The drop-in entry points as a caller sees them (host demux, upload,
    decode, resample, synchronise), median of 5: decode_ogg_to_cuda_tensor on
    the workload's first file and decode_ogg_multi on the many-file set."""
    import glob
    out = {}
    secs = cn.probe(blobs[0]) / 16000.0
    med, best = _median_ms(lambda: cn.decode_ogg_to_cuda_tensor(blobs[0]), torch)
    out["f16k_on"] = {"audio_s": round(secs, 1), "ms_median": round(med, 2), "ms_best": round(best, 2),
                      "value": round(secs / (med / 1000.0), 1), "unit": "audio-s/s"}
    _log("full call decode_ogg_to_cuda_tensor: %.1f ms (median of 5)" % med)
    paths = sorted(glob.glob(os.path.join(ROOT, WORKLOADS[multi_name]["glob"])))
    if paths:
        mb = [open(p, "rb").read() for p in paths]
        msecs = sum(cn.probe(b) for b in mb) / 16000.0
        med, best = _median_ms(lambda: cn.decode_ogg_multi(mb), torch)
        out["f16k_multi"] = {"workload": multi_name, "files": len(mb), "audio_s": round(msecs, 1),
                             "ms_median": round(med, 2), "ms_best": round(best, 2),
                             "value": round(msecs / (med / 1000.0), 1), "unit": "audio-s/s"}
        _log("full call decode_ogg_multi (%d files): %.1f ms (median of 5)" % (len(mb), med))
    return out


def _stereo(cn, torch, steps, name="st_celt_fb_64k_1h"):
    """Stereo stream through the stereo kernel build: the resident batch
    (persistent, f32 16 kHz one-channel render; two copies of the 1 h file,
    7200 chains) timed like the headline, and the drop-in calls on the file
    for the interleaved int16 48 kHz output and the f32 16 kHz output,
    median of 5."""
    path = os.path.join(ROOT, WORKLOADS[name]["files"][0])
    if not os.path.exists(path):
        return None
    blob = open(path, "rb").read()
    b = cn.Batch([blob, blob], mode=cn.MODE_PERSISTENT, profile=True)
    wall, dec_ms, _ = _time_batch(b, steps, 1, torch)
    out = {"workload": name, "batch_files": 2, "chains": b.chains, "audio_s": round(b.playable / 48000.0, 1),
           "batch_ms_per_step": round(1000.0 * wall / steps, 3),
           "batch_value": round(b.playable / 48000.0 * steps / wall, 1), "unit": "audio-s/s",
           "decode_kernel_ms": round(sum(dec_ms) / len(dec_ms), 3) if dec_ms else None}
    b.close()
    del b
    audio_s = cn.file_info(blob)["play48"] / 48000.0  # one copy
    med, best = _median_ms(lambda: cn.decode_ogg_to_s16_tensor(blob), torch)
    out["s16_call"] = {"ms_median": round(med, 2), "ms_best": round(best, 2),
                       "value": round(audio_s / (med / 1000.0), 1), "unit": "audio-s/s", "channels": 2}
    med, best = _median_ms(lambda: cn.decode_ogg_to_cuda_tensor(blob), torch)
    out["f16k_call"] = {"ms_median": round(med, 2), "ms_best": round(best, 2),
                        "value": round(audio_s / (med / 1000.0), 1), "unit": "audio-s/s", "channels": 1}
    _log("stereo %s: batch %.1f ms/step, s16 call %.1f ms, f16k call %.1f ms"
         % (name, out["batch_ms_per_step"], out["s16_call"]["ms_median"], out["f16k_call"]["ms_median"]))
    return out


def _logmel(cn, torch, x, n_mels=80):
    """The log-mel front-end on the workload's f32 16 kHz output: median of 5
    launches timed with events on the current stream; HBM roofline with 4 B
    per input sample read and 4 B per feature written."""
    import numpy as np
    frames = cn.log_mel_frames(x.numel())
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    cn.log_mel(x, n_mels)
    ts = []
    for _ in range(5):
        ev[0].record()
        cn.log_mel(x, n_mels)
        ev[1].record()
        torch.cuda.synchronize()
        ts.append(ev[0].elapsed_time(ev[1]))
    ms = float(np.median(ts))
    alg = 4 * x.numel() + 4 * frames * n_mels
    ach = alg / (ms * 1e-3) / 1e9
    return {"kernel": "k_log_mel", "n_mels": n_mels, "frames": frames, "ms_median": round(ms, 3),
            "audio_s_per_s": round(x.numel() / 16000.0 / (ms * 1e-3), 1),
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(alg)}}


def _batching(cn, torch, steps, name="uneven_50"):
    """Many files of uneven length (config 3 shape, >= 10k chains): every chain
    in one persistent launch vs one decode + one resample launch per file."""
    import glob
    paths = sorted(glob.glob(os.path.join(ROOT, WORKLOADS[name]["glob"])))
    if not paths:
        return None
    blobs = [open(p, "rb").read() for p in paths]
    out = {"workload": name, "files": len(blobs)}
    for key, mode in (("persistent", cn.MODE_PERSISTENT), ("plain", cn.MODE_PLAIN)):
        b = cn.Batch(blobs, mode=mode)
        wall, _, _ = _time_batch(b, steps, 1, torch)
        out["chains"] = b.chains
        out["audio_s"] = round(b.playable / 48000.0, 1)
        out[key + "_value"] = round(b.playable / 48000.0 * steps / wall, 1)
        out[key + "_ms_per_step"] = round(1000.0 * wall / steps, 2)
        b.close()
        del b
        _log("batching %s %s: %.1f ms/step" % (name, key, out[key + "_ms_per_step"]))
    out["batched_vs_plain"] = round(out["persistent_value"] / out["plain_value"], 3)
    return out


def _serve(cn, torch, qps_list, seconds, seed=0x5E7E):
    """Poisson arrivals of requests drawn from the serving pool (durations
    lognormal, median 45 s, sigma 1.0, clipped 2 s-2 h; mode/bitrate mix as the
    uneven set) against the continuous-batching server.  Latency = arrival to
    the poll that returns the finished tensor."""
    import glob
    import numpy as np
    paths = sorted(glob.glob(os.path.join(ROOT, "corpus", "bench", "serve_pool", "f*.opus")))
    if not paths:
        return None
    pool = [open(p, "rb").read() for p in paths]
    secs = [cn.probe(b) / 16000.0 for b in pool]
    out = []
    rs = np.random.default_rng(seed)
    for qps in qps_list:
        srv = cn.Server(n_streams=4, mem_budget=16 << 30)
        n = max(1, int(qps * seconds))
        gaps = rs.exponential(1.0 / qps, size=n)
        arrive = np.cumsum(gaps)
        pick = rs.integers(0, len(pool), size=n)
        lat = []
        audio = 0.0
        t0 = time.perf_counter()
        k = 0
        pending = {}
        while k < n or pending:
            now = time.perf_counter() - t0
            while k < n and arrive[k] <= now:
                rid = srv.submit(pool[pick[k]])
                pending[rid] = (arrive[k], secs[pick[k]])
                k += 1
            for rid, _t, rc in srv.poll():
                if rc:
                    raise RuntimeError("request failed rc=%d" % rc)
                ta, sa = pending.pop(rid)
                lat.append(time.perf_counter() - t0 - ta)
                audio += sa
            time.sleep(0.0002)
        wall = time.perf_counter() - t0
        srv.close()
        lat = np.array(lat) * 1000.0
        out.append({"qps": qps, "requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 2),
                    "p99_ms": round(float(np.percentile(lat, 99)), 2), "audio_s_per_s": round(audio / wall, 1),
                    "offered_audio_s_per_s": round(float(np.mean(secs)) * qps, 1)})
        _log("serve qps %g: p50 %.1f ms p99 %.1f ms, %.0f audio-s/s" % (qps, out[-1]["p50_ms"], out[-1]["p99_ms"],
                                                                       out[-1]["audio_s_per_s"]))
    return out


def main():
    a = _args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only (--gpus 1): there is no multi-GPU decode path; chains would shard "
              "by file or by contiguous chain range with a 24-sample resampler halo", file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[a.workload]
    if "glob" in w:
        import glob
        paths = sorted(glob.glob(os.path.join(ROOT, w["glob"])))
        if not paths:
            sys.exit("workload input %s missing: generate it with tests.oracle_lib.bench_set()" % w["glob"])
    else:
        paths = [os.path.join(ROOT, p) for p in w["files"]]
    for p in paths:
        if not os.path.exists(p):
            sys.exit("workload input %s missing: generate it with __graft_entry__.build()" % p)
    blobs = [open(p, "rb").read() for p in paths]
    masks = None

    import torch
    from cuopus_amd import cuopus_native as cn
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    torch.cuda.set_device(0)
    cn.warmup()
    if a.copy_calib_mib:
        n = a.copy_calib_mib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    modes = {"persistent": cn.MODE_PERSISTENT, "one": cn.MODE_ONE_LAUNCH, "plain": cn.MODE_PLAIN}
    mode, slots, lpt = modes[a.mode], a.slots, 3  # 3: the library's (tuned) default order
    tuned = None
    if a.tune:
        from cuopus_amd import autotune
        tuned = autotune.tune(blobs, verbose=True)
        mode, slots, lpt = tuned["mode"], tuned["n_slots"], tuned["lpt_order"]
        a.mode = {v: k for k, v in modes.items()}[mode]
        _log("tuned: %s" % {k: tuned[k] for k in ("mode", "n_slots", "order", "rs_variant", "ms")})
    if a.calibrate:
        from cuopus_amd import autotune
        cal = autotune.calibrate(blobs, verbose=True)
        _log("calibration: %s" % json.dumps(cal))
    if "loss" in w:
        # the seeded loss mask of each file, as the parity test draws it
        from tests import oracle_lib as ol
        p_loss, seed0 = w["loss"]
        masks = [ol.loss_mask(cn.packet_count(bl)[0], p_loss, seed0 + i) for i, bl in enumerate(blobs)]
        _log("loss masks: %d of %d packets lost" % (sum(int(m.sum()) for m in masks), sum(m.size for m in masks)))
    b = cn.Batch(blobs, lost_masks=masks, mode=mode, n_slots=slots, lpt_order=lpt, use_graph=a.graph,
                 profile=not a.graph)
    _log("batch ready: %d files, %d chains, %.0f audio-s" % (len(blobs), b.chains, b.playable / 48000.0))
    wall, dec_ms, rs_ms = _time_batch(b, a.steps, a.warmup, torch)
    _log("%s: %.1f ms/step" % (a.mode, 1000.0 * wall / a.steps))
    work = None
    if a.mode != "plain":
        import numpy as np
        s0, s1 = b.chain_times()
        ends = np.sort(s1)
        work = {"span_ms": round(float(ends[-1] - s0.min()) / 1e5, 3),
                "last_1pct_of_chains_ms": round(float(ends[-1] - ends[int(0.99 * (ends.size - 1))]) / 1e5, 3),
                "chain_ms_mean": round(float((s1 - s0).mean()) / 1e5, 4),
                "chain_ms_max": round(float((s1 - s0).max()) / 1e5, 4)}
    audio_s = b.playable / 48000.0
    ms = 1000.0 * wall / a.steps
    value = audio_s * a.steps / wall

    # dominant kernel roofline (decode), resampler beside it
    alg_dec = b.chain_packet_bytes + 8 * b.chain_packets + 2 * b.playable + 40 * b.chains
    roof = None
    rs_roof = None
    if dec_ms:
        d_avg = sum(dec_ms) / len(dec_ms)
        r_avg = sum(rs_ms) / len(rs_ms)
        ach = alg_dec / (d_avg * 1e-3) / 1e9
        traffic = _read_traffic(a.workload)
        roof = {"bound": "issue", "achieved": round(ach, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 6), "traffic": traffic,
                "frac_basis": "algorithmic HBM bytes per launch over the launch time, against 8 TB/s; the "
                              "kernel is bound by instruction issue and latency (see issue)",
                "kernel": "k_decode_persistent" if a.mode == "persistent" else "k_decode_chains",
                "kernel_ms": round(d_avg, 3), "algorithmic_bytes": int(alg_dec), "issue": _read_issue(a.workload)}
        rs_bytes = 10 * b.out16
        rs_ach = rs_bytes / (r_avg * 1e-3) / 1e9
        rs_roof = {"bound": "hbm", "achieved": round(rs_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(rs_ach / HBM_PEAK_GBS, 4), "kernel": "k_resample_48to16",
                   "kernel_ms": round(r_avg, 3), "algorithmic_bytes": int(rs_bytes)}
    logmel = None
    if b.outputs and b.outputs[0].numel() and not a.no_logmel:
        logmel = _logmel(cn, torch, b.outputs[0])
        _log("log-mel: %.3f ms for %d frames" % (logmel["ms_median"], logmel["frames"]))
    b.close()
    del b

    full = None if a.no_full_calls else _full_calls(cn, torch, blobs)

    plain = None
    if not a.no_plain and a.mode != "plain":
        bp = cn.Batch(blobs, lost_masks=masks, mode=cn.MODE_PLAIN, profile=True)
        pw, _, _ = _time_batch(bp, max(1, min(a.steps, 5)), 1, torch)
        plain = round(bp.playable / 48000.0 * max(1, min(a.steps, 5)) / pw, 1)
        _log("plain: %.1f ms/step" % (1000.0 * pw / max(1, min(a.steps, 5))))
        bp.close()
        del bp

    batching = None
    if not a.no_batching and not a.workload.startswith("uneven"):
        batching = _batching(cn, torch, max(1, min(a.steps, 3)))

    stereo = None
    if not a.no_stereo and not a.workload.startswith("st_"):
        stereo = _stereo(cn, torch, max(1, min(a.steps, 3)))

    serve = None
    if a.serve_qps:
        serve = _serve(cn, torch, [float(q) for q in a.serve_qps.split(",")], a.serve_seconds)

    cpu = None
    if not a.no_cpu:
        th = a.cpu_threads or min(BOX_CPU_SHARE, os.cpu_count() or 1)
        cpu = _cpu_baseline(blobs, th, a.cpu_audio_s, a.cpu_one_thread_s)
        _log("cpu baseline: %s" % cpu["sample"])
        cpu["stream_1thread"] = _cpu_stream(blobs[0], min(a.cpu_audio_s, 1800.0))
        _log("cpu stream: %s" % cpu["stream_1thread"]["sample"])

    line = {
        "metric": "audio-sec decoded/s on 1 MI355X, bit-exact: batched vs plain; serve p99 ms",
        "value": round(value, 1),
        "unit": "audio-s/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16",
        "data": "synthetic",
        "config": {"workload": a.workload, "desc": w["desc"], "mode": a.mode, "graph": a.graph,
                   "n_slots": slots or "default", "tuned": tuned is not None,
                   "files": len(blobs), "audio_s_per_step": round(audio_s, 3)},
        "work_distribution": work,
        "plain_value": plain,
        "batched_vs_plain": round(value / plain, 3) if plain else None,
        "roofline": roof,
        "resampler_roofline": rs_roof,
        "log_mel": logmel,
        "full_calls": full,
        "cpu_baseline": cpu,
        "batching": batching,
        "stereo": stereo,
        "serve": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
