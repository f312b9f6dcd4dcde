"""Autotuner of the batched decoder's launch parameters, and the calibration
of its work-order cost model.

Axes (times measured with HIP events on the launch stream, min of `reps`):

* resident wavefronts of the persistent decode kernel: 1024 .. 4096 (four per
  SIMD is the chip's limit at 10 KB of LDS and <= 128 VGPRs per wavefront),
  and the one-workgroup-per-chain launch;
* work order: file order, costliest first by the per-packet cost model, or
  longest first by samples and packets;
* resampler launch shape: threads per workgroup x outputs per thread
  (256x4, 256x8, 128x8, 512x4, 128x4).

Chains per wavefront is not an axis: the decoder runs one chain per
wavefront with its serial control flow in scalar registers; two chains per
wavefront would run both chains' divergent control flow one after the other
on 32 lanes each and need two LDS blocks per wavefront (two wavefronts per
SIMD instead of four), so it does less work per SIMD, not more.

Every shape must give bit-identical outputs.  The best shape is stored per
device in cuopus_tuned.txt beside libcuopus.so, which the library reads for
its defaults (batches, the server and the drop-in entry points), and with
the whole grid in tuned.json.

    from cuopus_amd import autotune
    best = autotune.tune(blobs)
    costs = autotune.calibrate(blobs)   # per-frame costs of the work-order model
"""
import json
import math
import os
import time

import numpy as np

from . import cuopus_native as cn

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned.json")
TUNED_TXT = os.path.join(HERE, "cuopus_tuned.txt")

SLOTS = (1024, 2048, 3072, 4096)
ORDERS = (0, 1, 2)
RS_VARIANTS = (0, 1, 2, 3, 4)
RS_NAMES = {0: "256x4", 1: "256x8", 2: "128x8", 3: "512x4", 4: "128x4"}


def device_key():
    """'<gfx arch>/<CU count>', the key of cuopus_tuned.txt."""
    import torch
    p = torch.cuda.get_device_properties(0)
    return "%s/%d" % (p.gcnArchName.split(":")[0], p.multi_processor_count)


def signature(chains):
    """Size class of a workload: chain count rounded to a power of two."""
    return "chains~%d" % (1 << max(0, int(round(math.log2(max(1, chains))))))


def _load_json(path):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(chains, path=CACHE):
    """Cached best shape for this device and workload size class, or None."""
    try:
        return _load_json(path).get(device_key(), {}).get(signature(chains))
    except Exception:
        return None


def _run(blobs, reps, ref, **kw):
    import torch
    b = cn.Batch(blobs, profile=True, **kw)
    b.run()
    b.check()
    dec, rs = [], []
    for _ in range(reps):
        b.run()
        torch.cuda.synchronize()
        d, r = b.kernel_ms()
        dec.append(d)
        rs.append(r)
    b.check()
    outs = [o.clone() for o in b.outputs] if ref is None else None
    same = ref is None or all(torch.equal(x, y) for x, y in zip(b.outputs, ref))
    chains = b.chains
    b.close()
    return min(dec), min(rs), same, outs, chains


def write_tuned(best, path=TUNED_TXT):
    """Store this device's defaults in cuopus_tuned.txt (one line per device)."""
    key = device_key()
    lines = []
    if os.path.exists(path):
        lines = [ln for ln in open(path).read().splitlines() if ln.strip() and not ln.startswith(key + " ")]
    lines.append("%s slots=%d rs=%d order=%d" % (key, best["n_slots"], best["rs_variant"], best["order"]))
    with open(path + ".tmp", "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(path + ".tmp", path)


def tune(blobs, reps=2, path=CACHE, write=True, verbose=False):
    """Time the decode shapes (slots x order, and one launch) and the
    resampler shapes on this workload; returns the best and stores it.
    Raises if any shape's output differs from the first one's."""
    ref = None
    grid = []
    chains = 0

    def log(msg):
        if verbose:
            print("[tune] " + msg, flush=True)

    for slots in SLOTS:
        for order in ORDERS:
            d, r, same, outs, chains = _run(blobs, reps, ref, mode=cn.MODE_PERSISTENT, n_slots=slots,
                                            lpt_order=order, rs_variant=0)
            ref = ref if ref is not None else outs
            if not same:
                raise RuntimeError("persistent slots %d order %d produced different output" % (slots, order))
            grid.append({"mode": cn.MODE_PERSISTENT, "n_slots": slots, "order": order, "decode_ms": round(d, 3)})
            log("persistent slots %d order %d: decode %.2f ms" % (slots, order, d))
    d, r, same, _, _ = _run(blobs, reps, ref, mode=cn.MODE_ONE_LAUNCH, lpt_order=1, rs_variant=0)
    if not same:
        raise RuntimeError("one-launch mode produced different output")
    grid.append({"mode": cn.MODE_ONE_LAUNCH, "n_slots": 0, "order": 1, "decode_ms": round(d, 3)})
    log("one workgroup per chain: decode %.2f ms" % d)
    best = min((g for g in grid if g["mode"] == cn.MODE_PERSISTENT), key=lambda g: g["decode_ms"])
    rs_grid = []
    for v in RS_VARIANTS:
        _, r, same, _, _ = _run(blobs, reps, ref, mode=cn.MODE_PERSISTENT, n_slots=best["n_slots"],
                                lpt_order=best["order"], rs_variant=v)
        if not same:
            raise RuntimeError("resampler shape %s produced different output" % RS_NAMES[v])
        rs_grid.append({"rs_variant": v, "shape": RS_NAMES[v], "resample_ms": round(r, 4)})
        log("resampler %s: %.3f ms" % (RS_NAMES[v], r))
    rs_best = min(rs_grid, key=lambda g: g["resample_ms"])
    out = {"mode": cn.MODE_PERSISTENT, "n_slots": best["n_slots"], "order": best["order"],
           "lpt_order": best["order"], "rs_variant": rs_best["rs_variant"], "decode_ms": best["decode_ms"],
           "resample_ms": rs_best["resample_ms"], "ms": round(best["decode_ms"] + rs_best["resample_ms"], 3),
           "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S"), "grid": grid, "rs_grid": rs_grid,
           "chains_per_wavefront": 1}
    if write:
        try:
            cache = _load_json(path)
            cache.setdefault(device_key(), {})[signature(chains)] = out
            with open(path + ".tmp", "w") as f:
                json.dump(cache, f, indent=1)
            os.replace(path + ".tmp", path)
            write_tuned(out)
        except OSError:
            pass
    return out


# ---------------------------------------------------------------- cost model
CLASSES = ["celt_2.5", "celt_5", "celt_10", "celt_20", "hybrid_10", "hybrid_20", "silk_10", "silk_20", "silk_40",
           "silk_60"]


def packet_class(p):
    """(class index, frames) of a packet from its TOC (the classes of the
    library's per-packet cost model)."""
    toc = p[0]
    code = toc & 3
    frames = 1 if code == 0 else (2 if code < 3 else (p[1] & 0x3F if len(p) > 1 else 1))
    if toc & 0x80:
        return (toc >> 3) & 3, frames
    if (toc & 0x60) == 0x60:
        return 4 + ((toc >> 3) & 1), frames
    return 6 + ((toc >> 3) & 3), frames


def calibrate(blobs, n_slots=0, verbose=False):
    """Fit per-frame decode costs (device clock ticks of one wavefront) by
    non-negative least squares of each chain's measured time on the frame
    counts of its packets; returns {class: ticks per frame} and the fit's
    relative residual."""
    import torch
    from scipy.optimize import nnls
    b = cn.Batch(blobs, mode=cn.MODE_PERSISTENT, n_slots=n_slots, lpt_order=1)
    b.run()
    b.check()
    b.run()
    torch.cuda.synchronize()
    s, e = b.chain_times()
    b.close()
    rows = []
    for blob in blobs:
        data, lens = cn.packets(blob)
        offs = np.concatenate([[0], np.cumsum(lens.astype(np.int64))[:-1]])
        for first, npk, _, _ in cn.chain_table(blob):
            row = np.zeros(len(CLASSES))
            for k in range(int(first), int(first) + int(npk)):
                if lens[k] == 0:
                    continue
                cls, fr = packet_class(data[offs[k]:offs[k] + lens[k]])
                row[cls] += fr
            rows.append(row)
    A = np.array(rows)
    y = (e - s).astype(np.float64)
    used = A.sum(axis=0) > 0
    x, res = nnls(A[:, used], y)
    costs = dict(zip([c for c, u in zip(CLASSES, used) if u], x))
    rel = float(np.linalg.norm(A[:, used] @ x - y) / np.linalg.norm(y))
    if verbose:
        for c, v in costs.items():
            print("[calibrate] %-10s %8.1f ticks (%.2f us) per frame" % (c, v, v / 100.0), flush=True)
        print("[calibrate] relative residual %.3f over %d chains" % (rel, len(y)), flush=True)
    return {"ticks_per_frame": costs, "relative_residual": rel, "chains": int(len(y)),
            "tail_ms": float((e.max() - np.sort(e)[int(0.99 * len(e))]) / 1e5) if len(e) else 0.0,
            "span_ms": float((e.max() - s.min()) / 1e5) if len(e) else 0.0}
