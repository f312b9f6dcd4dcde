"""Autotuner for the batched decoder's launch parameters.

For a workload (a list of Ogg/Opus streams) it times the resident batch over
a grid of launch shapes -- the persistent kernel with a given number of
resident wavefronts (two to four per SIMD), with and
without longest-chain-first ordering, and the one-workgroup-per-chain launch
-- checks that every shape produces bit-identical outputs, and keeps the
fastest.  Results are cached per device name and workload size class in a
JSON file so later runs reuse them.

    from cuopus_amd import autotune
    best = autotune.tune(blobs)          # {"mode": ..., "n_slots": ..., "lpt_order": ..., "ms": ...}
"""
import json
import math
import os
import time

from . import cuopus_native as cn

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")

# (mode, n_slots, lpt_order)
DEFAULT_GRID = [
    (cn.MODE_PERSISTENT, 2048, True),
    (cn.MODE_PERSISTENT, 3072, True),
    (cn.MODE_PERSISTENT, 4096, True),
    (cn.MODE_PERSISTENT, 4096, False),
    (cn.MODE_ONE_LAUNCH, 0, True),
]


def _device_name():
    import torch
    return torch.cuda.get_device_name(0)


def signature(chains):
    """Size class of a workload: chain count rounded to a power of two."""
    return "chains~%d" % (1 << max(0, int(round(math.log2(max(1, chains))))))


def _load_cache(path):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(chains, path=CACHE):
    """Cached best shape for this device and workload size class, or None."""
    try:
        return _load_cache(path).get(_device_name(), {}).get(signature(chains))
    except Exception:
        return None


def _time(blobs, mode, n_slots, lpt, reps, ref):
    import torch
    b = cn.Batch(blobs, mode=mode, n_slots=n_slots, lpt_order=lpt, profile=True)
    b.run()
    b.check()
    ms = []
    for _ in range(reps):
        b.run()
        torch.cuda.synchronize()
        ms.append(sum(b.kernel_ms()))
    b.check()
    same = True
    if ref is not None:
        same = all(torch.equal(x, y) for x, y in zip(b.outputs, ref))
        outs = None
    else:
        outs = [o.clone() for o in b.outputs]
    chains = b.chains
    b.close()
    return min(ms), same, outs, chains


def tune(blobs, grid=None, reps=2, path=CACHE, verbose=False):
    """Time every shape of the grid on this workload; returns the best shape
    and stores it in the cache.  Raises if any shape's output differs."""
    grid = grid or DEFAULT_GRID
    ref = None
    results = []
    chains = 0
    for (mode, n_slots, lpt) in grid:
        ms, same, outs, chains = _time(blobs, mode, n_slots, lpt, reps, ref)
        if ref is None:
            ref = outs
        if not same:
            raise RuntimeError("launch shape %s produced different output" % ((mode, n_slots, lpt),))
        results.append({"mode": mode, "n_slots": n_slots, "lpt_order": lpt, "ms": round(ms, 3)})
        if verbose:
            print("[tune] mode %d slots %d lpt %d: %.2f ms" % (mode, n_slots, lpt, ms), flush=True)
    best = min(results, key=lambda r: r["ms"])
    best = dict(best, tuned_at=time.strftime("%Y-%m-%d %H:%M:%S"), grid=results)
    try:
        cache = _load_cache(path)
        cache.setdefault(_device_name(), {})[signature(chains)] = best
        with open(path + ".tmp", "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(path + ".tmp", path)
    except OSError:
        pass
    return best
