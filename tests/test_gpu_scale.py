"""GPU parity at the benchmark sizes: every packet's sample count and final
range and every int16 sample of whole workloads equal fixed-point libopus
1.5.2 (the oracle, decoding the identical chain list on the host's cores), and
the float32 16 kHz outputs are within 1e-6 of the oracle's resampler.

* silk_wb_16k_60s: the reference's CPU-runnable plain case (60 chains).
* celt_fb_64k_2h: the headline workload (7200 chains, one launch).
* celt_fb_64k_2h through the resident batch exactly as bench.py times it
  (persistent kernel, the default 4096 wavefronts, cost-model order).
* st_celt_fb_64k_1h (stereo, 3600 chains, the stereo kernel build): every
  chain at two output channels through the gate, the interleaved int16
  entry point, and the f32 16 kHz one-channel render.
* uneven_1000 / uneven_250 / uneven_50 (whichever is present): every file.
* hybrid_fec_200k (2% seeded loss, FEC or concealment): the full 200 files
  when present; otherwise its first LOSS_SUBSET files (24,000 chains) are
  generated on the spot by the corpus recipe and checked.

The drop-in entry points (and the gate used here) run the persistent kernel
with the default resident wavefronts, the shape bench.py measures.
"""
import glob
import os

import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


def _check_stream(cn, blob, lost=None, tag=""):
    import torch
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    print("case scale", tag, "packets", len(off), "chains", len(chains), flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, chains, lost=lost)
    ref = ol.decode_chains(data, off, ln, chains, lost=lost, threads=THREADS)
    assert (ref["ret"] >= 0).all()
    assert rc == 0 and not stop.any()
    assert np.array_equal(nsamp, ref["nsamp"]), tag
    bad = np.nonzero(rng != ref["rng"])[0]
    assert bad.size == 0, (tag, "final range mismatch at", bad[:10])
    diff = np.nonzero(pcm != ref["pcm"])[0]
    assert diff.size == 0, (tag, "pcm mismatch at", diff[:10])
    out = cn.decode_ogg_lossy(blob, lost) if lost is not None else cn.decode_ogg_to_cuda_tensor(blob)
    torch.cuda.synchronize()
    want = ol.resample_threads(ref["pcm"], THREADS)
    got = out.cpu().numpy()
    assert got.shape == want.shape
    md = float(np.abs(got - want).max()) if got.size else 0.0
    print("  f16k max abs diff %.3g over %d outputs" % (md, got.size), flush=True)
    assert md <= 1e-6
    return len(chains), len(off)


def _bench_file(name):
    p = os.path.join(ol.BENCH, name + ".opus")
    if not os.path.exists(p):
        pytest.skip("%s not present" % p)
    return open(p, "rb").read()


def test_plain_silk_wb_60s(cn):
    _check_stream(cn, _bench_file("silk_wb_16k_60s"), tag="silk_wb_16k_60s")


def test_celt_fb_2h_all_chains(cn):
    nch, npk = _check_stream(cn, _bench_file("celt_fb_64k_2h"), tag="celt_fb_64k_2h")
    assert nch == 7200


def test_celt_fb_2h_benchmarked_batch(cn):
    """The exact benchmarked shape: Batch(mode=persistent) with the library's
    default slot count, several runs, every f32 output against the oracle."""
    import torch
    blob = _bench_file("celt_fb_64k_2h")
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, threads=THREADS)
    want = ol.resample_threads(ref["pcm"], THREADS)
    b = cn.Batch([blob], mode=cn.MODE_PERSISTENT)
    print("case batch persistent default slots, chains", b.chains, flush=True)
    for rep in range(2):
        b.outputs[0].fill_(float("nan"))
        b.run()
        assert b.check() == [want.size]
        got = b.outputs[0].cpu().numpy()
        md = float(np.abs(got - want).max())
        print("  run %d: f16k max abs diff %.3g" % (rep, md), flush=True)
        assert md <= 1e-6
    b.close()
    torch.cuda.synchronize()


def test_stereo_1h_all_chains(cn):
    import torch
    blob = _bench_file("st_celt_fb_64k_1h")
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    assert d.channels == 2
    print("case scale st_celt_fb_64k_1h packets", len(off), "chains", len(chains), flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, chains, channels=2)
    ref = ol.decode_chains(data, off, ln, chains, threads=THREADS, channels=2)
    assert rc == 0 and not stop.any() and (ref["ret"] >= 0).all()
    assert np.array_equal(nsamp, ref["nsamp"]) and np.array_equal(rng, ref["rng"])
    diff = np.nonzero(pcm != ref["pcm"])[0]
    assert diff.size == 0, ("pcm mismatch at", diff[:10])
    s16 = cn.decode_ogg_to_s16_tensor(blob).cpu().numpy()
    assert np.array_equal(s16.reshape(-1), ref["pcm"])
    mono = ol.decode_chains(data, off, ln, chains, threads=THREADS, channels=1)
    want = ol.resample_threads(mono["pcm"], THREADS)
    out = cn.decode_ogg_to_cuda_tensor(blob)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert got.shape == want.shape
    md = float(np.abs(got - want).max())
    print("  f16k (one-channel render) max abs diff %.3g over %d outputs" % (md, got.size), flush=True)
    assert md <= 1e-6


def _set_files(name):
    files = sorted(glob.glob(os.path.join(ol.BENCH, name, "f*.opus")))
    if not files:
        pytest.skip("%s set not present" % name)
    return files


def test_uneven_set_all_files(cn):
    name = next((n for n in ("uneven_1000", "uneven_250", "uneven_50")
                 if glob.glob(os.path.join(ol.BENCH, n, "f*.opus"))), "uneven_50")
    total = 0
    for f in _set_files(name):
        total += _check_stream(cn, open(f, "rb").read(), tag=os.path.basename(f))[0]
    print(name, "chains", total)


LOSS_SUBSET = 24


@pytest.fixture(scope="module")
def loss_files():
    """The config-5 set: all 200 files when generated, else its first
    LOSS_SUBSET files (identical to the full set's: same per-file seeds),
    generated here by the corpus recipe."""
    files = sorted(glob.glob(os.path.join(ol.BENCH, "hybrid_fec_200k", "f*.opus")))
    if len(files) >= 200:
        return files
    if not os.path.exists(ol.MKCORPUS):
        pytest.skip("corpus generator not built")
    return ol.bench_set("hybrid_fec_200k", jobs=THREADS, limit=LOSS_SUBSET)


def test_hybrid_fec_loss_set(cn, loss_files):
    total = lost_total = 0
    for i, f in enumerate(loss_files):
        blob = open(f, "rb").read()
        d, _ = ol.demux_chains(blob)
        lost = ol.loss_mask(len(d.packets), 0.02, seed=0x0C0FFEE5 + i)
        lost_total += int(lost.sum())
        total += _check_stream(cn, blob, lost=lost, tag=os.path.basename(f))[0]
    print("hybrid_fec_200k chains", total, "lost packets", lost_total)
