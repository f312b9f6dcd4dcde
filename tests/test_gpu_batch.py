"""GPU: the resident batch (plain per-file launches, one launch, persistent
work queue, HIP-graph replays) gives exactly the single-file path's output
for every file, and that output matches the fixed-point oracle.

Bar: identical float32 outputs between the batch modes and the single-file
path (same int16 PCM, same resampler); oracle resampler within 1e-6 abs.
"""
import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

RS_TOL = 1e-6
NAMES = ["silk_wb", "celt_fb", "hybrid_fb", "mixed", "silk_nb", "hybrid_fec"]


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


def _blobs(names):
    return [open(ol.small_corpus(n, ol.RECIPES[n]), "rb").read() for n in names]


def _oracle_f16k(blob, lost=None):
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    return ol.resample(ref["pcm"])


@pytest.mark.parametrize("mode,graph", [(0, False), (1, False), (2, False), (2, True), (1, True)])
def test_batch_modes_match_single_file(cn, mode, graph):
    import torch
    blobs = _blobs(NAMES)
    print("case batch mode", mode, "graph", graph, flush=True)
    singles = [cn.decode_ogg_to_cuda_tensor(b) for b in blobs]
    b = cn.Batch(blobs, mode=mode, n_slots=7 if mode == 2 else 0, use_graph=graph)
    for rep in range(3):
        for o in b.outputs:
            o.fill_(float("nan"))
        # eager work of the program between replays
        torch.cuda.synchronize()
        b.run()
        lens = b.check()
        for i, (o, s) in enumerate(zip(b.outputs, singles)):
            assert lens[i] == s.numel()
            assert torch.equal(o, s), (NAMES[i], rep)
    for i, blob in enumerate(blobs[:3]):
        want = _oracle_f16k(blob)
        got = b.outputs[i].cpu().numpy()
        assert got.shape == want.shape
        assert float(np.abs(got - want).max()) <= RS_TOL
    b.close()


def test_batch_with_loss_masks(cn):
    import torch
    names = ["hybrid_fec", "silk_wb", "celt_fb"]
    blobs = _blobs(names)
    masks = []
    for n, blob in zip(names, blobs):
        d, _ = ol.demux_chains(blob)
        masks.append(ol.loss_mask(len(d.packets), 0.05, seed=sum(n.encode())))
    masks[1] = None  # a file without losses in the same batch
    b = cn.Batch(blobs, lost_masks=masks, mode=2, n_slots=5)
    b.run()
    b.check()
    for i, blob in enumerate(blobs):
        want = _oracle_f16k(blob, masks[i])
        got = b.outputs[i].cpu().numpy()
        assert got.shape == want.shape
        assert float(np.abs(got - want).max()) <= RS_TOL, names[i]
        if masks[i] is not None:
            single = cn.decode_ogg_lossy(blob, masks[i])
            assert torch.equal(single, b.outputs[i])
    b.close()


def test_batch_profile_events(cn):
    blobs = _blobs(["celt_fb", "silk_wb"])
    b = cn.Batch(blobs, mode=2, n_slots=16, profile=True)
    b.run()
    b.check()
    dms, rms = b.kernel_ms()
    assert dms > 0 and rms > 0
    assert b.chains > 0 and b.packets > 0 and b.out16 == sum(b.n16)
    b.close()


def test_batch_errors(cn):
    with pytest.raises(cn.CuopusError) as e:
        cn.Batch([b"not an ogg stream"])
    assert e.value.rc == 1
    blobs = _blobs(["silk_wb"])
    d, _ = ol.demux_chains(blobs[0])
    with pytest.raises(cn.CuopusError) as e:
        cn.Batch(blobs, lost_masks=[np.zeros(len(d.packets) + 1, np.uint8)])
    assert e.value.rc == 3
