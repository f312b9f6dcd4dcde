"""GPU: the continuous-batching server returns, for every request, exactly the
output the single-file path gives for that stream alone, whether it runs
alone, mixed with other requests arriving while it is on the GPU, or under a
memory budget that forces requests to wait for admission.  Lossy requests
match the lossy single-file path.  Bar: identical float32 tensors.
"""
import time

import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

NAMES = ["silk_wb", "celt_fb", "hybrid_fb", "mixed", "silk_nb", "hybrid_fec", "hybrid_swb", "silk_long"]


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


def _blobs():
    return {n: open(ol.small_corpus(n, ol.RECIPES[n]), "rb").read() for n in NAMES}


def _drain(srv, want, timeout_s=60.0):
    got = {}
    t0 = time.time()
    while len(got) < want:
        for rid, out, rc in srv.poll():
            got[rid] = (out, rc)
        assert time.time() - t0 < timeout_s, ("server did not finish", srv.stats())
        time.sleep(0.001)
    return got


def test_alone_equals_single_file(cn):
    import torch
    blobs = _blobs()
    srv = cn.Server(n_streams=2)
    for name, blob in blobs.items():
        print("case serve alone", name, flush=True)
        rid = srv.submit(blob)
        got = _drain(srv, 1)
        out, rc = got[rid]
        assert rc == 0
        assert torch.equal(out, cn.decode_ogg_to_cuda_tensor(blob)), name
    srv.close()


def test_mix_equals_alone(cn):
    """Requests submitted while others are in flight (several streams, small
    launches) give the same tensors as alone."""
    import torch
    blobs = _blobs()
    ref = {n: cn.decode_ogg_to_cuda_tensor(b) for n, b in blobs.items()}
    srv = cn.Server(n_streams=3, n_waves=16)
    rs = np.random.default_rng(5)
    order = [NAMES[i] for i in rs.integers(0, len(NAMES), size=40)]
    ids = {}
    got = {}
    for k, name in enumerate(order):
        ids[srv.submit(blobs[name])] = name
        if k % 3 == 2:
            for rid, out, rc in srv.poll():
                got[rid] = (out, rc)
    got.update(_drain(srv, len(order) - len(got)))
    assert len(got) == len(order)
    for rid, (out, rc) in got.items():
        assert rc == 0
        assert torch.equal(out, ref[ids[rid]]), ids[rid]
    srv.close()


def test_memory_budget_admission(cn):
    """A budget that fits about one request at a time: requests queue, are
    admitted as memory frees, and still match."""
    import torch
    blobs = _blobs()
    ref = {n: cn.decode_ogg_to_cuda_tensor(b) for n, b in blobs.items()}
    # device bytes a request needs: packets + int16 48 kHz PCM + f32 output (+ tables)
    need = [len(b) + 2 * (3 * cn.probe(b) + 48) + 4 * cn.probe(b) + (64 << 10) for b in blobs.values()]
    srv = cn.Server(n_streams=4, mem_budget=int(1.5 * max(need)))
    ids = {srv.submit(b): n for n, b in blobs.items()}
    got = {}
    for rid, out, rc in srv.poll():  # first admission: only what the budget holds
        got[rid] = (out, rc)
    saw_queue = srv.stats()["queued"] > 0
    got.update(_drain(srv, len(ids) - len(got)))
    for rid, (out, rc) in got.items():
        assert rc == 0 and torch.equal(out, ref[ids[rid]])
    assert saw_queue
    srv.close()


def test_lossy_request(cn):
    import torch
    blob = _blobs()["hybrid_fec"]
    d, _ = ol.demux_chains(blob)
    lost = ol.loss_mask(len(d.packets), 0.05, seed=3)
    srv = cn.Server()
    rid = srv.submit(blob, lost_mask=lost)
    got = _drain(srv, 1)
    assert got[rid][1] == 0
    assert torch.equal(got[rid][0], cn.decode_ogg_lossy(blob, lost))
    srv.close()


def test_submit_errors(cn):
    srv = cn.Server(mem_budget=1 << 16)
    with pytest.raises(cn.CuopusError) as e:
        srv.submit(_blobs()["celt_fb"])  # larger than the whole budget
    assert e.value.rc == 3
    with pytest.raises(cn.CuopusError) as e:
        srv.submit(b"not an ogg stream")
    assert e.value.rc == 1
    srv.close()


def test_submit_orders_after_callers_stream(cn):
    """Requests submitted while the caller's stream still has queued work on
    recycled memory: the caching allocator hands the freed block of a
    tensor that a queued kernel still writes to the next request's output;
    the server's launch waits for the caller's stream, so the request's
    output is what it is alone."""
    import torch
    blobs = _blobs()
    ref = {n: cn.decode_ogg_to_cuda_tensor(b) for n, b in blobs.items()}
    srv = cn.Server(n_streams=4)
    ids = {}
    for rep in range(3):
        for name, blob in blobs.items():
            n16 = cn.probe(blob)
            junk = torch.empty(n16, dtype=torch.float32, device="cuda")
            for _ in range(4):
                junk.fill_(float("nan"))  # queued work on this block, on the caller's stream
            del junk                       # the next allocation of this size reuses the block
            ids[srv.submit(blob)] = name
    got = _drain(srv, len(ids))
    for rid, (out, rc) in got.items():
        assert rc == 0
        assert torch.equal(out, ref[ids[rid]]), ids[rid]
    srv.close()


def test_held_memory_within_budget(cn):
    """The server never holds more device memory than its budget (all launch
    slots' buffers, decoder heaps included)."""
    import torch
    blobs = _blobs()
    budget = 48 << 20
    srv = cn.Server(n_streams=3, mem_budget=budget)
    ids = {}
    for name, blob in list(blobs.items()) * 3:
        ids[srv.submit(blob)] = name
    got = {}
    t0 = time.time()
    while len(got) < len(ids):
        for rid, out, rc in srv.poll():
            got[rid] = (out, rc)
        assert srv.stats()["held_bytes"] <= budget
        assert time.time() - t0 < 60.0, ("server did not finish", srv.stats())
        time.sleep(0.0005)
    for rid, (out, rc) in got.items():
        assert rc == 0
    srv.close()
    torch.cuda.synchronize()
