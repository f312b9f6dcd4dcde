"""GPU: the allocation-free entry point (cuopus_decode_f16k_noalloc_on),
called through the C ABI, gives exactly what cuopus_decode_f16k_on gives on
the same bytes, on two calls in a row (the second with a smaller input, so
the static scratch holds the first call's contents), without the library
allocating device memory; streams larger than its scratch are refused.
"""
import ctypes
import os

import numpy as np
import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native as cn
    lib = ctypes.CDLL(cn.library_path())
    sig = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64),
           ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    for name in ("cuopus_decode_f16k_on", "cuopus_decode_f16k_noalloc_on"):
        getattr(lib, name).argtypes = sig
        getattr(lib, name).restype = ctypes.c_int
    lib.cuopus_workspace_bytes.restype = ctypes.c_longlong
    return torch, cn, lib


def _call(torch, lib, name, blob, cap):
    out = torch.full((max(cap, 1),), float("nan"), dtype=torch.float32, device="cuda")
    n = ctypes.c_int64(-1)
    err = ctypes.create_string_buffer(512)
    rc = getattr(lib, name)(blob, len(blob), out.data_ptr(), cap, ctypes.byref(n),
                            torch.cuda.current_stream().cuda_stream, err, 512)
    torch.cuda.synchronize()
    return rc, out[:max(n.value, 0)].clone(), err.value.decode()


def test_noalloc_matches_allocating_entry_point(env):
    torch, cn, lib = env
    first = open(ol.small_corpus("mixed", ol.recipe("mixed")), "rb").read()        # 20 s, mode switches
    second = open(ol.small_corpus("silk_wb", ol.recipe("silk_wb")), "rb").read()   # 12 s: smaller
    ref = {}
    for tag, blob in (("first", first), ("second", second)):
        cap = cn.probe(blob)
        rc, out, msg = _call(torch, lib, "cuopus_decode_f16k_on", blob, cap)
        assert rc == 0, msg
        ref[tag] = out
    # the allocating path's workspaces are released; the allocation-free path must not bring any back
    lib.cuopus_release_workspaces()
    assert lib.cuopus_workspace_bytes() == 0
    for tag, blob in (("first", first), ("second", second)):
        print("case noalloc", tag, len(blob), "bytes", flush=True)
        rc, out, msg = _call(torch, lib, "cuopus_decode_f16k_noalloc_on", blob, cn.probe(blob))
        assert rc == 0, msg
        assert out.numel() == ref[tag].numel()
        assert torch.equal(out, ref[tag]), tag
        assert lib.cuopus_workspace_bytes() == 0
    # and against the oracle for the smaller one (resampler tolerance 1e-6)
    d, chains = ol.demux_chains(second)
    data, off, ln = d.blob()
    want = ol.resample(ol.decode_chains(data, off, ln, chains)["pcm"])
    assert float(np.abs(out.cpu().numpy() - want).max()) <= 1e-6


def test_noalloc_refuses_oversize_stream(env):
    torch, cn, lib = env
    path = os.path.join(ol.BENCH, "celt_fb_64k_2h.opus")
    if not os.path.exists(path):
        pytest.skip("2 h benchmark stream not generated")
    blob = open(path, "rb").read()
    # full-size output: even a call that did not refuse would stay inside it
    out = torch.empty(cn.probe(blob), dtype=torch.float32, device="cuda")
    n = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(512)
    rc = lib.cuopus_decode_f16k_noalloc_on(blob, len(blob), out.data_ptr(), cn.probe(blob), ctypes.byref(n),
                                           torch.cuda.current_stream().cuda_stream, err, 512)
    assert rc == 3, err.value
    assert b"scratch" in err.value


def test_noalloc_stereo_stream(env):
    """A stereo stream (the stereo kernel build, its larger heaps from the
    same static scratch) gives the allocating entry point's output."""
    torch, cn, lib = env
    blob = open(ol.small_corpus("st_mixed", ol.recipe("st_mixed")), "rb").read()
    cap = cn.probe(blob)
    rc, ref, msg = _call(torch, lib, "cuopus_decode_f16k_on", blob, cap)
    assert rc == 0, msg
    lib.cuopus_release_workspaces()
    print("case noalloc stereo", len(blob), "bytes", flush=True)
    rc, out, msg = _call(torch, lib, "cuopus_decode_f16k_noalloc_on", blob, cap)
    assert rc == 0, msg
    assert torch.equal(out, ref)
    assert lib.cuopus_workspace_bytes() == 0
