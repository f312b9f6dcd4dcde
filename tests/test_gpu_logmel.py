"""GPU: the log-mel front-end (cuopus_log_mel) on the resident f32 16 kHz
output against the float64 oracle (oracle/logmel.py) on the same input.

Tolerance: 2e-3 absolute in natural-log units (0.009 dB) on every entry.
The kernel computes in float32 (window, 256-point complex FFT, power,
filter sums); a float32 FFT of the same frames differs from the float64
evaluation by at most ~1.5e-4 on these corpora, the rest is headroom.
"""
import numpy as np
import pytest

from oracle import logmel as lm
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL = 2e-3


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


@pytest.mark.parametrize("name,n_mels", [("silk_wb", 80), ("celt_fb", 80), ("mixed", 128), ("silk_dtx", 40),
                                         ("len_1s", 80)])
def test_log_mel_matches_float64(cn, name, n_mels):
    import torch
    blob = open(ol.small_corpus(name, ol.recipe(name)), "rb").read()
    x = cn.decode_ogg_to_cuda_tensor(blob)
    print("case logmel", name, n_mels, x.numel(), flush=True)
    got = cn.log_mel(x, n_mels=n_mels)
    torch.cuda.synchronize()
    want = lm.log_mel(x.cpu().numpy(), n_mels=n_mels)
    assert got.shape == want.shape == (lm.n_frames(x.numel()), n_mels)
    diff = np.abs(got.cpu().numpy().astype(np.float64) - want)
    print("  max abs diff %.3g over %d values" % (diff.max(), diff.size), flush=True)
    assert diff.max() <= TOL


def test_log_mel_edges(cn):
    import torch
    for n in (0, 399, 400, 559, 560, 1601):
        x = torch.randn(n, dtype=torch.float32, device="cuda") * 0.1
        got = cn.log_mel(x)
        want = lm.log_mel(x.cpu().numpy())
        assert got.shape == want.shape
        if want.size:
            assert np.abs(got.cpu().numpy() - want).max() <= TOL
    z = cn.log_mel(torch.zeros(2000, device="cuda"))
    assert float((z - float(np.log(1e-10))).abs().max()) <= 1e-5
    with pytest.raises(cn.CuopusError) as e:
        cn.log_mel(torch.zeros(1000, device="cuda"), n_mels=129)
    assert e.value.rc == 3
