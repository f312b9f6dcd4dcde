"""bench.py pieces that run without a GPU: the one-GPU refusal, the readers of
the committed rocprofv3 summary, and the single-stream CPU rate (fixed-point
libopus over the first packets of a file as one stream)."""
import os
import subprocess
import sys

import numpy as np

from tests import oracle_lib as ol

sys.path.insert(0, ol.ROOT)
import bench  # noqa: E402

GOLD = os.path.join(ol.ROOT, "tests", "golden")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ol.ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_profile_readers():
    traffic = bench._read_traffic("celt_fb_64k_2h")
    assert isinstance(traffic, int) and traffic > 0
    issue = bench._read_issue("celt_fb_64k_2h")
    for k in ("valu", "salu", "any", "wait"):
        assert 0.0 < issue[k] < 1.0
    assert issue["any"] >= max(issue["valu"], issue["salu"])
    assert bench._read_issue("no_such_workload") is None


def test_cpu_stream_rate_and_first_chain():
    blob = open(os.path.join(GOLD, "g_celt_fb.opus"), "rb").read()
    r = bench._cpu_stream(blob, 1e9)
    assert r["value"] > 0 and r["cores"] == 1 and r["kind"] == "reference"
    # the stream decodes every packet from a fresh decoder, so its first
    # samples are chain 0's (chain 0 starts at packet 0 without preroll)
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    whole = ol.decode_chains(data, off, ln, [(0, off.size, d.preskip, 0)], caps=[off.size * 5760 + 5760])
    first = ol.decode_chains(data, off, ln, chains[:1])
    keep0 = chains[0][3]
    assert chains[0][0] == 0 and chains[0][2] == d.preskip
    assert np.array_equal(whole["full"][0][d.preskip:d.preskip + keep0], first["pcm"])
