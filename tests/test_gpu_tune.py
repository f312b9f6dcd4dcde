"""GPU: the autotuner's shapes give bit-identical output, the library reads
its per-device defaults file, and the work-order cost model's calibration
runs (per-chain device times)."""
import os
import subprocess
import sys

import pytest

from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu
NAMES = ["silk_wb", "celt_fb", "hybrid_fb", "mixed"]


@pytest.fixture(scope="module")
def cn():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native
    return cuopus_native


def test_tuner_shapes_bit_identical(cn, monkeypatch):
    from cuopus_amd import autotune
    monkeypatch.setattr(autotune, "SLOTS", (8, 64))
    blobs = [open(ol.small_corpus(n, ol.RECIPES[n]), "rb").read() for n in NAMES]
    best = autotune.tune(blobs, reps=1, write=False)  # raises if any shape's output differs
    assert best["n_slots"] in (8, 64) and best["order"] in (0, 1, 2) and 0 <= best["rs_variant"] < 5
    assert len(best["grid"]) == 2 * 3 + 1 and len(best["rs_grid"]) == 5


def test_tuned_file_is_read(cn, tmp_path):
    from cuopus_amd import autotune
    path = tmp_path / "tuned.txt"
    path.write_text("%s slots=96 rs=2 order=0\n" % autotune.device_key())
    code = ("from cuopus_amd import cuopus_native as cn; import torch; torch.cuda.init(); "
            "print(cn.tuned_defaults())")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, cwd=ol.ROOT,
                       env=dict(os.environ, CUOPUS_TUNED=str(path)))
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1] == "(96, 2, 0)"


def test_calibration_fit(cn):
    from cuopus_amd import autotune
    blobs = [open(ol.small_corpus(n, ol.RECIPES[n]), "rb").read() for n in NAMES]
    cal = autotune.calibrate(blobs)
    assert cal["chains"] > 0 and all(v >= 0 for v in cal["ticks_per_frame"].values())
