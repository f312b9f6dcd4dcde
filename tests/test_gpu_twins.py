"""Lane-parallel decoder stages against their serial twins, on the GPU.

The SILK pulse, parameter and synthesis stages and the CELT codeword decode
have wavefront-parallel versions that only compile for the device; the host
build (the CPU gate, and its AddressSanitizer + UndefinedBehaviorSanitizer
build run over the malformed sets) runs their serial twins instead.  The
library built with CUO_SERIAL_TWINS (tests/hostgate/libcuopus_twins.so) runs
those serial twins on the GPU.  Every case here decodes the same packets
through both builds' cuopus_gate_chains and requires identical per-packet
sample counts and final ranges, per-chain stop codes and PCM: the lane
versions compute what the sanitizer-checked serial code computes, on every
corpus recipe (mono and stereo, with and without loss) and on the malformed
sets.
"""
import os

import numpy as np
import pytest

from tests import malformed as mf
from tests import oracle_lib as ol

pytestmark = pytest.mark.gpu

TWINS = os.path.join(ol.ROOT, "tests", "hostgate", "libcuopus_twins.so")


@pytest.fixture(scope="module")
def libs():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native as cn
    assert os.path.exists(TWINS), "serial-twin build missing: run __graft_entry__.build()"
    return cn, cn._Native(TWINS)


def _same(cn, twin, tag, data, off, ln, chains, lost=None, channels=1, stereo_build=False):
    print("case twins", tag, "packets", len(off), "chains", len(chains), "channels", channels, flush=True)
    a = cn.gate_chains(data, off, ln, chains, lost=lost, channels=channels, stereo_build=stereo_build)
    b = cn.gate_chains(data, off, ln, chains, lost=lost, channels=channels, stereo_build=stereo_build, lib=twin)
    for name, x, y in zip(("pcm", "nsamp", "rng", "stop", "rc"), a, b):
        if name == "rc":
            assert x == y, (tag, x, y)
        else:
            bad = np.nonzero(x != y)[0]
            assert bad.size == 0, (tag, name, bad[:10])
    return a


def _corpus(name):
    path = ol.small_corpus(name, ol.recipe(name))
    d, chains = ol.demux_chains(open(path, "rb").read())
    data, off, ln = d.blob()
    return data, off, ln, chains


MONO = sorted(ol.RECIPES) + sorted(ol.SEED_SWEEP) + sorted(ol.GAIN_RECIPES)


@pytest.mark.parametrize("name", MONO)
@pytest.mark.parametrize("lossy", [False, True])
def test_twins_mono_corpora(libs, name, lossy):
    cn, twin = libs
    data, off, ln, chains = _corpus(name)
    lost = ol.loss_mask(len(off), 0.05, seed=sum(name.encode())) if lossy else None
    _same(cn, twin, name + (" lossy" if lossy else ""), data, off, ln, chains, lost=lost)


@pytest.mark.parametrize("name", sorted(ol.ST_RECIPES))
@pytest.mark.parametrize("channels", [1, 2])
def test_twins_stereo_corpora(libs, name, channels):
    cn, twin = libs
    data, off, ln, chains = _corpus(name)
    lost = ol.loss_mask(len(off), 0.05, seed=sum(name.encode()))
    _same(cn, twin, name, data, off, ln, chains, channels=channels, stereo_build=True)
    _same(cn, twin, name + " lossy", data, off, ln, chains, lost=lost, channels=channels, stereo_build=True)


@pytest.mark.parametrize("name", sorted(mf.SETS))
@pytest.mark.parametrize("lossy", [False, True])
def test_twins_malformed(libs, name, lossy):
    cn, twin = libs
    data, off, ln, chains = mf.SETS[name]()
    _same(cn, twin, "malformed " + name, data, off, ln, chains, lost=mf.loss_for(off) if lossy else None)


@pytest.mark.parametrize("name", sorted(mf.STEREO_SETS))
@pytest.mark.parametrize("channels", [1, 2])
def test_twins_malformed_stereo(libs, name, channels):
    cn, twin = libs
    data, off, ln, chains = mf.STEREO_SETS[name]()
    _same(cn, twin, "malformed " + name, data, off, ln, chains, channels=channels, stereo_build=True)
    _same(cn, twin, "malformed %s lossy" % name, data, off, ln, chains, lost=mf.loss_for(off), channels=channels,
          stereo_build=True)
