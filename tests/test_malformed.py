"""Malformed input: random, mis-framed, truncated and bit-flipped packets
(tests/malformed.py, after libopus's test_opus_decode.c:295-367) decoded
chain by chain, each chain from a fresh decoder, against the oracle.

CPU: the decoder source compiled for the host under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/hostgate/fuzz_gate) -- no memory error,
no undefined behaviour, and every return code, final range and sample up to
a chain's first error equal to libopus.  The lane-parallel device versions
of the SILK pulse/parameter/synthesis stages and the CELT codeword decode
are not part of this build (they compile for the GPU only): the GPU tests
here cover them, and tests/test_gpu_twins.py compares them with the serial
twins this build checks, on these sets and every corpus recipe.

GPU: the same sets through cuopus_gate_chains: per-packet sample counts and
final ranges, per-chain stop codes and the PCM each chain produced before
its first error, all equal to the oracle.

Stereo sets (tests/malformed.py STEREO_SETS: every stereo TOC
configuration, configuration x channel-count pairs, framing with random
channel flags, damaged packets of stereo streams with the stereo flag open to
flips) go through the stereo build (fuzz_gate_st on the CPU, the stereo
kernel on the GPU) at one and two output channels, against libopus opened
with as many channels.
"""
import os
import subprocess

import numpy as np
import pytest

from tests import malformed as mf
from tests import oracle_lib as ol

FUZZ = os.path.join(ol.ROOT, "tests", "hostgate", "fuzz_gate")
FUZZ_ST = os.path.join(ol.ROOT, "tests", "hostgate", "fuzz_gate_st")
CASES = [(name, lossy) for name in sorted(mf.SETS) for lossy in (False, True)]
ST_CASES = [(name, lossy, ch) for name in sorted(mf.STEREO_SETS) for lossy in (False, True) for ch in (1, 2)]


def _set(name, lossy):
    data, off, ln, chains = (mf.SETS.get(name) or mf.STEREO_SETS[name])()
    lost = mf.loss_for(off) if lossy else None
    return data, off, ln, lost, chains


def _run_fuzz(exe, name, lossy, channels, tmp_path):
    if not os.path.exists(exe):
        pytest.skip("sanitizer build of the host gate not present")
    data, off, ln, lost, chains = _set(name, lossy)
    path = tmp_path / "set.bin"
    mf.write_set(str(path), data, off, ln, lost, chains)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(path), str(channels)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.parametrize("name,lossy", CASES)
def test_sanitized_host_gate_matches_libopus(name, lossy, tmp_path):
    _run_fuzz(FUZZ, name, lossy, 1, tmp_path)


@pytest.mark.parametrize("name,lossy,channels", ST_CASES)
def test_sanitized_stereo_host_gate_matches_libopus(name, lossy, channels, tmp_path):
    _run_fuzz(FUZZ_ST, name, lossy, channels, tmp_path)


def _gpu_case(name, lossy, channels=1, stereo_build=False):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cuopus_amd import cuopus_native as cn
    data, off, ln, lost, chains = _set(name, lossy)
    chains, ref = mf.oracle_keep(data, off, ln, chains, lost, channels=channels)
    print("case malformed", name, "lossy" if lossy else "", "channels", channels, "chains", len(chains), flush=True)
    pcm, nsamp, rng, stop, rc = cn.gate_chains(data, off, ln, chains, lost=lost, channels=channels,
                                               stereo_build=stereo_build)
    assert rc in (0, 1)
    want_stop = np.minimum(ref["ret"], 0)
    assert np.array_equal(stop, want_stop), np.nonzero(stop != want_stop)[0][:10]
    assert np.array_equal(nsamp, ref["nsamp"]), np.nonzero(nsamp != ref["nsamp"])[0][:10]
    assert np.array_equal(rng, ref["rng"]), np.nonzero(rng != ref["rng"])[0][:10]
    assert np.array_equal(pcm, ref["pcm"]), np.nonzero(pcm != ref["pcm"])[0][:10]


@pytest.mark.gpu
@pytest.mark.parametrize("name,lossy", CASES)
def test_gpu_gate_matches_libopus(name, lossy):
    _gpu_case(name, lossy)


@pytest.mark.gpu
@pytest.mark.parametrize("name,lossy,channels", ST_CASES)
def test_gpu_stereo_gate_matches_libopus(name, lossy, channels):
    _gpu_case(name, lossy, channels, stereo_build=True)
