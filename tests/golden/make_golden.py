"""Writes the golden fixtures of tests/golden/ from the oracle.

For a few short seeded streams (SILK WB, CELT FB with all frame sizes, hybrid
FB with in-band FEC and a seeded loss mask, mixed-mode with transitions) it
stores the Ogg/Opus bytes (*.opus) and, in <name>.npz (no pickles), the chain
table, the per-packet sample counts and final ranges and the compacted int16
48 kHz PCM of fixed-point libopus 1.5.2 (built
from the reference's vendored sources under oracle/) with the chain protocol
(1 s chains, 250 ms preroll, fresh decoder per chain).

    python -m tests.golden.make_golden      # needs oracle/_ref (build())
"""
import os
import shlex
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tests import oracle_lib as ol  # noqa: E402

GOLDEN = {
    "g_silk_wb": ("--seed 101 --dur 3 --signal speech --app voip --bw wb --force silk --vbr 0 --bitrate 16000",
                  None),
    "g_celt_fb": ("--seed 102 --dur 3 --signal music --app lowdelay --bw fb --bitrate 64000 "
                  "--frameset 120,240,480,960", None),
    "g_hybrid_fec": ("--seed 103 --dur 3 --signal speech --app voip --bw fb --force hybrid --bitrate 32000 "
                     "--fec 1 --loss 5", 0.08),
    "g_mixed": ("--seed 104 --dur 4 --signal mixed --app audio --bw fb --brange 12000 64000", None),
}


def make(name, args, loss_p):
    path = os.path.join(HERE, name + ".opus")
    subprocess.run([ol.MKCORPUS, "--out", path] + shlex.split(args), check=True)
    blob = open(path, "rb").read()
    d, chains = ol.demux_chains(blob)
    data, off, ln = d.blob()
    lost = ol.loss_mask(len(off), loss_p, seed=sum(name.encode())) if loss_p else None
    ref = ol.decode_chains(data, off, ln, chains, lost=lost)
    assert (ref["ret"] >= 0).all()
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        chains=np.array(chains, dtype=np.uint32).reshape(-1, 4),
                        nsamp=ref["nsamp"], rng=ref["rng"], pcm=ref["pcm"],
                        lost=lost if lost is not None else np.zeros(0, np.uint8))
    print(name, len(off), "packets", len(chains), "chains", ref["pcm"].size, "samples")


def main():
    for name, (args, loss_p) in GOLDEN.items():
        make(name, args, loss_p)


if __name__ == "__main__":
    main()
