"""CPU: the one-step searches the device decoder uses in place of serial ones
give the serial results for every input their tables allow.

* CELT leaves (celt.h qp_leaf, device path): bits2pulses' six-step bisection
  over a pulse-cache row (rate.h:48-76 of libopus) ends at the lowest entry
  index >= bits -- or the row's count when there is none -- and the one below
  it.  Checked here for every row of the cache (tables_gen.h kCacheIndex /
  kCacheBits, generated from libopus' static mode) and every bit count that
  can differ (the entries are bytes, so all counts above 256 behave alike).
* icdf tables searched by a ballot (silk.h silk_icdf, ec_icdf_ballot): a
  symbol is the lowest index j with val >= r * icdf[j]; the serial loop
  (entdec.c ec_dec_icdf) stops at the same j because each table is
  non-increasing and ends in 0.  Checked for every table the ballot path
  reads, over the length the decoder passes.
"""
import os
import re

import pytest

TABLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuopus_amd", "csrc", "dec",
                      "tables_gen.h")


def _arr(name):
    s = open(TABLES).read()
    m = re.search(name + r"\[(\d+)\]\s*=\s*\{([^}]*)\}", s)
    assert m, name
    return [int(x) for x in m.group(2).replace("\n", " ").split(",") if x.strip()]


def _bisect(c, bits):
    lo, hi = 0, c[0]
    for _ in range(6):
        mid = (lo + hi + 1) >> 1
        if c[mid] >= bits:
            hi = mid
        else:
            lo = mid
    return lo if bits - (-1 if lo == 0 else c[lo]) <= c[hi] - bits else hi


def _ballot(c, bits):
    mx = c[0]
    cand = [j for j in range(1, mx + 1) if c[j] >= bits]
    hi = cand[0] if cand else mx
    lo = hi - 1
    return lo if bits - (-1 if lo == 0 else c[lo]) <= c[hi] - bits else hi


def test_pulse_cache_ballot_equals_bisection():
    cache_bits, cache_index = _arr("kCacheBits"), _arr("kCacheIndex")
    rows = 0
    for ci in cache_index:
        if ci < 0:
            continue
        c = cache_bits[ci:]
        assert 4 <= c[0] <= 40, "row count outside the 64 lanes / six bisection steps"
        rows += 1
        for bits in range(-8, 300):
            assert _ballot(c, bits) == _bisect(c, bits), (ci, bits)
    assert rows == sum(1 for ci in cache_index if ci >= 0) and rows > 80


@pytest.mark.parametrize("name,n", [("kDeltaGainIcdf", 41), ("kUniform8Icdf", 8), ("kNlsfExtIcdf", 7),
                                    ("kNlsfInterpFactorIcdf", 5), ("kPitchDeltaIcdf", 21), ("kPitchLagIcdf", 32),
                                    ("kUniform6Icdf", 6), ("kLtpGainIcdf0", 8), ("kLtpGainIcdf1", 16),
                                    ("kLtpGainIcdf2", 32), ("kPitchContourIcdf", 34), ("kPitchContourNBIcdf", 11),
                                    ("kPitchContour10msIcdf", 12)])
def test_icdf_tables_searchable_in_one_step(name, n):
    t = _arr(name)
    assert len(t) == n and t[-1] == 0
    assert all(t[i] >= t[i + 1] for i in range(n - 1))


@pytest.mark.parametrize("name,rows,n", [("kGainIcdf", 3, 8), ("kNlsfWB_CB1_iCDF", 2, 32),
                                         ("kNlsfNBMB_CB1_iCDF", 2, 32), ("kNlsfWB_ec_iCDF", 8, 9),
                                         ("kNlsfNBMB_ec_iCDF", 8, 9), ("kPulsesPerBlockIcdf", 10, 18)])
def test_icdf_table_rows_searchable_in_one_step(name, rows, n):
    t = _arr(name)
    assert len(t) == rows * n
    for r in range(rows):
        row = t[r * n:(r + 1) * n]
        assert row[-1] == 0 and all(row[i] >= row[i + 1] for i in range(n - 1)), (name, r)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_shell_tables_searchable_in_one_step(k):
    """Split of p pulses: the p + 1 entries from byte p(p+1)/2 - 1."""
    t = _arr("kShellTable%d" % k)
    for p in range(1, 17):
        off = (p * (p + 1) >> 1) - 1
        row = t[off:off + p + 1]
        assert len(row) == p + 1 and row[-1] == 0 and all(row[i] >= row[i + 1] for i in range(p)), p


def _dec_icdf_serial(rng, val, icdf, ftb=8):
    s, r, ret = rng, rng >> ftb, -1
    while True:
        t = s
        ret += 1
        s = r * icdf[ret]
        if not val < s:
            return ret, t - s, val - s


def _dec_icdf_ballot(rng, val, icdf, n, ftb=8):
    r = rng >> ftb
    tj = [r * icdf[j] for j in range(n)]
    ret = min(j for j in range(n) if val >= tj[j])
    tv = tj[ret - 1] if ret else rng
    return ret, tv - tj[ret], val - tj[ret]


def test_icdf_ballot_equals_serial_search():
    import random
    rnd = random.Random(7)
    t = _arr("kDeltaGainIcdf")
    for _ in range(20000):
        rng = rnd.randrange(1 << 23, 1 << 32)
        val = rnd.randrange(0, rng)
        assert _dec_icdf_ballot(rng, val, t, len(t)) == _dec_icdf_serial(rng, val, t)
