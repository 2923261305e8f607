"""Pins of oracle functions that the first golden set left open (CPU only).

  orc_rhs_mix        CpuSat<float>::GetDerivatives (AnalogSAT/cpu_sat/CpuSat.cpp:197-241,
                     float instantiation at :391): bit-exact against the
                     reference build's float outputs (rhs_cases_f32.npz) on the six
                     RHS cases, the bias case included (serial float mean)
  orc_randomsat      RandomSatMaker (AnalogSAT/problem/RandomSatMaker.cpp:59-125):
                     literal streams bit-exact against the reference build's
                     (randomsat_cases.npz), including instances where most
                     draws repeat an earlier clause; the generated clauses are
                     distinct, and without repeats the stream is FastSatMaker's
  orc_philox4x32_10  Philox4x32-10 (Salmon et al., SC'11): the three known-answer
                     vectors published with the Random123 library
  orc_philox_uniform the library's uniform stream built on it: top-bit
                     conversions, counter/key packing and offsets checked
                     against an independent numpy Philox
"""
import os

import numpy as np
import pytest

import oracle as O


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name), allow_pickle=False)


@pytest.mark.parametrize("case", ["k3", "k4", "mixed", "k12", "bias", "n2000"])
def test_mix_rhs_equals_reference_float_instantiation(golden_dir, case):
    d = _load(golden_dir, "rhs_cases.npz")
    f = _load(golden_dir, "rhs_cases_f32.npz")
    p = O.Problem(d[case + "_lits"], d[case + "_lens"])
    x = d[case + "_x"].astype(np.float32)
    got = O.rhs(p, x, mode="mix", b=float(d[case + "_b"]))
    want = f[case + "_rhs32"]
    assert got.dtype == np.float32 and want.dtype == np.float32
    assert np.array_equal(got, want), np.abs(got - want).max()
    # the float results are not the double ones rounded: the fixture tests
    # the float operation order, not just its precision
    assert not np.array_equal(want, d[case + "_rhs"].astype(np.float32))


def test_randomsat_matches_reference_streams(golden_dir):
    d = _load(golden_dir, "randomsat_cases.npz")
    assert len(d.files) >= 6
    for key in d.files:
        n, m, k, sd = map(int, key.split("_")[1:])
        p = O.randomsat(n, m, k, sd)
        assert np.array_equal(p.C, d[key]), key
        rows = [tuple(r) for r in d[key].reshape(m, k)]
        assert len(set(rows)) == m, key                  # no clause twice
        for r in rows:                                   # sorted by variable, distinct variables
            v = [abs(x) for x in r]
            assert v == sorted(v) and len(set(v)) == k, key


def test_randomsat_without_repeats_is_fastsat_stream():
    # C(2000,3) 2^3 clauses: 8500 draws almost surely never repeat, and then
    # both makers consume the same stream
    a = O.randomsat(2000, 8500, 3, 4)
    b = O.fastsat(2000, 8500, 3, 4)
    assert np.array_equal(a.C, b.C)


def test_randomsat_rejects_impossible_counts():
    with pytest.raises(ValueError):
        O.randomsat(4, 9, 1, 0)      # only 8 distinct 1-clauses over 4 variables


# Random123 kat_vectors, philox4x32_10 (counter words, key words, output words)
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert [int(v) for v in O.philox4x32_10(ctr, key)] == list(want)


def _np_philox(ctr, key):
    """Independent numpy Philox4x32-10 over arrays of counters (uint64 math)."""
    c = [np.asarray(w, np.uint64) for w in ctr]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    M0, M1, mask = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57), np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & mask, p1 & mask,
             ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & mask, p0 & mask]
        k0, k1 = (k0 + np.uint64(0x9E3779B9)) & mask, (k1 + np.uint64(0xBB67AE85)) & mask
    return c


@pytest.mark.parametrize("seed,offset", [(0, 0), (12345, 7), ((1 << 40) + 3, (1 << 33) - 5)])
def test_philox_uniform_stream_packing(seed, offset):
    n = 5000
    idx = np.arange(n, dtype=np.uint64) + np.uint64(offset)
    w = _np_philox([idx & np.uint64(0xFFFFFFFF), idx >> np.uint64(32), np.full(n, 0x5EED, np.uint64),
                    np.zeros(n, np.uint64)], (seed & 0xFFFFFFFF, seed >> 32))
    want64 = ((w[0] << np.uint64(32) | w[1]) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    want32 = ((w[0] >> np.uint64(8)).astype(np.float64) * 2.0 ** -24).astype(np.float32)
    got64 = O.philox_uniform(seed, offset, n, np.float64)
    got32 = O.philox_uniform(seed, offset, n, np.float32)
    assert np.array_equal(got64, want64)
    assert np.array_equal(got32, want32)
    assert got64.min() >= 0.0 and got64.max() < 1.0 and got32.max() < 1.0
    # a continuing stream: two draws equal one draw of the joint length
    assert np.array_equal(np.concatenate([O.philox_uniform(seed, offset, 100),
                                          O.philox_uniform(seed, offset + 100, 50)]),
                          O.philox_uniform(seed, offset, 150))
