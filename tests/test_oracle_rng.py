"""Pins the oracle's random-number layer.

- mt19937_64: the C++ standard's known answer ([rand.predef]: the 10000th output of a
  default-constructed std::mt19937_64 is 9981545732273789042) and libstdc++ itself
  (oracle/stdrng_probe, compiled with g++ from the system headers).
- generate_canonical / uniform_real_distribution(a, b) / uniform_int_distribution:
  libstdc++ draws on the same engine (MCIntegrator.hpp:32-34, TrialMoveInterface.hpp:120-124,
  SRRDVecMove.hpp:27).
- Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mcitest import ROOT, load

PROBE = os.path.join(ROOT, "oracle", "build", "stdrng_probe")


def probe(cmd):
    out = subprocess.run([PROBE], input=cmd + "\n", capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()
    assert lines[-1] == "end"
    return lines[:-1]


def test_mt19937_64_known_answer():
    lib = load("orc")
    assert lib.orc_mt64_nth(5489, 10000) == 9981545732273789042


@pytest.mark.parametrize("seed", [0, 1, 1337, 5649871, 2 ** 64 - 1])
def test_mt19937_64_words_match_libstdcxx(seed):
    lib = load("orc")
    n = 1000
    out = np.zeros(n, dtype=np.uint64)
    lib.orc_mt64_words(seed, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    ref = np.array([int(v) for v in probe(f"words {seed} {n}")], dtype=np.uint64)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("a,b", [(0.0, 1.0), (-1.0, 1.0), (0.5, 2.0)])
def test_uniform_real_matches_libstdcxx(a, b):
    lib = load("orc")
    n, seed = 2000, 4242
    words = np.zeros(n, dtype=np.uint64)
    lib.orc_mt64_words(seed, n, words.ctypes.data_as(C.POINTER(C.c_uint64)))
    ours = np.array([lib.orc_canonical53(int(w)) * (b - a) + a for w in words])
    ref = np.array([float.fromhex(v) for v in probe(f"real {seed} {n} {a!r} {b!r}")])
    assert np.array_equal(ours, ref)


def test_canonical_clamps_to_below_one():
    lib = load("orc")
    # words that round to 2^64 as doubles must give nextafter(1, 0)
    top = 2 ** 64 - 1
    assert lib.orc_canonical53(top) == np.nextafter(1.0, 0.0)
    assert lib.orc_canonical53(0) == 0.0
    assert lib.orc_canonical53(2 ** 63) == 0.5


@pytest.mark.parametrize("hi", [0, 1, 2, 3, 15, 63, 1000, 2 ** 31 - 1])
def test_uniform_int_matches_libstdcxx(hi):
    lib = load("orc")
    n, seed = 500, 99
    words = np.zeros(4 * n, dtype=np.uint64)
    lib.orc_mt64_words(seed, 4 * n, words.ctypes.data_as(C.POINTER(C.c_uint64)))
    ours, k = [], 0
    used = C.c_int64()
    for _ in range(n):
        v = lib.orc_lemire64(words[k:].ctypes.data_as(C.POINTER(C.c_uint64)), 4 * n - k, hi + 1, C.byref(used))
        k += used.value
        ours.append(v)
    ref = [int(v) for v in probe(f"int {seed} {n} 0 {hi}")]
    assert ours == ref


def test_lemire_rejects_below_threshold():
    lib = load("orc")
    # range 3: threshold = 2^64 mod 3 = 1; word 0 gives low 0 < 1 -> rejected, next word used
    words = np.array([0, 2 ** 63], dtype=np.uint64)
    used = C.c_int64()
    v = lib.orc_lemire64(words.ctypes.data_as(C.POINTER(C.c_uint64)), 2, 3, C.byref(used))
    assert used.value == 2 and v == 1


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox4x32_10_known_answers(ctr, key, expect):
    lib = load("orc")
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    assert tuple(o) == expect


def test_det_sum_order_and_value():
    lib = load("orc")
    rng = np.random.default_rng(1)
    for n in [1, 2, 255, 256, 257, 1000, 65536]:
        v = rng.normal(size=n)
        s = lib.orc_det_sum(v.ctypes.data_as(C.POINTER(C.c_double)), n)
        # restated order: chunks of 256, pairwise tree, chunks in order
        tot = None
        for c0 in range(0, n, 256):
            buf = np.zeros(256)
            chunk = v[c0:c0 + 256]
            buf[:len(chunk)] = chunk
            st = 128
            while st > 0:
                buf[:st] = buf[:st] + buf[st:2 * st]
                st //= 2
            tot = buf[0] if tot is None else tot + buf[0]
        assert s == tot
        assert abs(s - np.sum(v)) <= 1e-12 * max(1.0, np.sum(np.abs(v)))
