"""The chunk rule of the specialised FP64 kernel (mcintegratorplusplus_amd/csrc/mci_fast_plan.hpp),
checked on the host for runs far longer than a GPU test could take: blocks (a Simple
accumulator's block is the whole run) of more than 2^31 steps must still be flushed exactly once
per block end, with every inner loop at most 2^30 steps."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "mcintegratorplusplus_amd", "csrc")

PROG = r"""
#include "mci_fast_plan.hpp"
#include <cstdio>
#include <cstdlib>
int main(int argc, char **argv)
{
    const int64_t nsteps = std::atoll(argv[1]), B = std::atoll(argv[2]);
    const bool acc = std::atoi(argv[3]) != 0;
    int64_t s = 0, inblk = 0, flushes = 0, chunks = 0, maxlen = 0, minlen = INT64_MAX;
    while (s < nsteps) {
        const int len = mcid::fast_chunk_len(nsteps, s, B, inblk, acc);
        if (len < 1) { std::printf("bad %d\n", len); return 1; }
        maxlen = len > maxlen ? len : maxlen;
        minlen = len < minlen ? len : minlen;
        s += len;
        inblk += len;
        ++chunks;
        if (acc && inblk > B) { std::printf("crossed\n"); return 1; }
        if (acc && inblk == B) { ++flushes; inblk = 0; }
    }
    std::printf("%lld %lld %lld %lld %lld\n", (long long)s, (long long)flushes, (long long)chunks, (long long)maxlen, (long long)minlen);
    return 0;
}
"""


@pytest.fixture(scope="module")
def plan(tmp_path_factory):
    d = tmp_path_factory.mktemp("plan")
    src = d / "plan.cpp"
    src.write_text(PROG)
    exe = d / "plan"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, str(src), "-o", str(exe)], check=True)

    def run(nsteps, B, acc):
        out = subprocess.run([str(exe), str(nsteps), str(B), str(int(acc))], capture_output=True, text=True, check=True)
        return [int(v) for v in out.stdout.split()]
    return run


@pytest.mark.parametrize("nsteps,B,acc", [
    (10240, 160, True),              # the benchmark's blocks
    (2 ** 31 + 7, 2 ** 31 + 7, True),  # Simple accumulator beyond 2^31 (wrapped negative as an int)
    (2 ** 32 + 3, 2 ** 32 + 3, True),  # Simple accumulator beyond 2^32 (wrapped to 3 as an int)
    (3 * (2 ** 31 - 1), 2 ** 31 - 1, True),  # the largest Block size, three blocks
    (5, 5, True),
    (2 ** 33 + 1, 1, False),         # no accumulation (decorrelation)
])
def test_chunks_stay_32bit_and_flush_at_block_ends(plan, nsteps, B, acc):
    s, flushes, chunks, maxlen, minlen = plan(nsteps, B, acc)
    assert s == nsteps
    assert 1 <= minlen and maxlen <= 2 ** 30
    if acc:
        assert flushes == nsteps // B
    assert chunks >= -(-nsteps // 2 ** 30)


BITS = r"""
#include "mci_fast_plan.hpp"
#include <cstdio>
int main()
{
    for (long long n = 0; n <= 5000; ++n)
        std::printf("%lld %d %d %d %d\n", n, mcid::shared_index_bits(true, true, false, n), mcid::shared_index_bits(false, true, false, n),
                    mcid::shared_index_bits(true, false, false, n), mcid::shared_index_bits(true, true, true, n));
    return 0;
}
"""


def test_shared_index_bits_match_the_oracle(tmp_path):
    """The index bits the library takes from a native step's shared word (mci_fast_plan.hpp) and the
    oracle's rule (orc_shared_index_bits) agree: m for 2^m vectors up to 2^11, else -1 (Lemire);
    only native streams, vector moves, outside a MultiStepMove."""
    import ctypes as C

    from mcitest import Chain
    src = tmp_path / "bits.cpp"
    src.write_text(BITS)
    exe = tmp_path / "bits"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    rows = [list(map(int, ln.split())) for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n") if ln]
    for n, m, mt, allmove, ms in rows:
        exp = -1
        if n >= 1 and n & (n - 1) == 0 and n.bit_length() - 1 <= 11:
            exp = n.bit_length() - 1
        assert (m, mt, allmove, ms) == (exp, -1, -1, -1), n
    # the oracle, through contexts of 2^k coordinates with single-coordinate moves
    from mcitest import RNG_PHILOX, load
    lib = load("orc")
    lib.orc_shared_index_bits.argtypes = [C.c_void_p]
    for d in (1, 2, 3, 16, 24, 2048, 4096):
        c = Chain("orc", d, 1)
        c.setRng(RNG_PHILOX)
        c.setTrialMove(0, 1, 1, None)
        assert lib.orc_shared_index_bits(c.ctx) == rows[d][1], d
        c.setTrialMoveMultistep(2, 0, 1, 1, None)
        assert lib.orc_shared_index_bits(c.ctx) == -1
