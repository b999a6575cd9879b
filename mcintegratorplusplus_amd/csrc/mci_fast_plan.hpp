// mci_fast_plan.hpp — how the specialised kernels split a run into chunks.  Plain C++ with
// no HIP dependency (constexpr, so device code calls it too), so that the host-side tests can
// check the rule on runs far longer than a GPU test could take.
#pragma once

#include <cstdint>

namespace mcid
{
// Longest chunk the inner step loop runs (its counters are 32-bit).
constexpr int64_t FAST_MAX_CHUNK = (int64_t)1 << 30;

// Steps of the next chunk of a run of nsteps steps after s steps, inblk of them in the current
// block of B steps (B = nsteps for a Simple accumulator).  With accumulation (acc) a chunk never
// crosses a block end, so every block is flushed exactly when inblk reaches B; the result is in
// [1, 2^30] while s < nsteps.
constexpr int fast_chunk_len(int64_t nsteps, int64_t s, int64_t B, int64_t inblk, bool acc)
{
    int64_t lim = nsteps - s;
    if (acc && B - inblk < lim) { lim = B - inblk; }
    return (int)(lim < FAST_MAX_CHUNK ? lim : FAST_MAX_CHUNK);
}
// Bits of a native step's shared word (main word 0) that the index of a single-vector move over
// nvecs vectors takes: m when nvecs = 2^m with m <= 11 -- the FP64 acceptance uniform keeps the 53
// bits below them, the FP32 one (m <= 8) 24 -- else -1, and the index is drawn by Lemire's method
// from a word of its own.  Only the chain's own move shares (not one inside a MultiStepMove), and
// only in the native stream.
constexpr int shared_index_bits(bool philox, bool vec_move, bool multistep, int64_t nvecs)
{
    if (!philox || !vec_move || multistep || nvecs < 1 || (nvecs & (nvecs - 1)) != 0) { return -1; }
    int m = 0;
    while (((int64_t)1 << m) < nvecs) { ++m; }
    return m <= 11 ? m : -1;
}
// The FP32 chain shares only up to 8 bits (its 24-bit acceptance uniform sits below them).
constexpr int shared_index_bits32(int ibits) { return ibits <= 8 ? ibits : -1; }

// Steps between recomputations of |x|^2 from the position in the FP32 kernel.
constexpr int FAST32_REFRESH = 1024;

// Steps of the FP32 kernel's next chunk after s steps of nsteps, j of them in the current block of
// B (at most 2^31 - 1): at most FAST32_REFRESH, never past the block end.
constexpr int fast32_chunk_len(int64_t nsteps, int64_t s, int B, int j, bool acc)
{
    int64_t lim = nsteps - s;
    if (acc && (int64_t)(B - j) < lim) { lim = B - j; }
    return (int)(lim < FAST32_REFRESH ? lim : FAST32_REFRESH);
}
} // namespace mcid
