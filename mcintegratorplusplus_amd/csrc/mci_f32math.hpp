// mci_f32math.hpp — this library's own FP32 exp and log and the symmetric 24-bit grid, written
// with correctly rounded operations only (fused multiply-add, rint, IEEE divide), so that a CPU
// evaluating the same operations in the same order reproduces them bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcif
{
// a ^ b ^ c in one instruction (v_bitop3_b32 with the truth table of a three-input xor): the
// Philox rounds' xors of a product half, a counter word and a key word
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }

// e^x: Cody–Waite reduction by ln 2 (two-part constant, fused multiply-add), the degree-7
// Taylor polynomial of e^r on |r| <= ln2/2 (Horner, fused), exact scaling by 2^n.
// 0 below -87 (and for NaN), +inf above 88; at most 0.92 ulp from e^x on [-87, 88].
// Branch-free: the range cases are selected at the end.
__device__ __forceinline__ float bexp(float x)
{
    const float xc = fminf(fmaxf(x, -87.0f), 88.0f);
    const float n = rintf(xc * 0x1.715476p+0f);
    float r = fmaf(-n, 0x1.62e43p-1f, xc);
    r = fmaf(-n, -0x1.05c61p-29f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0x1p-1f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    float res = p * __uint_as_float((uint32_t)((int)n + 127) << 23);
    res = (x >= -87.0f) ? res : 0.0f;
    return (x > 88.0f) ? __builtin_huge_valf() : res;
}

// log x for positive normal x: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s)
// with s = (m - 1)/(m + 1) as an odd series to s^9 (Horner in s^2, fused); about 1 ulp.
__device__ __forceinline__ float blog(float x)
{
    const uint32_t b = __float_as_uint(x);
    int e = (int)(b >> 23) - 127;
    float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float z = s * s;
    float p = 0x1.c71c72p-4f;
    p = fmaf(p, z, 0x1.24924ap-3f);
    p = fmaf(p, z, 0x1.99999ap-3f);
    p = fmaf(p, z, 0x1.555556p-2f);
    const float ls = s + s;
    const float lm = fmaf(ls * z, p, ls);
    const float fe = (float)e;
    return fmaf(fe, 0x1.62e43p-1f, fmaf(fe, -0x1.05c61p-29f, lm));
}

// sin and cos of 2 pi u for u in [0, 1): the quadrant from 4u (exact), the angle within it
// t = (4u - quadrant) pi/2 in [0, pi/2), then sin t and cos t by their Taylor series to t^11 and
// t^12 (Horner in t^2, fused; truncation below 6e-8 on the quadrant), signs and roles swapped by
// quadrant.  A few ulp from the true values.
__device__ __forceinline__ void bsincos2pi(float u, float &s, float &c)
{
    const float t4 = 4.0f * u;
    const float q = floorf(t4);
    const float th = (t4 - q) * 1.57079632679f;
    const float z = th * th;
    float ps = -2.50521084e-8f;
    ps = fmaf(ps, z, 2.75573192e-6f);
    ps = fmaf(ps, z, -1.98412698e-4f);
    ps = fmaf(ps, z, 8.33333333e-3f);
    ps = fmaf(ps, z, -1.66666667e-1f);
    ps = fmaf(ps, z, 1.0f);
    const float sn = th * ps;
    float pc = 2.08767570e-9f;
    pc = fmaf(pc, z, -2.75573192e-7f);
    pc = fmaf(pc, z, 2.48015873e-5f);
    pc = fmaf(pc, z, -1.38888889e-3f);
    pc = fmaf(pc, z, 4.16666667e-2f);
    pc = fmaf(pc, z, -0.5f);
    pc = fmaf(pc, z, 1.0f);
    const int qi = (int)q;
    s = (qi == 0) ? sn : (qi == 1) ? pc : (qi == 2) ? -sn : -pc;
    c = (qi == 0) ? pc : (qi == 1) ? -sn : (qi == 2) ? -pc : sn;
}

// symmetric grid (2m + 1 - 2^24) 2^-24 in (-1, 1) from the high 24 bits of a word; never 0
__device__ __forceinline__ float sym24(uint32_t w) { return (float)(2 * (int32_t)(w >> 8) + 1 - 16777216) * 0x1p-24f; }

} // namespace mcif
