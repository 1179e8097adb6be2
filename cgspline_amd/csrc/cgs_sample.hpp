// Random numbers, sample positions and angle functions of the sampled
// (Monte Carlo) render, device side.
//
// The reference's Monte Carlo build (USE_MONTECARLO_SAMPLING, rt/StaticConfig.h:5)
// copies a Philox state per pixel from a 256×256 grid seeded with 42
// (rt/RayTracer.cu:23-25,183-192,212-218), looks through (i + s/n, j + vdc(s))
// for sample s of n (RayTracer.cu:229-246, PBR.h:145-157), and picks a GGX
// microfacet normal for each bounce from two uniform draws with atan2f, sinf
// and cosf (PBR.h:159-176, RayTracer.cu:473-490).  Here:
//   * the generator is Philox4x32-10 (Salmon, Moraes, Dror and Shaw,
//     "Parallel random numbers: as easy as 1, 2, 3", SC'11; Random123's
//     philox4x32 at 10 rounds);
//   * draw m of a pixel's stream is word (m mod 4) of the block with counter
//     (m / 4, 0, stream, 0) and key (seed, 0), stream = (i mod 256) +
//     256 (j mod 256); a bounce takes draws m and m + 1 with m even, two
//     words of one block;
//   * a word w is the float (2 (w >> 9) + 1) 2^-24, in (0, 1);
//   * sin, cos and atan2 run in double precision (Cody-Waite reduction and
//     Taylor series; a table of atan(k/8)) and round to float.
// Every step is plain IEEE arithmetic without contraction, so the host
// restatement of the same steps gives the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#ifndef CGS_DEV
#define CGS_DEV __device__ __forceinline__
#endif

namespace cgs
{

// Philox4x32-10: per round the 32×32→64 products of counter words 0 and 2 by
// 0xD2511F53 and 0xCD9E8D57; the key advances by 0x9E3779B9 and 0xBB67AE85
// between rounds.
CGS_DEV uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++)
    {
        if (r > 0)
        {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}

CGS_DEV float unitFloat(uint32_t w) { return (float)(2u * (w >> 9) + 1u) * 0x1p-24f; }

// Draws m and m + 1 (m even) of a stream as floats in (0, 1).
CGS_DEV void drawPair(uint32_t seed, uint32_t stream, uint32_t m, float& x, float& y)
{
    const uint4 b = philox4x32_10(make_uint4(m >> 2, 0u, stream, 0u), seed, 0u);
    const bool hi = (m & 2u) != 0;
    x = unitFloat(hi ? b.z : b.x);
    y = unitFloat(hi ? b.w : b.y);
}

// The base-2 radical inverse: the bit reversal of i times 2^-32.
CGS_DEV float vdc(uint32_t i) { return 0x1p-32f * (float)__builtin_bitreverse32(i); }

// sin and cos of a float angle, |a| < 2^20: a - k π/2 with π/2 in two parts
// (33 significant bits in the first, so k times it is exact), then sin's
// Taylor series to degree 17 and cos's to degree 18 on |r| <= π/4.
CGS_DEV void sinCosAngle(float a, float& s, float& c)
{
    const double x = a;
    const double k = __builtin_rint(x * 0x1.45f306dc9c883p-1);
    const double r = (x - k * 0x1.921fb544p+0) - k * 0x1.0b4611a626331p-34;
    const double r2 = r * r;
    const double ps =
        -0x1.5555555555555p-3 +
        r2 * (0x1.1111111111111p-7 +
              r2 * (-0x1.a01a01a01a01ap-13 +
                    r2 * (0x1.71de3a556c734p-19 +
                          r2 * (-0x1.ae64567f544e4p-26 +
                                r2 * (0x1.6124613a86d09p-33 + r2 * (-0x1.ae7f3e733b81fp-41 + r2 * 0x1.952c77030ad4ap-49))))));
    const double pc =
        -0x1p-1 +
        r2 * (0x1.5555555555555p-5 +
              r2 * (-0x1.6c16c16c16c17p-10 +
                    r2 * (0x1.a01a01a01a01ap-16 +
                          r2 * (-0x1.27e4fb7789f5cp-22 +
                                r2 * (0x1.1eed8eff8d898p-29 +
                                      r2 * (-0x1.93974a8c07c9dp-37 + r2 * (0x1.ae7f3e733b81fp-45 + r2 * -0x1.6827863b97d97p-53)))))));
    const double sr = r + r * r2 * ps;
    const double cr = 1.0 + r2 * pc;
    const int q = (int)((int64_t)k & 3);
    const double sv = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    const double cv = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
    s = (float)sv;
    c = (float)cv;
}

// atan2(y, x) for y, x >= 0, not both zero: t = min/max, atan t =
// atan(k/8) + atan((t - k/8) / (1 + t k/8)) with k = rint(8 t), the series
// argument at most 1/16, atan's Taylor series to degree 15; π/2 - that for
// y > x.
CGS_DEV float atan2Quadrant(float y, float x)
{
    const double Y = y, X = x;
    const bool swap = Y > X;
    const double t = swap ? X / Y : Y / X;
    const double kd = __builtin_rint(t * 8.0);
    const int k = (int)kd;
    const double atanK = k == 0 ? 0.0
                       : k == 1 ? 0x1.fd5ba9aac2f6ep-4
                       : k == 2 ? 0x1.f5b75f92c80ddp-3
                       : k == 3 ? 0x1.6f61941e4def1p-2
                       : k == 4 ? 0x1.dac670561bb4fp-2
                       : k == 5 ? 0x1.1e00babdefeb4p-1
                       : k == 6 ? 0x1.4978fa3269ee1p-1
                       : k == 7 ? 0x1.700a7c5784634p-1
                                : 0x1.921fb54442d18p-1;
    const double ck = kd * 0.125;
    const double z = (t - ck) / (1.0 + t * ck);
    const double z2 = z * z;
    const double p =
        -0x1.5555555555555p-2 +
        z2 * (0x1.999999999999ap-3 +
              z2 * (-0x1.2492492492492p-3 +
                    z2 * (0x1.c71c71c71c71cp-4 +
                          z2 * (-0x1.745d1745d1746p-4 + z2 * (0x1.3b13b13b13b14p-4 + z2 * -0x1.1111111111111p-4)))));
    double a = atanK + (z + z * z2 * p);
    if (swap) a = 0x1.921fb54442d18p+0 - a;
    return (float)a;
}

} // namespace cgs
