// Test infrastructure only (the oracle): the random numbers, sample positions
// and angle functions of the Monte Carlo render, restated on the CPU.
//
// The reference's Monte Carlo build (USE_MONTECARLO_SAMPLING, off in
// rt/StaticConfig.h:5) draws from cuRAND's Philox4_32_10 generator: one state
// per cell of a 256×256 grid, initialised with seed 42 and the cell index as
// subsequence (rt/RayTracer.cu:23-25,183-192), copied into each pixel whose
// (i mod 256, j mod 256) is that cell and advanced by every draw of the
// pixel's samples (RayTracer.cu:212-246); sample s of n looks through
// (i + s/n, j + vdc(s)) (PBR.h:145-157), and a bounce picks a GGX microfacet
// normal from two uniform draws (PBR.h:159-176, RayTracer.cu:473-490) with
// sinf, cosf and atan2f.
//
// cuRAND's state handling and CUDA's math functions are not restated.  This
// build's stream, defined here and by the same steps in the library:
//   * the generator is Philox4x32-10 as published (Salmon, Moraes, Dror and
//     Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11; the
//     Random123 library's philox4x32 at 10 rounds), pinned by that library's
//     known-answer vectors;
//   * draw m of a stream is word (m mod 4) of the block with counter
//     (m / 4, 0, stream, 0) and key (seed, 0); stream = (i mod 256) +
//     256 (j mod 256), the reference's grid cell;
//   * a draw w becomes the float (2 (w >> 9) + 1) 2^-24, in (0, 1) exactly;
//   * a bounce takes x then y (C++ leaves the reference's argument order
//     open);
//   * sin, cos and atan2 are evaluated in double precision (Cody-Waite
//     reduction and the Taylor series; a table of atan(k/8)) and rounded to
//     float: within an ulp of the correctly rounded value.
// Results are this build's and equal between library and oracle bit for
// bit; against the reference's images they agree in distribution only
// ("parity unpinned" for the draws themselves).
#pragma once

#include <cmath>
#include <cstdint>

namespace orc
{

// Philox4x32-10 (Salmon et al. 2011, Random123 philox4x32_R with R = 10):
// per round, the two 32×32→64 products of counter words 0 and 2 by the
// multipliers 0xD2511F53 and 0xCD9E8D57, and the key bumped by the Weyl
// constants 0x9E3779B9 and 0xBB67AE85 between rounds.
inline void philox4x32_10(const uint32_t ctrIn[4], const uint32_t keyIn[2], uint32_t out[4])
{
    uint32_t c0 = ctrIn[0], c1 = ctrIn[1], c2 = ctrIn[2], c3 = ctrIn[3];
    uint32_t k0 = keyIn[0], k1 = keyIn[1];
    for (int r = 0; r < 10; r++)
    {
        if (r > 0)
        {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// The random word of draw m of a stream.
inline uint32_t sampleWord(uint32_t seed, uint32_t stream, uint32_t m)
{
    const uint32_t ctr[4] {m >> 2, 0u, stream, 0u};
    const uint32_t key[2] {seed, 0u};
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    return out[m & 3];
}

// A word as a uniform float in (0, 1): 23 bits on the odd multiples of 2^-24.
inline float unitFloat(uint32_t w) { return (float)(2u * (w >> 9) + 1u) * 0x1p-24f; }

// vdc (PBR.h:145-157): the Van der Corput radical inverse in base 2, the bit
// reversal of i scaled by 2^-32.
inline float vdc(uint32_t i)
{
    uint32_t r = 0;
    for (int b = 0; b < 32; b++) r |= ((i >> b) & 1u) << (31 - b);
    return 0x1p-32f * (float)r;
}

// sin and cos of a float angle (|a| < 2^20): a - k π/2 with π/2 split in two
// parts (the first with 33 significant bits, so k times it is exact), then
// the Taylor series of sin to degree 17 and of cos to degree 18 on
// |r| <= π/4, in double precision.
inline void sinCosAngle(float a, float& s, float& c)
{
    const double x = a;
    const double k = std::rint(x * 0x1.45f306dc9c883p-1);
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

// atan2(y, x) for y >= 0, x >= 0, not both zero (the microfacet angle's
// quadrant): t = min/max in [0, 1], atan t = atan(k/8) + atan((t - k/8) /
// (1 + t k/8)) with k = rint(8 t), so the series argument is at most 1/16 and
// the Taylor series of atan to degree 15 suffices; π/2 - atan for y > x.
inline float atan2Quadrant(float y, float x)
{
    static constexpr double atanK8[9] {0.0,
                                       0x1.fd5ba9aac2f6ep-4,
                                       0x1.f5b75f92c80ddp-3,
                                       0x1.6f61941e4def1p-2,
                                       0x1.dac670561bb4fp-2,
                                       0x1.1e00babdefeb4p-1,
                                       0x1.4978fa3269ee1p-1,
                                       0x1.700a7c5784634p-1,
                                       0x1.921fb54442d18p-1};
    const double Y = y, X = x;
    const bool swap = Y > X;
    const double t = swap ? X / Y : Y / X;
    const double kd = std::rint(t * 8.0);
    const double ck = kd * 0.125;
    const double z = (t - ck) / (1.0 + t * ck);
    const double z2 = z * z;
    const double p =
        -0x1.5555555555555p-2 +
        z2 * (0x1.999999999999ap-3 +
              z2 * (-0x1.2492492492492p-3 +
                    z2 * (0x1.c71c71c71c71cp-4 +
                          z2 * (-0x1.745d1745d1746p-4 + z2 * (0x1.3b13b13b13b14p-4 + z2 * -0x1.1111111111111p-4)))));
    double a = atanK8[(int)kd] + (z + z * z2 * p);
    if (swap) a = 0x1.921fb54442d18p+0 - a;
    return (float)a;
}

} // namespace orc
