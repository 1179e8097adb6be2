// SPDX-License-Identifier: MIT
//
// Device building blocks for gfx950 (wave64): ray–bicubic-patch Bézier
// clipping with the reference's host semantics, the hashed stackless BVH
// traversal and the per-object ray transform.
//
// Numerics.  The library is compiled with -ffp-contract=off and IEEE
// (correctly rounded) f32 division and square root, and every expression keeps
// the reference's evaluation order, so an FP32 result equals the one the
// reference's host path computes for the same inputs bit for bit.  Keep that
// in mind before "simplifying" an expression here: a*b+c*d written in another
// order, a dropped "+ 0.0f" or a reciprocal multiply in place of a division
// changes results.
//
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cgs.h"

#define CGS_DEV __device__ __forceinline__

namespace cgs
{

constexpr uint32_t EMPTY = 0xFFFFFFFFu;

// Event counts of a launch, for the work model: a build with
// CGS_WORK_COUNTS=1 counts each event in a per-workgroup LDS slot, and the
// kernels that call workBegin/workEnd add the slots to a device total.  In
// the default build every CGS_COUNT is empty.
#ifndef CGS_WORK_COUNTS
#define CGS_WORK_COUNTS 0
#endif
enum WorkEvent : int
{
    kWorkCullNode,    // cull-tree node box tested
    kWorkCullElem,    // cull-tree element (object leaf box) tested
    kWorkNode,        // BVH inner node visited (two child boxes tested)
    kWorkProbe,       // perfect-hash lookup on an ascent
    kWorkLeaf,        // leaf element visited
    kWorkPatch,       // patch test (projection)
    kWorkClip,        // clipping pass that clipped the region
    kWorkSplit,       // clipping pass that split it
    kWorkPrune,       // clipping pass whose hull missed
    kWorkConverged,   // region under tolerance
    kWorkOnHit,       // candidate evaluated
    kWorkObject,      // object-local ray set up
    kWorkEvents
};
#if CGS_WORK_COUNTS
CGS_DEV unsigned long long* workSlots()
{
    __shared__ unsigned long long slots[kWorkEvents];
    return slots;
}
#define CGS_COUNT(e) atomicAdd(::cgs::workSlots() + (e), 1ull)
#else
#define CGS_COUNT(e) ((void)0)
#endif
constexpr float FLT_EPS = 1.1920928955078125e-07f; // 2^-23

struct V2 { float x, y; };
struct V3 { float x, y, z; };

// glibc x86-64 fminf/fmaxf (the reference's host fmin/fmax): a NaN operand
// yields the other operand, ties return the second operand.
CGS_DEV float fmin_(float a, float b) { return (a < b || b != b) ? a : b; }
CGS_DEV float fmax_(float a, float b) { return (a > b || b != b) ? a : b; }

CGS_DEV V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
CGS_DEV V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CGS_DEV V3 mul3(float s, V3 v) { return {s * v.x, s * v.y, s * v.z}; }
CGS_DEV V3 mulv3(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
CGS_DEV V3 neg3(V3 a) { return {-a.x, -a.y, -a.z}; }
CGS_DEV float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CGS_DEV V3 cross3(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
CGS_DEV float len3(V3 v) { return __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
CGS_DEV float max3(V3 v) { return fmax_(v.x, fmax_(v.y, v.z)); }

// Vector::versor (Ds): scale by the reciprocal length unless the length is
// within machine epsilon of zero.
CGS_DEV V3 versor3(V3 v)
{
    float len = len3(v);
    if (!(__builtin_fabsf(len) <= FLT_EPS))
    {
        float inv = 1.0f / len;
        v = {v.x * inv, v.y * inv, v.z * inv};
    }
    return v;
}
CGS_DEV V2 normalize2(V2 v)
{
    float len = __builtin_sqrtf(v.x * v.x + v.y * v.y);
    if (!(__builtin_fabsf(len) <= FLT_EPS))
    {
        float inv = 1.0f / len;
        v = {v.x * inv, v.y * inv};
    }
    return v;
}

// 4×4 column-major matrix times (x, y, z, w) (AlignedTypes.h:185-188).
struct M4
{
    float4 c[4];
};
CGS_DEV float4 m4mul(const M4& m, float4 v)
{
    float4 r;
    r.x = ((v.x * m.c[0].x + v.y * m.c[1].x) + v.z * m.c[2].x) + v.w * m.c[3].x;
    r.y = ((v.x * m.c[0].y + v.y * m.c[1].y) + v.z * m.c[2].y) + v.w * m.c[3].y;
    r.z = ((v.x * m.c[0].z + v.y * m.c[1].z) + v.z * m.c[2].z) + v.w * m.c[3].z;
    r.w = ((v.x * m.c[0].w + v.y * m.c[1].w) + v.z * m.c[2].w) + v.w * m.c[3].w;
    return r;
}
CGS_DEV M4 loadM4(const float4* __restrict__ p)
{
    M4 m;
    m.c[0] = p[0]; m.c[1] = p[1]; m.c[2] = p[2]; m.c[3] = p[3];
    return m;
}
// mat3(M) * v
CGS_DEV V3 m3mul(const M4& m, V3 v)
{
    return {(v.x * m.c[0].x + v.y * m.c[1].x) + v.z * m.c[2].x,
            (v.x * m.c[0].y + v.y * m.c[1].y) + v.z * m.c[2].y,
            (v.x * m.c[0].z + v.y * m.c[1].z) + v.z * m.c[2].z};
}
// mat3(M).transposed() * v: columns of the transposed are the rows of M.
CGS_DEV V3 m3tmul(const M4& m, V3 v)
{
    return {(v.x * m.c[0].x + v.y * m.c[0].y) + v.z * m.c[0].z,
            (v.x * m.c[1].x + v.y * m.c[1].y) + v.z * m.c[1].z,
            (v.x * m.c[2].x + v.y * m.c[2].y) + v.z * m.c[2].z};
}
CGS_DEV V3 project4(float4 p) { return {p.x / p.w, p.y / p.w, p.z / p.w}; }

// ---------------------------------------------------------------------------
// Bicubic evaluation (SplineMat.h:22-44, 101-187), power-basis matrix form
// with every term of the constant matrices kept (so signed zeros and the
// rounding of each partial sum match the reference).
// ---------------------------------------------------------------------------

CGS_DEV float4 basisMul(float4 v) // CUBIC_BASE * v
{
    float4 r;
    r.x = ((v.x * 1.0f + v.y * -3.0f) + v.z * 3.0f) + v.w * -1.0f;
    r.y = ((v.x * 0.0f + v.y * 3.0f) + v.z * -6.0f) + v.w * 3.0f;
    r.z = ((v.x * 0.0f + v.y * 0.0f) + v.z * 3.0f) + v.w * -3.0f;
    r.w = ((v.x * 0.0f + v.y * 0.0f) + v.z * 0.0f) + v.w * 1.0f;
    return r;
}
CGS_DEV float4 dbasisMul(float4 v) // CUBIC_DERIVATIVE_BASE * v
{
    float4 r;
    r.x = ((v.x * -3.0f + v.y * 6.0f) + v.z * -3.0f) + v.w * 0.0f;
    r.y = ((v.x * 3.0f + v.y * -12.0f) + v.z * 9.0f) + v.w * 0.0f;
    r.z = ((v.x * 0.0f + v.y * 6.0f) + v.z * -9.0f) + v.w * 0.0f;
    r.w = ((v.x * 0.0f + v.y * 0.0f) + v.z * 3.0f) + v.w * 0.0f;
    return r;
}
CGS_DEV float4 evalVec(float t) { float t2 = t * t; return {1.0f, t, t2, t2 * t}; }
CGS_DEV float4 evalVecDt(float t) { return {1.0f, t, t * t, 0.0f}; }

// x.x*a + x.y*b + x.z*c + x.w*d over float4 points (mat4(a,b,c,d) * x)
CGS_DEV float4 comb4(float4 a, float4 b, float4 c, float4 d, float4 x)
{
    float4 r;
    r.x = ((x.x * a.x + x.y * b.x) + x.z * c.x) + x.w * d.x;
    r.y = ((x.x * a.y + x.y * b.y) + x.z * c.y) + x.w * d.y;
    r.z = ((x.x * a.z + x.y * b.z) + x.z * c.z) + x.w * d.z;
    r.w = ((x.x * a.w + x.y * b.w) + x.z * c.w) + x.w * d.w;
    return r;
}

// mat::interpolate (SplineMat.h:125-140)
CGS_DEV float4 interpolateMat(const float4* __restrict__ S, float u, float v)
{
    float4 x = basisMul(evalVec(u));
    float4 y = basisMul(evalVec(v));
    float4 a0 = comb4(S[0], S[1], S[2], S[3], x);
    float4 a1 = comb4(S[4], S[5], S[6], S[7], x);
    float4 a2 = comb4(S[8], S[9], S[10], S[11], x);
    float4 a3 = comb4(S[12], S[13], S[14], S[15], x);
    return comb4(a0, a1, a2, a3, y);
}
// mat::derivativeU / derivativeV / normal (SplineMat.h:142-187)
CGS_DEV V3 normalMat(const float4* __restrict__ S, float u, float v)
{
    float4 x = dbasisMul(evalVecDt(u));
    float4 y = basisMul(evalVec(v));
    float4 a0 = comb4(S[0], S[1], S[2], S[3], x);
    float4 a1 = comb4(S[4], S[5], S[6], S[7], x);
    float4 a2 = comb4(S[8], S[9], S[10], S[11], x);
    float4 a3 = comb4(S[12], S[13], S[14], S[15], x);
    float4 dU = comb4(a0, a1, a2, a3, y);
    float4 xb = basisMul(evalVec(u));
    float4 yd = dbasisMul(evalVecDt(v));
    float4 b0 = comb4(S[0], S[4], S[8], S[12], yd);
    float4 b1 = comb4(S[1], S[5], S[9], S[13], yd);
    float4 b2 = comb4(S[2], S[6], S[10], S[14], yd);
    float4 b3 = comb4(S[3], S[7], S[11], S[15], yd);
    float4 dV = comb4(b0, b1, b2, b3, xb);
    return cross3({dU.x, dU.y, dU.z}, {dV.x, dV.y, dV.z});
}

// ---------------------------------------------------------------------------
// Curve slicing (Spline.h:324-364): split1 at u, then split0 at (v-u)/(1-u).
// Both coordinates of a 2D point are sliced with the same scalars.
// ---------------------------------------------------------------------------
CGS_DEV void slice1(float& c0, float& c1, float& c2, float& c3, float u, float iu, float vv, float ivv)
{
    float p5, p6, p8;
    p6 = u * c1 + iu * c0;
    p5 = u * c2 + iu * c1;
    c2 = u * c3 + iu * c2;
    p8 = u * p5 + iu * p6;
    c1 = u * c2 + iu * p5;
    c0 = u * c1 + iu * p8;
    p6 = ivv * c2 + vv * c3;
    p5 = ivv * c1 + vv * c2;
    c1 = ivv * c0 + vv * c1;
    p8 = ivv * p5 + vv * p6;
    c2 = ivv * c1 + vv * p5;
    c3 = ivv * c2 + vv * p8;
}

// subpatchU (rows, Spline.h:366-373) of a 2D patch held in registers; every
// point index is a compile-time constant so the patch never leaves VGPRs.
// subpatchV (columns, :375-382) is the same slice applied to the columns: the
// clipper gets it by keeping the patch transposed while cutting in v.
CGS_DEV void subpatchU2(V2 (&p)[16], float a, float b)
{
    float iu = 1.0f - a;
    float vv = (b - a) / iu;
    float ivv = 1.0f - vv;
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        slice1(p[4 * k].x, p[4 * k + 1].x, p[4 * k + 2].x, p[4 * k + 3].x, a, iu, vv, ivv);
        slice1(p[4 * k].y, p[4 * k + 1].y, p[4 * k + 2].y, p[4 * k + 3].y, a, iu, vv, ivv);
    }
}
// ---------------------------------------------------------------------------
// Hull helpers (Spline_BezierClipping.cu:127-165, Spline.h:709-737)
// ---------------------------------------------------------------------------
CGS_DEV bool xAxisIntersection(float& lo, float& hi, V2 A, V2 B)
{
    float q = A.y - B.y;
    if (__builtin_fabsf(q) > FLT_EPS)
    {
        float t = A.y / q;
        if (0.0f <= t && t <= 1.0f)
        {
            float s = (1.0f - t) * A.x + t * B.x;
            lo = s;
            hi = s;
            return true;
        }
    }
    else if (__builtin_fabsf(A.y) <= FLT_EPS)
    {
        lo = fmin_(A.x, B.x);
        hi = fmax_(A.x, B.x);
        return true;
    }
    return false;
}

CGS_DEV float isCCW(V2 P, V2 Q) { return P.x * Q.y - P.y * Q.x; }

CGS_DEV bool triangleOrigin(V2& coord, V2 P, V2 A, V2 B)
{
    V2 p {A.x - P.x, A.y - P.y};
    V2 q {B.x - P.x, B.y - P.y};
    float det = p.x * q.y - q.x * p.y;
    if (__builtin_fabsf(det) < 0x1p-80f) return false;
    det = 1.0f / det;
    // m0 = (q.y, -q.x), m1 = (-p.y, p.x); dot with -P
    float u = det * (q.y * -P.x + -q.x * -P.y);
    if (u < 0.0f || u > 1.0f) return false;
    float v = det * (-p.y * -P.x + p.x * -P.y);
    if (v < 0.0f || u + v > 1.0f) return false;
    coord = {u, v};
    return true;
}

// ---------------------------------------------------------------------------
// Ray frame for clipping: the patch is projected on two planes containing the
// ray (Spline_BezierClipping.cu:57-83).  `dir` is the Ray3f direction.
// ---------------------------------------------------------------------------
struct ClipRay
{
    V3 o, d, ax0, ax1;
};
CGS_DEV ClipRay makeClipRay(V3 o, V3 dir)
{
    ClipRay r;
    r.o = o;
    r.d = dir;
    V3 e {dir.x + 0.0f, dir.y + 0.0f, dir.z + 8.0f};
    r.ax0 = versor3(cross3(dir, e));
    r.ax1 = versor3(cross3(dir, r.ax0));
    return r;
}

// Region stack: per entry the 16 working points and the region (min, max,
// size, side) — exactly the state the reference's host clipper keeps per
// std::vector entry.  It lives in a library-owned HBM buffer, one slice per
// resident lane, interleaved by lane so the 64 lanes of a wave pushing the
// same word touch one contiguous 512-B run.  (In private memory the 5 KB per
// lane would make the runtime cap the number of resident waves.)
constexpr int kStackWords = 20; // float2 words per entry: 16 points + 3 region + pad
struct ClipStack
{
    float2* __restrict__ base; // this lane's first word
    uint32_t stride;           // lanes sharing the buffer
    CGS_DEV void put(int sp, int w, float2 v) const { base[(size_t)(sp * kStackWords + w) * stride] = v; }
    CGS_DEV float2 get(int sp, int w) const { return base[(size_t)(sp * kStackWords + w) * stride]; }
};
CGS_DEV ClipStack clipStackFor(float2* buf, uint32_t lane, uint32_t lanes) { return {buf + lane, lanes}; }

// doBezierClipping with host semantics (Spline_BezierClipping.cu:46-114,
// 174-445).  S: the 16 control points (global).  best/bu/bv: current hit,
// updated by accepted candidates (t < best && V·d > 0).  With ANY the test
// stops at the first accepted candidate (the caller only needs "found").
// Returns found; sets overflow when the region stack is exhausted.
// CAPPED (an opt-in mode with no counterpart in the reference): the test
// stops after maxIters passes of its loop and keeps what it found so far.
template<bool ANY, bool CAPPED = false>
CGS_DEV bool clipPatch(const float4* __restrict__ S, const ClipRay& R, float tol, float& best, float& bu, float& bv,
                       const ClipStack& stk, bool& overflow, uint32_t maxIters = 0)
{
    CGS_COUNT(kWorkPatch);
    V2 p[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
    {
        float4 q = S[i];
        V3 D {q.x - R.o.x, q.y - R.o.y, q.z - R.o.z};
        p[i] = {dot3(R.ax0, D), dot3(R.ax1, D)};
    }
    V2 mn {0.0f, 0.0f}, mx {1.0f, 1.0f}, sz {1.0f, 1.0f};
    bool vside = false;
    int sp = 0; // entries stored below the current one
    bool found = false;

    auto onHit = [&](float eu, float ev) -> bool
    {
        CGS_COUNT(kWorkOnHit);
        float4 X = interpolateMat(S, eu, ev);
        V3 V = sub3(project4(X), R.o);
        float t = len3(V);
        if (t < best && dot3(V, R.d) > 0.0f)
        {
            best = t;
            bu = eu;
            bv = ev;
            return true;
        }
        return false;
    };

    // The working patch is kept in the orientation of the current cut side:
    // natural (p[4j+i] = point(i,j)) while clipping in u, transposed while
    // clipping in v, so the curves being cut are always the rows p[4r..4r+3]
    // and one code path serves both sides.  Transposing moves values only;
    // every formula below gives the natural-orientation result bit for bit:
    // the tolerance diagonals p3-p12 and p15-p0 change at most by an exact
    // negation, L = (p15-p0) + (p12-p3) in transposed storage equals
    // (p15-p0) - (p12-p3) in natural storage (x + (-y) == x - y), and the
    // envelope visits the same points in the same order.
    uint32_t iters = 0;
    while (true)
    {
        if (CAPPED && iters++ >= maxIters) break;
        bool pop = false;
        V2 t0 {p[3].x - p[12].x, p[3].y - p[12].y};
        V2 t1 {p[15].x - p[0].x, p[15].y - p[0].y};
        float d0 = __builtin_sqrtf(t0.x * t0.x + t0.y * t0.y);
        float d1 = __builtin_sqrtf(t1.x * t1.x + t1.y * t1.y);
        if (d0 < tol && d1 < tol)
        {
            CGS_COUNT(kWorkConverged);
            // approximateIntersection (Spline.h:739-765) on the natural corners
            const V2 c30 = vside ? p[12] : p[3], c03 = vside ? p[3] : p[12];
            V2 c;
            if (triangleOrigin(c, p[0], c30, c03))
                found |= onHit(mn.x + c.x * sz.x, mn.y + c.y * sz.y);
            if (triangleOrigin(c, p[15], c03, c30))
                found |= onHit(mx.x - c.x * sz.x, mx.y - c.y * sz.y);
            if (ANY && found) break;
            pop = true;
        }
        else
        {
            V2 L {p[15].x - p[0].x, p[15].y - p[0].y};
            const V2 Dv {p[12].x - p[3].x, p[12].y - p[3].y};
            L = {L.x + Dv.x, L.y + Dv.y};
            L = normalize2(L);
            const V2 N {-L.y, L.x};
            // signed distances N·p folded into the envelope as they are computed
            float ytop[4], ybot[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const float d0 = N.x * p[i].x + N.y * p[i].y;
                ybot[i] = d0;
                ytop[i] = d0;
#pragma unroll
                for (int k = 1; k < 4; k++)
                {
                    const float dk = N.x * p[4 * k + i].x + N.y * p[4 * k + i].y;
                    ytop[i] = fmax_(ytop[i], dk);
                    ybot[i] = fmin_(ybot[i], dk);
                }
            }
            float lower = 1.0f, upper = 0.0f;
            const float o1 = 1.0f / 3.0f, o2 = 2.0f / 3.0f;
            const V2 hull[7] {{o2, ybot[2]}, {1.0f, ybot[3]}, {1.0f, ytop[3]}, {o2, ytop[2]},
                              {o1, ytop[1]}, {0.0f, ytop[0]}, {0.0f, ybot[0]}};
            V2 P1 {0.0f, ybot[0]};
            V2 P2 {o1, ybot[1]};
#pragma unroll
            for (int h = 0; h < 7; h++)
            {
                const V2 P3 = hull[h];
                if (isCCW({P2.x - P1.x, P2.y - P1.y}, {P3.x - P1.x, P3.y - P1.y}) >= 0.0f)
                {
                    float lo, hi;
                    if (xAxisIntersection(lo, hi, P1, P2))
                    {
                        lower = fmin_(lower, lo);
                        upper = fmax_(upper, hi);
                    }
                    P1 = P2;
                }
                P2 = P3;
            }
            {
                float lo, hi;
                if (xAxisIntersection(lo, hi, P1, P2))
                {
                    lower = fmin_(lower, lo);
                    upper = fmax_(upper, hi);
                }
            }
            lower = lower * 0x0FFp-8f;
            upper = upper + (1.0f - upper) * 0x001p-8f;
            const float delta = upper - lower;
            if (delta < 0.0f)
            {
                CGS_COUNT(kWorkPrune);
                pop = true;
            }
            else
            {
                // delta > 0.8: split in half, keep the upper half and push the
                // lower half (Spline_BezierClipping.cu:373-409); otherwise clip
                // to [lower, upper] (:410-432).  Both keep a slice of the rows.
                const bool split = delta > 0.8f;
                CGS_COUNT(split ? kWorkSplit : kWorkClip);
                if (split && sp >= CGS_CLIP_STACK)
                {
                    overflow = true;
                    break;
                }
                // region along the cut side (u: x, v: y)
                const float smn = vside ? mn.y : mn.x, smx = vside ? mx.y : mx.x, ssz = vside ? sz.y : sz.x;
                float nmn, nmx, nsz;
                if (split)
                {
                    // the lower half, pushed for later: rows sliced to [0, 1/2],
                    // stored transposed (the popped entry cuts the other side:
                    // stored word 4c + r holds row r, point c)
                    {
                        // subpatchU2(copy, 0, 1/2) one row at a time
                        const float a = 0.0f, b = 0.5f;
                        const float iu = 1.0f - a, vv = (b - a) / iu, ivv = 1.0f - vv;
#pragma unroll
                        for (int r = 0; r < 4; r++)
                        {
                            V2 c0 = p[4 * r], c1 = p[4 * r + 1], c2 = p[4 * r + 2], c3 = p[4 * r + 3];
                            slice1(c0.x, c1.x, c2.x, c3.x, a, iu, vv, ivv);
                            slice1(c0.y, c1.y, c2.y, c3.y, a, iu, vv, ivv);
                            stk.put(sp, r, make_float2(c0.x, c0.y));
                            stk.put(sp, 4 + r, make_float2(c1.x, c1.y));
                            stk.put(sp, 8 + r, make_float2(c2.x, c2.y));
                            stk.put(sp, 12 + r, make_float2(c3.x, c3.y));
                        }
                    }
                    const float hs = 0.5f * ssz;
                    const float hm = smn + hs;
                    const V2 emx = vside ? V2 {mx.x, hm} : V2 {hm, mx.y};
                    const V2 psz = vside ? V2 {sz.x, hs} : V2 {hs, sz.y};
                    stk.put(sp, 16, make_float2(mn.x, mn.y));
                    stk.put(sp, 17, make_float2(emx.x, emx.y));
                    stk.put(sp, 18, make_float2(psz.x, psz.y));
                    stk.put(sp, 19, make_float2(vside ? 0.0f : 1.0f, 0.0f));
                    sp++;
                    nmn = hm;
                    nmx = smx;
                    nsz = hs;
                }
                else
                {
                    nmn = smn + ssz * lower;
                    nmx = smx - ssz * (1.0f - upper);
                    nsz = nmx - nmn;
                }
                subpatchU2(p, split ? 0.5f : lower, split ? 1.0f : upper);
                if (vside)
                {
                    mn.y = nmn;
                    mx.y = nmx;
                    sz.y = nsz;
                }
                else
                {
                    mn.x = nmn;
                    mx.x = nmx;
                    sz.x = nsz;
                }
                // the other side is cut next: transpose
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int c = r + 1; c < 4; c++)
                    {
                        const V2 t = p[4 * r + c];
                        p[4 * r + c] = p[4 * c + r];
                        p[4 * c + r] = t;
                    }
                vside = !vside;
                continue;
            }
        }
        if (pop)
        {
            if (sp == 0) break;
            sp--;
#pragma unroll
            for (int i = 0; i < 16; i++)
            {
                float2 q = stk.get(sp, i);
                p[i] = {q.x, q.y};
            }
            float2 r0 = stk.get(sp, 16), r1 = stk.get(sp, 17), r2 = stk.get(sp, 18), r3 = stk.get(sp, 19);
            mn = {r0.x, r0.y};
            mx = {r1.x, r1.y};
            sz = {r2.x, r2.y};
            vside = r3.x != 0.0f;
        }
    }
    return found;
}

// Newton polish of a closest hit (an opt-in refinement with no counterpart
// in the reference).  The clipper stops when the projected sub-patch is
// below an absolute 2^-12, which leaves (u, v) uncertain by about
// 2^-12 / |dS/du|; and in FP32 an object-local ray far from a small object
// (its origin thousands of local units away) is itself uncertain by an ulp
// of that distance.  So the polish runs in FP64 (the GPU's FP64 vector rate
// is that of its unpacked FP32): the local ray o_l + t M3 d is rebuilt in
// FP64 from the FP32 world ray and world→local matrix, with t the world ray
// parameter, and up to six Newton steps on S(u,v) - o_l - t M3 d = 0
// (Cramer's rule for the 3x3 step, the quotient rule for w) refine (u, v, t).
// The result is kept when the steps converge (the last below 1e-9 in u and
// v) inside the parameter square with t > 0; otherwise the hit stands.
struct D3 { double x, y, z; };
CGS_DEV D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CGS_DEV D3 dmul(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
CGS_DEV double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CGS_DEV D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// x.x*a + x.y*b + x.z*c + x.w*d of the points' four components, in FP64
CGS_DEV void dcomb4(const double* a, const double* b, const double* c, const double* d, const double* x, double* r)
{
#pragma unroll
    for (int k = 0; k < 4; k++) r[k] = ((x[0] * a[k] + x[1] * b[k]) + x[2] * c[k]) + x[3] * d[k];
}
CGS_DEV void dbasis(double t, double* b, double* db)
{
    const double t2 = t * t, t3 = t2 * t;
    // the cubic Bernstein polynomials and their derivatives in power form
    b[0] = ((1.0 - 3.0 * t) + 3.0 * t2) - t3;
    b[1] = (3.0 * t - 6.0 * t2) + 3.0 * t3;
    b[2] = 3.0 * t2 - 3.0 * t3;
    b[3] = t3;
    db[0] = (-3.0 + 6.0 * t) - 3.0 * t2;
    db[1] = (3.0 - 12.0 * t) + 9.0 * t2;
    db[2] = 6.0 * t - 9.0 * t2;
    db[3] = 3.0 * t2;
}

CGS_DEV bool polishHit64(const float4* __restrict__ S, const float4 (&w2l)[4], V3 ow, V3 dw, float& t, float& u, float& v)
{
    // local origin (project of M (o, 1)) and the unnormalised local direction
    double M[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
    {
        M[c][0] = w2l[c].x; M[c][1] = w2l[c].y; M[c][2] = w2l[c].z; M[c][3] = w2l[c].w;
    }
    const double ox = ow.x, oy = ow.y, oz = ow.z, dx = dw.x, dy = dw.y, dz = dw.z;
    double O4[4];
#pragma unroll
    for (int k = 0; k < 4; k++) O4[k] = ((ox * M[0][k] + oy * M[1][k]) + oz * M[2][k]) + M[3][k];
    const D3 o {O4[0] / O4[3], O4[1] / O4[3], O4[2] / O4[3]};
    const D3 D {(dx * M[0][0] + dy * M[1][0]) + dz * M[2][0], (dx * M[0][1] + dy * M[1][1]) + dz * M[2][1],
                (dx * M[0][2] + dy * M[1][2]) + dz * M[2][2]};
    double pu = u, pv = v, pt = t;
    bool converged = false;
    for (int it = 0; it < 6; it++)
    {
        double bu[4], dbu[4], bv[4], dbv[4];
        dbasis(pu, bu, dbu);
        dbasis(pv, bv, dbv);
        // row j of the control net, combined along u (read per step rather
        // than held: 64 doubles of registers)
        double a[4][4], ad[4][4];
#pragma unroll
        for (int j = 0; j < 4; j++)
        {
            double q[4][4];
#pragma unroll
            for (int i = 0; i < 4; i++)
            {
                const float4 c = S[4 * j + i];
                q[i][0] = c.x; q[i][1] = c.y; q[i][2] = c.z; q[i][3] = c.w;
            }
            dcomb4(q[0], q[1], q[2], q[3], bu, a[j]);
            dcomb4(q[0], q[1], q[2], q[3], dbu, ad[j]);
        }
        double X[4], Xu[4], Xv[4];
        dcomb4(a[0], a[1], a[2], a[3], bv, X);
        dcomb4(ad[0], ad[1], ad[2], ad[3], bv, Xu);
        dcomb4(a[0], a[1], a[2], a[3], dbv, Xv);
        const double iw = 1.0 / X[3];
        const D3 p {X[0] * iw, X[1] * iw, X[2] * iw};
        const D3 Pu {(Xu[0] - p.x * Xu[3]) * iw, (Xu[1] - p.y * Xu[3]) * iw, (Xu[2] - p.z * Xu[3]) * iw};
        const D3 Pv {(Xv[0] - p.x * Xv[3]) * iw, (Xv[1] - p.y * Xv[3]) * iw, (Xv[2] - p.z * Xv[3]) * iw};
        const D3 F = dsub(dsub(p, o), dmul(pt, D));
        const D3 mD {-D.x, -D.y, -D.z};
        const double det = ddot(Pu, dcross(Pv, mD));
        if (!(det > 1e-300 || det < -1e-300)) return false;
        const double idet = 1.0 / det;
        const D3 nF {-F.x, -F.y, -F.z};
        const double du = ddot(nF, dcross(Pv, mD)) * idet;
        const double dv = ddot(Pu, dcross(nF, mD)) * idet;
        const double dt = ddot(Pu, dcross(Pv, nF)) * idet;
        pu += du;
        pv += dv;
        pt += dt;
        if ((du < 1e-9 && du > -1e-9) && (dv < 1e-9 && dv > -1e-9))
        {
            converged = true;
            break;
        }
    }
    if (!converged || !(pu >= 0.0 && pu <= 1.0 && pv >= 0.0 && pv <= 1.0 && pt > 0.0)) return false;
    t = (float)pt;
    u = (float)pu;
    v = (float)pv;
    return true;
}

// The same with the mode chosen per lane at run time.
CGS_DEV bool clipPatchRT(const bool any, const float4* __restrict__ S, const ClipRay& R, float tol, float& best,
                         float& bu, float& bv, const ClipStack& stk, bool& overflow)
{
    return any ? clipPatch<true>(S, R, tol, best, bu, bv, stk, overflow)
               : clipPatch<false>(S, R, tol, best, bu, bv, stk, overflow);
}

// ---------------------------------------------------------------------------
// BVH traversal (rt/BVH.h:147-350)
// ---------------------------------------------------------------------------

// Exact 32-bit remainder by an invariant divisor (Lemire, Kaser & Kurz,
// "Faster remainder by direct computation", 2019): M = floor((2^64-1)/d) + 1,
// x mod d = hi64(lo64(M·x) · d).  Exact for every 32-bit x and d >= 1.
struct FastMod
{
    uint64_t M;
    uint32_t d;
};
CGS_DEV uint32_t fastmod(uint32_t x, const FastMod& f)
{
    uint64_t low = f.M * (uint64_t)x;
    return (uint32_t)__umul64hi(low, (uint64_t)f.d);
}

// One BVH with its perfect hash, as laid out in device memory.  Node and
// table entries are absolute offsets into the scene's node array.
struct BVHRef
{
    const float4* __restrict__ nodes;    // 4 float4 (64 B) per node
    const uint32_t* __restrict__ indices;
    const uint32_t* __restrict__ table;
    const uint32_t* __restrict__ displ;
    uint32_t root;                       // absolute node offset of the root
    uint32_t A, B, mask;
    FastMod T;
    uint32_t nodeCount;                  // size of `nodes`, bounds table lookups
    uint32_t finiteBoxes;                // 1: every child box of an inner node is finite
    uint32_t wide;                       // 1: the wide hash family (exact arithmetic)
};

struct NodeBoxes
{
    V3 lmin, lmax, rmin, rmax;
    uint32_t left, right;
};
CGS_DEV NodeBoxes loadNode(const float4* __restrict__ nodes, uint32_t i)
{
    // three 16-B loads for the boxes and one 8-B load for the child offsets
    // (the node's last 8 bytes are padding).  The offsets' element index goes
    // through an empty asm so the compiler cannot see that the address is
    // 16-B aligned: otherwise it widens the 8-B load to 16 B, and the widened
    // load's registers overlap the boxes', which serialises the node fetch
    // into two memory round trips (s_waitcnt vmcnt(0) between the loads).
    const float4 a = nodes[4 * (size_t)i], b = nodes[4 * (size_t)i + 1], c = nodes[4 * (size_t)i + 2];
    size_t w = 8 * (size_t)i + 6;
    asm("" : "+v"(w));
    const uint2 d = reinterpret_cast<const uint2*>(nodes)[w];
    NodeBoxes n;
    n.lmin = {a.x, a.y, a.z};
    n.lmax = {a.w, b.x, b.y};
    n.rmin = {b.z, b.w, c.x};
    n.rmax = {c.y, c.z, c.w};
    n.left = d.x;
    n.right = d.y;
    return n;
}

// loadNode for a BVH whose nodes lie in the first 4 GiB of the node array
// (the top-level BVH, placed first): 32-bit byte offsets from the array's
// wave-uniform base, so each load takes the base in scalar registers and the
// address costs one shift instead of 64-bit arithmetic.  The child offsets
// go through an empty asm for the reason given in loadNode.
CGS_DEV NodeBoxes loadNodeLow(const float4* __restrict__ nodes, uint32_t i)
{
    const char* base = reinterpret_cast<const char*>(nodes);
    const uint32_t off = i * 64u;
    const float4 a = *reinterpret_cast<const float4*>(base + off);
    const float4 b = *reinterpret_cast<const float4*>(base + off + 16);
    const float4 c = *reinterpret_cast<const float4*>(base + off + 32);
    uint32_t offd = off + 48u;
    asm("" : "+v"(offd));
    const uint2 d = *reinterpret_cast<const uint2*>(base + offd);
    NodeBoxes n;
    n.lmin = {a.x, a.y, a.z};
    n.lmax = {a.w, b.x, b.y};
    n.rmin = {b.z, b.w, c.x};
    n.rmax = {c.y, c.z, c.w};
    n.left = d.x;
    n.right = d.y;
    return n;
}

// _node_intersect (rt/BVH.h:147-176).  The reference's "if (tn > tmin) tmin =
// tn" and "if (tf < tmax) tmax = tf" are computed as maxNum/minNum: a NaN
// operand (0·inf) leaves the bound unchanged in both forms, and a ±0
// difference only reaches comparisons (tmin < tmax, L0 < limit, L0 < R0),
// where -0 == +0; tMin is used by nothing but those comparisons.
//
// OCT >= 0 states the sign of every component of Dinv for the whole wave
// (bit i set: component i negative), so the near and far planes are chosen
// at compile time; OCT < 0 reads the signs per lane.  Both compute the same
// values.  OCT >= 0 also drops the -inf / +inf seeds of tmin and tmax:
// maxNum/minNum over the three axes equal the seeded chains unless all
// three distances are NaN, and waveOctant picks OCT >= 0 only for a wave
// whose every ray has an axis with a finite origin and a finite, non-zero
// Dinv against a BVH of finite boxes, where that axis's distances are never
// NaN.
template<int OCT = -1>
CGS_DEV bool nodeIntersect(float& tMin, V3 bmin, V3 bmax, V3 o, V3 Dinv)
{
    const float b0[3] = {bmin.x, bmin.y, bmin.z}, b1[3] = {bmax.x, bmax.y, bmax.z};
    const float oo[3] = {o.x, o.y, o.z}, di[3] = {Dinv.x, Dinv.y, Dinv.z};
    float tn[3], tf[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        const bool s = OCT >= 0 ? ((OCT >> i) & 1) != 0 : __builtin_signbit(di[i]);
        float ta = (b0[i] - oo[i]) * di[i];
        float tb = (b1[i] - oo[i]) * di[i];
        tn[i] = s ? tb : ta;
        tf[i] = s ? ta : tb;
    }
    float tmin, tmax;
    if (OCT >= 0)
    {
        tmin = __builtin_fmaxf(__builtin_fmaxf(tn[0], tn[1]), tn[2]);
        tmax = __builtin_fminf(__builtin_fminf(tf[0], tf[1]), tf[2]);
    }
    else
    {
        tmin = __builtin_fmaxf(__builtin_fmaxf(__builtin_fmaxf(-__builtin_inff(), tn[0]), tn[1]), tn[2]);
        tmax = __builtin_fminf(__builtin_fminf(__builtin_fminf(__builtin_inff(), tf[0]), tf[1]), tf[2]);
    }
    tMin = tmin;
    return tmin < tmax && tmax > 0.0001f;
}

// Upper levels of a BVH staged in LDS, indexed by heap key: node `key` is
// lds[4 * key .. 4 * key + 3] for key < keys (keys = 0: nothing staged).
// Reading a node from here instead of HBM/L2 and skipping the hash lookup on
// an ascent to such a key changes neither the visit order nor any result.
typedef __attribute__((address_space(3))) const float lds_float;

struct NodeCache
{
    lds_float* lds;     // 16 floats per key
    uint32_t keys;
};

CGS_DEV NodeBoxes loadNodeLds(lds_float* lds, uint32_t k)
{
    lds_float* q = lds + 16 * k;
    NodeBoxes n;
    n.lmin = {q[0], q[1], q[2]};
    n.lmax = {q[3], q[4], q[5]};
    n.rmin = {q[6], q[7], q[8]};
    n.rmax = {q[9], q[10], q[11]};
    n.left = __float_as_uint(q[12]);
    n.right = __float_as_uint(q[13]);
    return n;
}

template<bool LOW = false>
CGS_DEV NodeBoxes loadNodeCached(const BVHRef& b, const NodeCache& nc, uint32_t node, uint32_t key)
{
    if (key < nc.keys) return loadNodeLds(nc.lds, key);
    return LOW ? loadNodeLow(b.nodes, node) : loadNode(b.nodes, node);
}

// PerfectHashFunction::hash (PerfectHashFunction.h:58-63): 32-bit
// arithmetic, or for a BVH hashed with the wide family the same function in
// exact arithmetic, (A·(k mod T) + B + D) mod T.
template<bool BIG = true>
CGS_DEV uint32_t hashKey(const BVHRef& b, uint32_t key)
{
    if (BIG && b.wide)
    {
        // x mod T through a double quotient: x < 2^63 and T < 2^31, so the
        // quotient is off by at most one and one correction each way is exact
        const uint64_t x = (uint64_t)b.A * fastmod(key, b.T) + b.B + b.displ[key & b.mask];
        const double q = __builtin_floor((double)x / (double)b.T.d);
        int64_t r = (int64_t)x - (int64_t)q * (int64_t)b.T.d;
        r += r < 0 ? (int64_t)b.T.d : 0;
        r -= r >= (int64_t)b.T.d ? (int64_t)b.T.d : 0;
        return (uint32_t)r;
    }
    uint32_t x = b.A * key + b.B + b.displ[key & b.mask];
    return fastmod(x, b.T);
}

// Node indices of a traversal's postponed siblings, kept per lane in LDS so
// that an ascent usually reads the sibling's index instead of hashing its
// key (a displacement load and a table load, dependent, before the node
// load).  The ring holds the most recent `cap` postponed siblings in order:
// a descent that postpones one writes it (overwriting the oldest when full),
// an ascent takes the newest.  An ascent that finds the ring empty while
// siblings are still postponed looks the key up in the perfect hash as
// before.  The visited nodes and their order are unchanged.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
constexpr uint32_t kRingStride = 256;   // threads per workgroup: slot k of a lane is base[k * 256]

struct SiblingRing
{
    lds_u32* base;   // this lane's slot 0
    uint32_t cap;    // 0: no ring, every ascent hashes (the same for all lanes)
};

struct RingState
{
    uint32_t next = 0, live = 0;

    CGS_DEV void push(const SiblingRing& r, uint32_t node)
    {
        if (!r.cap) return;
        r.base[next * kRingStride] = node;
        next = next + 1 == r.cap ? 0 : next + 1;
        live += live < r.cap ? 1u : 0u;
    }
    CGS_DEV bool pop(const SiblingRing& r, uint32_t& node)
    {
        if (!live) return false;
        next = next ? next - 1 : r.cap - 1;
        node = r.base[next * kRingStride];
        --live;
        return true;
    }
};

// One lane's place in a hashed stackless traversal (rt/BVH.h:178-350): the
// current node and its heap key, a bit per level whose far child is still
// postponed, and the sibling ring.
struct TravState
{
    uint32_t node, key, postponed;
    RingState rs;
};

// Moves to the most recently postponed sibling, after a leaf or a subtree the
// ray misses.  Returns false at the end of the traversal (nothing postponed),
// or on a table entry past the node array (a corrupt table ends the ray
// instead of faulting).
template<bool BIG = true>
CGS_DEV bool ascend(const BVHRef& b, const NodeCache& nc, const SiblingRing& ring, TravState& st)
{
    if (!st.postponed) return false;
    const int levels = __builtin_ctz(st.postponed);
    st.postponed >>= levels;
    st.key = (st.key >> levels) ^ 1u;
    st.postponed ^= 1;
    if (st.rs.pop(ring, st.node)) return true;
    if (st.key < nc.keys) return true;    // staged in LDS: read by key, no lookup
    CGS_COUNT(kWorkProbe);
    st.node = b.table[hashKey<BIG>(b, st.key)];
    return st.node < b.nodeCount;
}

// Walks interior nodes (descending into the nearer child hit, postponing the
// other when both are hit, right on a tie; ascending when neither is) until
// it reaches a leaf, whose first element index it returns, or the end
// (EMPTY).  The caller tests the leaf's elements outside this loop and then
// ascends: the per-lane order of node visits and leaf tests is the
// reference's, while the loop itself stays small enough to keep its
// registers, and lanes that reach leaves together test them together.
template<int OCT = -1, bool LOW = false, bool BIG = true>
CGS_DEV uint32_t nextLeaf(const BVHRef& b, const NodeCache& nc, const SiblingRing& ring, V3 o, V3 Dinv,
                          const float& limit, TravState& st)
{
    while (true)
    {
        const NodeBoxes n = loadNodeCached<LOW>(b, nc, st.node, st.key);
        if (n.left == EMPTY) return n.right;
        CGS_COUNT(kWorkNode);
        float L0, R0;
        const bool iL = nodeIntersect<OCT>(L0, n.lmin, n.lmax, o, Dinv) && L0 < limit;
        const bool iR = nodeIntersect<OCT>(R0, n.rmin, n.rmax, o, Dinv) && R0 < limit;
        if (iL || iR)
        {
            const bool both = iL && iR;
            const bool goLeft = both ? L0 < R0 : iL;
            if (both) st.rs.push(ring, goLeft ? n.right : n.left);
            st.postponed = (st.postponed << 1) | (both ? 1u : 0u);
            st.key = (st.key << 1) | (goLeft ? 0u : 1u);
            st.node = goLeft ? n.left : n.right;
            continue;
        }
        if (!ascend<BIG>(b, nc, ring, st)) return EMPTY;
    }
}

// The direction octant shared by every active lane of the wave (bit i: Dinv
// component i negative, by sign bit as nodeIntersect reads it), or -1 when
// the lanes differ, when some lane's ray has no axis with a finite origin
// and a finite, non-zero Dinv, or when the BVH has a non-finite box (the
// conditions under which nodeIntersect<OCT >= 0> equals the generic form).
// Rays of one 8×8 pixel tile nearly always share an octant.
CGS_DEV int waveOctant(V3 o, V3 Dinv, uint32_t finiteBoxes)
{
    const int oct = (__builtin_signbit(Dinv.x) ? 1 : 0) | (__builtin_signbit(Dinv.y) ? 2 : 0) |
                    (__builtin_signbit(Dinv.z) ? 4 : 0);
    auto axis = [](float oi, float di) {
        return __builtin_isfinite(oi) && __builtin_isfinite(di) && di != 0.0f;
    };
    const bool ok = axis(o.x, Dinv.x) || axis(o.y, Dinv.y) || axis(o.z, Dinv.z);
    const int first = __builtin_amdgcn_readfirstlane(oct);
    return (__ballot(oct != first || !ok) == 0 && finiteBoxes) ? first : -1;
}

// nextLeaf with the near/far planes fixed at compile time when the wave
// shares a direction octant (a wave-uniform branch), per lane otherwise.
template<bool LOW, bool BIG = true>
CGS_DEV uint32_t nextLeafOct(int oct, const BVHRef& b, const NodeCache& nc, const SiblingRing& ring, V3 o, V3 Dinv,
                             const float& limit, TravState& st)
{
    switch (oct)
    {
    case 0: return nextLeaf<0, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 1: return nextLeaf<1, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 2: return nextLeaf<2, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 3: return nextLeaf<3, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 4: return nextLeaf<4, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 5: return nextLeaf<5, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 6: return nextLeaf<6, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    case 7: return nextLeaf<7, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    default: return nextLeaf<-1, LOW, BIG>(b, nc, ring, o, Dinv, limit, st);
    }
}

// Stackless closest-hit traversal (rt/BVH.h:178-266).  `limit` is the
// caller's hit.t (read through a reference so leaf hits tighten the culling);
// `leaf(index)` tests one element and returns whether it improved the hit.
// OCTANTS: specialise the node loop on the wave's direction octant (worth its
// code size where a wave's rays are coherent: the top level, in world space).
// LOW: the BVH's nodes lie in the first 4 GiB of the node array (loadNodeLow).
template<bool OCTANTS = false, bool LOW = false, bool BIG = true, class Leaf>
CGS_DEV bool hashIntersect(const BVHRef& b, V3 o, V3 dir, const float& limit, Leaf&& leaf,
                           const NodeCache& nc = NodeCache {nullptr, 0}, const SiblingRing& ring = {nullptr, 0})
{
    const V3 Dinv {1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
    TravState st {b.root, 1u, 0u, RingState {}};
    const int oct = OCTANTS ? waveOctant(o, Dinv, b.finiteBoxes) : -1;
    bool result = false;
    uint32_t i;
    while ((i = (OCTANTS ? nextLeafOct<LOW, BIG>(oct, b, nc, ring, o, Dinv, limit, st)
                         : nextLeaf<-1, LOW, BIG>(b, nc, ring, o, Dinv, limit, st))) != EMPTY)
    {
        uint32_t e;
        while ((e = b.indices[i]) != EMPTY)
        {
            CGS_COUNT(kWorkLeaf);
            result |= leaf(e);
            i++;
        }
        if (!ascend<BIG>(b, nc, ring, st)) break;
    }
    return result;
}

// Stackless any-hit traversal against `tmax` (rt/BVH.h:268-350).
template<bool OCTANTS = false, bool LOW = false, bool BIG = true, class Leaf>
CGS_DEV bool hashIntersectAny(const BVHRef& b, V3 o, V3 dir, float tmax, Leaf&& leaf,
                              const NodeCache& nc = NodeCache {nullptr, 0}, const SiblingRing& ring = {nullptr, 0})
{
    const V3 Dinv {1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
    TravState st {b.root, 1u, 0u, RingState {}};
    const int oct = OCTANTS ? waveOctant(o, Dinv, b.finiteBoxes) : -1;
    uint32_t i;
    while ((i = (OCTANTS ? nextLeafOct<LOW, BIG>(oct, b, nc, ring, o, Dinv, tmax, st)
                         : nextLeaf<-1, LOW, BIG>(b, nc, ring, o, Dinv, tmax, st))) != EMPTY)
    {
        uint32_t e;
        while ((e = b.indices[i]) != EMPTY)
        {
            CGS_COUNT(kWorkLeaf);
            if (leaf(e)) return true;
            i++;
        }
        if (!ascend<BIG>(b, nc, ring, st)) break;
    }
    return false;
}

} // namespace cgs
