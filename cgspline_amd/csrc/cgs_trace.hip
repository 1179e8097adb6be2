// SPDX-License-Identifier: MIT
//
// Ray-query and render kernels for gfx950.
//
// One ray per lane; 256-lane workgroups (4 waves).  A lane runs the reference's
// per-ray algorithm — top-level traversal, per-object local ray, per-object
// hashed traversal, Bézier clipping at each leaf — so every result equals the
// CPU restatement's bit for bit.  The objects a ray reaches are found through
// a cull tree where the scene has one, with the reference's top-level hashed
// traversal for the rays whose answer could depend on the visit order.
//
#include <hip/hip_runtime.h>

#include "cgs_internal.hpp"
#include "cgs_scene_dev.hpp"
#include "cgs_sample.hpp"

namespace cgs
{

constexpr float kClipTol = 0x1p-12f; // Spline.h:493

// Leaf method of the ray queries.  Clip: Bézier clipping (the reference's
// leaf call, Primitive.cu:234,260); ClipCapped: the same stopped after a
// number of loop passes; Subdiv: subdivision (the commented-out alternative
// at Primitive.cu:236,263, Spline_Subdivision.cu:168-288).  The last two,
// and the Newton polish, are opt-in modes; the default is Clip, unpolished.
enum LeafMethod : int
{
    Clip = 0,
    ClipCapped = 1,
    Subdiv = 2,
};

struct TraceParams
{
    float tol;          // clipping tolerance (Spline.h:493 default 2^-12)
    float threshold;    // subdivision threshold (Spline.h:500 default 2^-4)
    uint32_t maxIters;  // ClipCapped: loop passes per patch test
};

struct LocalRay
{
    V3 o, d;
    float max, dinv;
};

// RayTracer.cu:305-314,333: world ray -> object-local ray; t scales by |M d|.
CGS_DEV LocalRay toLocal(const DevObject& ob, V3 o, V3 d, float tmax)
{
    M4 M;
    M.c[0] = ob.w2l[0]; M.c[1] = ob.w2l[1]; M.c[2] = ob.w2l[2]; M.c[3] = ob.w2l[3];
    LocalRay l;
    l.o = project4(m4mul(M, make_float4(o.x, o.y, o.z, 1.0f)));
    V3 ld = m3mul(M, d);
    float len = len3(ld);
    l.max = tmax * len;
    float dinv = 1.0f / len;
    l.d = mul3(dinv, ld);
    l.dinv = dinv;
    return l;
}

// ---------------------------------------------------------------------------
// Subdivision intersection, host semantics (Spline_Subdivision.cu:168-288)
// ---------------------------------------------------------------------------

// Slab test over [0, +inf) (Ds Bounds3::intersect as restated).
CGS_DEV bool boxHit(V3 bmin, V3 bmax, V3 o, V3 d)
{
    float tmin = 0.0f, tmax = __builtin_inff();
    const float b0[3] = {bmin.x, bmin.y, bmin.z}, b1[3] = {bmax.x, bmax.y, bmax.z};
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
    {
        float inv = 1.0f / dd[i];
        float t0 = (b0[i] - oo[i]) * inv;
        float t1 = (b1[i] - oo[i]) * inv;
        if (inv < 0.0f)
        {
            float s = t0;
            t0 = t1;
            t1 = s;
        }
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
        if (tmin > tmax) return false;
    }
    return true;
}

// Möller–Trumbore (Ds triangle::intersect as restated): t in (0, +inf).
CGS_DEV bool triHit(V3 o, V3 d, V3 p0, V3 p1, V3 p2, float& t)
{
    V3 e1 = sub3(p1, p0), e2 = sub3(p2, p0);
    V3 s1 = cross3(d, e2);
    float det = dot3(s1, e1);
    if (__builtin_fabsf(det) < 1.17549435e-38f) return false;
    float inv = 1.0f / det;
    V3 s = sub3(o, p0);
    float b1 = dot3(s, s1) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    V3 s2 = cross3(s, e1);
    float b2 = dot3(d, s2) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    float tt = dot3(e2, s2) * inv;
    if (!(tt > 0.0f)) return false;
    t = tt;
    return true;
}

CGS_DEV void subpatch4(float4 (&p)[16], bool vside, float a, float b)
{
    float iu = 1.0f - a;
    float vv = (b - a) / iu;
    float ivv = 1.0f - vv;
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        const int i0 = vside ? k : 4 * k;
        const int st = vside ? 4 : 1;
        slice1(p[i0].x, p[i0 + st].x, p[i0 + 2 * st].x, p[i0 + 3 * st].x, a, iu, vv, ivv);
        slice1(p[i0].y, p[i0 + st].y, p[i0 + 2 * st].y, p[i0 + 3 * st].y, a, iu, vv, ivv);
        slice1(p[i0].z, p[i0 + st].z, p[i0 + 2 * st].z, p[i0 + 3 * st].z, a, iu, vv, ivv);
        slice1(p[i0].w, p[i0 + st].w, p[i0 + 2 * st].w, p[i0 + 3 * st].w, a, iu, vv, ivv);
    }
}

CGS_DEV void projBox(const float4 (&p)[16], V3& mn, V3& mx)
{
    const float big = 3.40282347e+38f;
    mn = {big, big, big};
    mx = {-big, -big, -big};
#pragma unroll
    for (int i = 0; i < 16; i++)
    {
        V3 q = project4(p[i]);
        if (q.x < mn.x) mn.x = q.x;
        if (q.x > mx.x) mx.x = q.x;
        if (q.y < mn.y) mn.y = q.y;
        if (q.y > mx.y) mx.y = q.y;
        if (q.z < mn.z) mn.z = q.z;
        if (q.z > mx.z) mx.z = q.z;
    }
}

constexpr int kSubdivStack = 64;

// The subdivision test of one patch against a ray (o, d unit), host
// semantics: returns found; best/bu/bv updated by a nearer triangle hit;
// overflow when the region stack is exhausted.
template<bool ANY>
CGS_DEV bool subdividePatch(const float4* __restrict__ S, V3 o, V3 d, float threshold, float& best, float& bu,
                            float& bv, bool& overflow)
{
    bool found = false;
    float4 sub[16];
#pragma unroll
    for (int i = 0; i < 16; i++) sub[i] = S[i];
    V3 bmn, bmx;
    projBox(sub, bmn, bmx);
    if (!boxHit(bmn, bmx, o, d)) return false;
    float4 st[kSubdivStack]; // (umin, vmin, umax, vmax)
    uint8_t side[kSubdivStack];
    int sp = 0;
    st[sp] = {0.0f, 0.0f, 0.5f, 1.0f}; side[sp++] = 0;
    st[sp] = {0.5f, 0.0f, 1.0f, 1.0f}; side[sp++] = 0;
    while (sp > 0)
    {
        sp--;
        const float4 e = st[sp];
        const bool es = side[sp] != 0;
#pragma unroll
        for (int i = 0; i < 16; i++) sub[i] = S[i];
        subpatch4(sub, true, e.y, e.w);
        subpatch4(sub, false, e.x, e.z);
        projBox(sub, bmn, bmx);
        if (!boxHit(bmn, bmx, o, d)) continue;
        float x = 0.5f * (e.x + e.z);
        float y = 0.5f * (e.y + e.w);
        if ((e.z - e.x) < threshold && (e.w - e.y) < threshold)
        {
            V3 p00 = project4(sub[0]), p10 = project4(sub[3]), p01 = project4(sub[12]), p11 = project4(sub[15]);
            float t = 0.0f, t1 = 0.0f;
            bool b0 = triHit(o, d, p10, p01, p00, t);
            bool b1 = triHit(o, d, p01, p10, p11, t1);
            if (b0)
            {
                if (b1 && t1 < t) t = t1;
            }
            else if (b1)
            {
                t = t1;
                b0 = true;
            }
            if (b0 && t < best)
            {
                best = t;
                bu = x;
                bv = y;
                found = true;
                if (ANY) break;
            }
        }
        else
        {
            if (sp + 2 > kSubdivStack)
            {
                overflow = true;
                break;
            }
            if (es)
            {
                st[sp] = {e.x, e.y, x, e.w}; side[sp++] = !es;
                st[sp] = {x, e.y, e.z, e.w}; side[sp++] = !es;
            }
            else
            {
                st[sp] = {e.x, e.y, e.z, y}; side[sp++] = !es;
                st[sp] = {e.x, y, e.z, e.w}; side[sp++] = !es;
            }
        }
    }
    return found;
}

// One patch test of a scene query by the chosen leaf method.
template<int METHOD, bool ANY>
CGS_DEV bool leafTest(const float4* __restrict__ S, const ClipRay& cr, const TraceParams& tp, float& best, float& bu,
                      float& bv, const ClipStack& stk, bool& overflow)
{
    bool f;
    if constexpr (METHOD == Subdiv)
        f = subdividePatch<ANY>(S, cr.o, cr.d, tp.threshold, best, bu, bv, overflow);
    else if constexpr (METHOD == ClipCapped)
        f = clipPatch<ANY, true>(S, cr, tp.tol, best, bu, bv, stk, overflow, tp.maxIters);
    else
        f = clipPatch<ANY>(S, cr, tp.tol, best, bu, bv, stk, overflow);
    return f;
}

struct ClosestHit
{
    float t, u, v;
    uint32_t patch;
    int32_t object;
};

// Sibling rings of a lane for the two traversal levels: the top-level BVH
// over objects (CGS_RING_TOP entries; 4096 objects make it 13 levels deep)
// and the patch BVH of one object (CGS_RING_BOT; 256 patches, 9 levels).
#ifndef CGS_RING_TOP
#define CGS_RING_TOP 8
#endif
#ifndef CGS_RING_BOT
#define CGS_RING_BOT 8
#endif
// Octant-specialised node loops in the patch BVHs too (object-local rays:
// lanes in different objects rarely share an octant).
#ifndef CGS_BOT_OCTANTS
#define CGS_BOT_OCTANTS false
#endif
struct Rings
{
    SiblingRing top, bot;
};
constexpr uint32_t kRingWords = CGS_RING_TOP + CGS_RING_BOT ? (CGS_RING_TOP + CGS_RING_BOT) * kRingStride : 1;

CGS_DEV Rings laneRings(uint32_t* lds)
{
    lds_u32* lane = (lds_u32*)lds + threadIdx.x;
    return Rings {{CGS_RING_TOP ? lane : nullptr, CGS_RING_TOP},
                  {CGS_RING_BOT ? lane + CGS_RING_TOP * kRingStride : nullptr, CGS_RING_BOT}};
}

// ---------------------------------------------------------------------------
// Cull-tree enumeration (cgs_cull.cpp): the elements (objects or patches)
// whose reference leaf box a ray passes, in the tree's near-to-far order
// ---------------------------------------------------------------------------

// The cull tree's node and element records, read per lane.
struct CullIter
{
    uint32_t node, elem, elemEnd;
};

CGS_DEV bool cullable(V3 o, V3 Dinv)
{
    auto ok = [](float oi, float di) { return __builtin_isfinite(oi) && __builtin_isfinite(di) && di != 0.0f; };
    return ok(o.x, Dinv.x) && ok(o.y, Dinv.y) && ok(o.z, Dinv.z);
}

// The next object whose reference leaf box the ray passes with entry
// below tmax and at most lim; its entry in L0.  EMPTY at the end.
CGS_DEV uint32_t cullNext(const SceneDev& sc, int oct, V3 o, V3 Dinv, float lim, float tmax, CullIter& it, float& L0)
{
    while (true)
    {
        while (it.elem < it.elemEnd)
        {
            CGS_COUNT(kWorkCullElem);
            const float4 a = sc.cullElems[2 * (size_t)it.elem], b = sc.cullElems[2 * (size_t)it.elem + 1];
            it.elem++;
            float l;
            if (nodeIntersect<-1>(l, {a.x, a.y, a.z}, {b.x, b.y, b.z}, o, Dinv) && l < tmax && l <= lim)
            {
                L0 = l;
                return __float_as_uint(a.w);
            }
        }
        if (it.node == EMPTY) return EMPTY;
        CGS_COUNT(kWorkCullNode);
        const float4 a = sc.cullNodes[2 * (size_t)it.node], b = sc.cullNodes[2 * (size_t)it.node + 1];
        const uint2 l = sc.cullLinks[8 * (size_t)it.node + oct];
        float tn;
        const bool pass = nodeIntersect<-1>(tn, {a.x, a.y, a.z}, {b.x, b.y, b.z}, o, Dinv) && tn <= lim;
        const uint32_t count = __float_as_uint(b.w);
        if (pass && count)
        {
            it.elem = __float_as_uint(a.w);
            it.elemEnd = it.elem + count;
        }
        it.node = pass && !count ? l.x : l.y;
    }
}

// How far beyond the nearest hit so far an object's box entry may lie and
// the object still be tested: 2^-10 of the distance plus the origin's
// magnitude, orders of magnitude above the FP32 rounding of a hit distance
// or a box entry.
CGS_DEV float cullLimit(float best, float mag) { return best + (best + mag) * 0x1p-10f; }

CGS_DEV int dirOctant(V3 Dinv)
{
    return (__builtin_signbit(Dinv.x) ? 1 : 0) | (__builtin_signbit(Dinv.y) ? 2 : 0) | (__builtin_signbit(Dinv.z) ? 4 : 0);
}

// The patches of a large surface through its cull tree: the argument of the
// objects' enumeration (the comment above closestCulled) one level down.  A
// patch's clip accepts candidates by the running distance only (onHit's
// t < best), so a patch's own nearest hit does not depend on the distance
// its test starts from; every patch whose leaf box the local ray passes near
// or before the nearest hit so far is clipped from the local ray's max, and
// the nearest wins.  Returns false, its outputs untouched, where the
// reference's visit order could decide: a tie, L0_W >= t_W, a region-stack
// overflow, a ray with a zero or non-finite component.  Clipping only (the
// subdivision leaf keeps the reference's traversal).
template<int METHOD>
CGS_DEV bool surfaceClosestCulled(const SceneDev& sc, const DevBVH& surf, const LocalRay& lr, const ClipRay& cr,
                                  const ClipStack& stk, const TraceParams& tp, float& lt, float& lu, float& lv,
                                  uint32_t& lp, bool& hit)
{
    const V3 Dinv {1.0f / lr.d.x, 1.0f / lr.d.y, 1.0f / lr.d.z};
    if (!cullable(lr.o, Dinv)) return false;
    const int oct = dirOctant(Dinv);
    const float mag = fmax_(__builtin_fabsf(lr.o.x), fmax_(__builtin_fabsf(lr.o.y), __builtin_fabsf(lr.o.z)));
    float bt = lr.max, bu = 0.0f, bv = 0.0f, L0W = 0.0f, L0;
    uint32_t bp = EMPTY, pid;
    bool tie = false, overflow = false;
    CullIter it {surf.cullRoot, 0u, 0u};
    while ((pid = cullNext(sc, oct, lr.o, Dinv, cullLimit(bt, mag), lr.max, it, L0)) != EMPTY)
    {
        const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + pid);
        float t = lr.max, u = 0.0f, v = 0.0f;
        bool f;
        if constexpr (METHOD == Clip)
            f = clipPatch<false>(S, cr, kClipTol, t, u, v, stk, overflow);
        else
            f = leafTest<METHOD, false>(S, cr, tp, t, u, v, stk, overflow);
        if (overflow) return false;
        if (!f) continue;
        if (t < bt)
        {
            bt = t;
            bu = u;
            bv = v;
            bp = pid;
            L0W = L0;
            tie = false;
        }
        else if (t == bt)
            tie = true;
    }
    if (tie || (bp != EMPTY && !(L0W < bt))) return false;
    hit = bp != EMPTY;
    if (hit)
    {
        lt = bt;
        lu = bu;
        lv = bv;
        lp = bp;
    }
    return true;
}

// A shadow ray's own surface patch, clipped last.  An any-hit answer is the
// OR of the patch tests the traversal reaches, in whatever order, so a query
// may pass over one patch, note that the traversal reached it (`seen`), and
// clip it only if no other patch is hit (originAny).  A shadow ray that
// grazes its origin patch clips it for up to thousands of passes; one that
// another patch occludes then never does.
struct OriginSkip
{
    int32_t object;
    uint32_t patch;
    bool seen;
};

// rt::intersect(BezierSurface, ray) through the surface's cull tree: any
// patch whose leaf box the ray passes below the local max and that has a
// hit, in any order (the reference's any-hit walk never narrows its limit).
// Returns false where an overflow or the ray's components call for the
// reference's traversal.
template<int METHOD>
CGS_DEV bool surfaceAnyCulled(const SceneDev& sc, const DevBVH& surf, const LocalRay& lr, const ClipRay& cr,
                              const ClipStack& stk, const TraceParams& tp, bool& occluded, uint32_t skip = EMPTY,
                              OriginSkip* origin = nullptr)
{
    const V3 Dinv {1.0f / lr.d.x, 1.0f / lr.d.y, 1.0f / lr.d.z};
    if (!cullable(lr.o, Dinv)) return false;
    const int oct = dirOctant(Dinv);
    float L0;
    uint32_t pid;
    bool overflow = false;
    CullIter it {surf.cullRoot, 0u, 0u};
    while ((pid = cullNext(sc, oct, lr.o, Dinv, lr.max, lr.max, it, L0)) != EMPTY)
    {
        if (pid == skip)
        {
            origin->seen = true;
            continue;
        }
        const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + pid);
        float best = lr.max, bu, bv;
        bool f;
        if constexpr (METHOD == Clip)
            f = clipPatch<true>(S, cr, kClipTol, best, bu, bv, stk, overflow);
        else
            f = leafTest<METHOD, true>(S, cr, tp, best, bu, bv, stk, overflow);
        if (overflow) return false;
        if (f)
        {
            occluded = true;
            return true;
        }
    }
    occluded = false;
    return true;
}

// rt::intersect(BezierSurface, hit, ray) (Primitive.cu:222-248) for object
// i, from RayTracer.cu:305-333: the object-local ray, the per-object hashed
// traversal from t = ray.max·|M d|, and the leaf test at each patch.  On a
// hit, t is the world distance (local t / |M d|).
template<int METHOD, bool BIG = true>
CGS_DEV bool objectClosest(const SceneDev& sc, const Rings& rg, uint32_t i, V3 o, V3 d, float tmax,
                           const ClipStack& stk, bool& overflow, const TraceParams& tp, float& t, float& u, float& v,
                           uint32_t& patch)
{
    const DevObject& ob = sc.objects[i];
    if (ob.surface == EMPTY) return false;   // no geometry (empty box: only a NaN ray gets here)
    CGS_COUNT(kWorkObject);
    LocalRay lr = toLocal(ob, o, d, tmax);
    const DevBVH surf = sc.surfaces[ob.surface];
    const BVHRef b = bvhRef(sc, surf);
    const ClipRay cr = makeClipRay(lr.o, versor3(lr.d));
    float lt = lr.max, lu = 0.0f, lv = 0.0f;
    uint32_t lp = EMPTY;
    bool hit = false;
    if constexpr (BIG && METHOD != Subdiv)
        if (surf.cullRoot != EMPTY && surfaceClosestCulled<METHOD>(sc, surf, lr, cr, stk, tp, lt, lu, lv, lp, hit))
        {
            if (!hit) return false;
            t = lt * lr.dinv;
            u = lu;
            v = lv;
            patch = lp;
            return true;
        }
    hit = hashIntersect<CGS_BOT_OCTANTS, false, BIG>(b, lr.o, lr.d, lt, [&](uint32_t pid) -> bool {
        const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + pid);
        bool f;
        // the default keeps the direct call: routing it through leafTest
        // costs the render and trace kernels hundreds of spilled VGPRs
        if constexpr (METHOD == Clip)
            f = clipPatch<false>(S, cr, kClipTol, lt, lu, lv, stk, overflow);
        else
            f = leafTest<METHOD, false>(S, cr, tp, lt, lu, lv, stk, overflow);
        if (f) lp = pid;
        return f;
    }, NodeCache {nullptr, 0}, rg.bot);
    if (!hit) return false;
    t = lt * lr.dinv;
    u = lu;
    v = lv;
    patch = lp;
    return true;
}

// rt::intersect(BezierSurface, ray) (Primitive.cu:250-268) for object i.
template<int METHOD, bool BIG = true>
CGS_DEV bool objectAny(const SceneDev& sc, const Rings& rg, uint32_t i, V3 o, V3 d, float tmax, const ClipStack& stk,
                       bool& overflow, const TraceParams& tp, OriginSkip* origin = nullptr)
{
    const DevObject& ob = sc.objects[i];
    if (ob.surface == EMPTY) return false;   // no geometry (empty box: only a NaN ray gets here)
    CGS_COUNT(kWorkObject);
    LocalRay lr = toLocal(ob, o, d, tmax);
    const DevBVH surf = sc.surfaces[ob.surface];
    const BVHRef b = bvhRef(sc, surf);
    const ClipRay cr = makeClipRay(lr.o, versor3(lr.d));
    const uint32_t skip = origin && (int32_t)i == origin->object ? origin->patch : EMPTY;
    if constexpr (BIG && METHOD != Subdiv)
    {
        bool occluded;
        if (surf.cullRoot != EMPTY && surfaceAnyCulled<METHOD>(sc, surf, lr, cr, stk, tp, occluded, skip, origin))
            return occluded;
    }
    return hashIntersectAny<CGS_BOT_OCTANTS, false, BIG>(b, lr.o, lr.d, lr.max, [&](uint32_t pid) -> bool {
        if (pid == skip)
        {
            origin->seen = true;
            return false;
        }
        const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + pid);
        float best = lr.max, bu, bv;
        if constexpr (METHOD == Clip)
            return clipPatch<true>(S, cr, kClipTol, best, bu, bv, stk, overflow);
        else
            return leafTest<METHOD, true>(S, cr, tp, best, bu, bv, stk, overflow);
    }, NodeCache {nullptr, 0}, rg.bot);
}

// ---------------------------------------------------------------------------
// Object enumeration through the cull tree (cgs_cull.cpp)
// ---------------------------------------------------------------------------
//
// The reference's top-level traversal reaches object X only if the ray
// passes X's leaf box with entry L0 below the hit distance at that moment
// (rt/BVH.h:147-266), and every ancestor box contains X's box, so it passes
// there whenever it passes X's (the slab test is monotone in the box).
// Each object's own result does not depend on the visit order (its traversal
// starts from ray.max·|M d|, RayTracer.cu:311-314).  So:
//  - any hit: the reference's answer is whether some object whose leaf box
//    the ray passes with L0 < ray.max has a hit, in any order;
//  - closest: enumerating every such object whose L0 is not far beyond the
//    nearest hit found so far (a hit lies inside its object's box, up to
//    rounding, for which the slack below leaves a wide margin) finds the
//    nearest hit W.  The reference reaches W when L0_W < t_W (its limit is
//    never below t_W), and then reports W unless another object has a hit at
//    exactly t_W, where its visit order decides.
// A lane falls back to the reference's traversal for a tie, for L0_W >= t_W,
// for a clip-stack overflow on the way (so that overflows are reported as in
// the reference order), and for rays on which the slab test could meet a
// NaN (a zero or non-finite direction component, a non-finite origin).

template<int METHOD, bool BIG = true>
CGS_DEV ClosestHit closestCulled(const SceneDev& sc, const Rings& rg, V3 o, V3 d, V3 Dinv, float tmax,
                                 const ClipStack& stk, const TraceParams& tp, bool& exact)
{
    ClosestHit h {tmax, 0.0f, 0.0f, EMPTY, -1};
    const int oct = dirOctant(Dinv);
    const float mag = fmax_(__builtin_fabsf(o.x), fmax_(__builtin_fabsf(o.y), __builtin_fabsf(o.z)));
    float L0W = 0.0f, L0;
    bool tie = false, overflow = false;
    CullIter it {sc.top.cullRoot, 0u, 0u};
    uint32_t i;
    while ((i = cullNext(sc, oct, o, Dinv, cullLimit(h.t, mag), tmax, it, L0)) != EMPTY)
    {
        float t, u, v;
        uint32_t p;
        if (!objectClosest<METHOD, BIG>(sc, rg, i, o, d, tmax, stk, overflow, tp, t, u, v, p)) continue;
        if (t < h.t)
        {
            h = {t, u, v, p, (int32_t)i};
            L0W = L0;
            tie = false;
        }
        else if (t == h.t)
            tie = true;
    }
    exact = overflow || tie || (h.object >= 0 && !(L0W < h.t));
    return h;
}

template<int METHOD, bool BIG = true>
CGS_DEV bool anyCulled(const SceneDev& sc, const Rings& rg, V3 o, V3 d, V3 Dinv, float tmax, const ClipStack& stk,
                       const TraceParams& tp, bool& exact, OriginSkip* origin = nullptr)
{
    const int oct = dirOctant(Dinv);
    bool overflow = false;
    CullIter it {sc.top.cullRoot, 0u, 0u};
    float L0;
    uint32_t i;
    while ((i = cullNext(sc, oct, o, Dinv, tmax, tmax, it, L0)) != EMPTY)
    {
        const bool hit = objectAny<METHOD, BIG>(sc, rg, i, o, d, tmax, stk, overflow, tp, origin);
        if (overflow) break;
        if (hit) return true;
    }
    exact = overflow;
    return false;
}

// Context::intersect(hit, ray) (RayTracer.cu:296-350): the nearest object
// hit, through the cull tree where it applies, else (or where its answer
// could depend on the visit order) by the reference's top-level traversal.
template<int METHOD = Clip, bool BIG = true>
CGS_DEV ClosestHit sceneClosest(const SceneDev& sc, const NodeCache& nc, const Rings& rg, V3 o, V3 d, float tmax,
                                const ClipStack& stk, bool& overflow, const TraceParams& tp = {kClipTol, 0x1p-4f, 0})
{
    if (sc.top.cullRoot != EMPTY)
    {
        const V3 Dinv {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        if (cullable(o, Dinv))
        {
            bool exact = false;
            const ClosestHit h = closestCulled<METHOD, BIG>(sc, rg, o, d, Dinv, tmax, stk, tp, exact);
            if (!exact) return h;
        }
    }
    ClosestHit h {tmax, 0.0f, 0.0f, EMPTY, -1};
    const BVHRef top = bvhRef(sc, sc.top);
    hashIntersect<true, true, BIG>(top, o, d, h.t, [&](uint32_t i) -> bool {
        float t, u, v;
        uint32_t p;
        if (objectClosest<METHOD, BIG>(sc, rg, i, o, d, tmax, stk, overflow, tp, t, u, v, p) && t < h.t)
        {
            h = {t, u, v, p, (int32_t)i};
            return true;
        }
        return false;
    }, nc, rg.top);
    return h;
}

// Context::intersect(ray) (RayTracer.cu:352-383), through the cull tree
// where it applies.  With `origin`, that patch is passed over and
// origin->seen tells whether the traversal reached it (OriginSkip).
template<int METHOD = Clip, bool BIG = true>
CGS_DEV bool sceneAny(const SceneDev& sc, const NodeCache& nc, const Rings& rg, V3 o, V3 d, float tmax,
                      const ClipStack& stk, bool& overflow, const TraceParams& tp = {kClipTol, 0x1p-4f, 0},
                      OriginSkip* origin = nullptr)
{
    if (sc.top.cullRoot != EMPTY)
    {
        const V3 Dinv {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        if (cullable(o, Dinv))
        {
            bool exact = false;
            if (anyCulled<METHOD, BIG>(sc, rg, o, d, Dinv, tmax, stk, tp, exact, origin)) return true;
            if (!exact) return false;
            if (origin) origin->seen = false;   // the traversal below decides
        }
    }
    const BVHRef top = bvhRef(sc, sc.top);
    return hashIntersectAny<true, true, BIG>(top, o, d, tmax, [&](uint32_t i) -> bool {
        return objectAny<METHOD, BIG>(sc, rg, i, o, d, tmax, stk, overflow, tp, origin);
    }, nc, rg.top);
}

// The test of a shadow ray's origin patch that sceneAny passed over: object
// i's local ray and patch test, as objectAny's leaf test.
CGS_DEV bool originAny(const SceneDev& sc, const OriginSkip& origin, V3 o, V3 d, float tmax, const ClipStack& stk,
                       bool& overflow)
{
    const DevObject& ob = sc.objects[origin.object];
    const LocalRay lr = toLocal(ob, o, d, tmax);
    const DevBVH surf = sc.surfaces[ob.surface];
    const ClipRay cr = makeClipRay(lr.o, versor3(lr.d));
    const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + origin.patch);
    float best = lr.max, bu, bv;
    return clipPatch<true>(S, cr, kClipTol, best, bu, bv, stk, overflow);
}

// Device totals of the work events (CGS_WORK_COUNTS builds).  The sampled
// render's translation unit keeps a copy of its own, which nothing reads.
#if CGS_SAMPLED_TU
static
#endif
__device__ unsigned long long g_workCounts[kWorkEvents];

// Clears the workgroup's event slots (the caller then passes a barrier
// before any event) / adds them to g_workCounts at the end of a kernel.
CGS_DEV void workBegin()
{
#if CGS_WORK_COUNTS
    if (threadIdx.x < kWorkEvents) workSlots()[threadIdx.x] = 0;
#endif
}
CGS_DEV void workEnd()
{
#if CGS_WORK_COUNTS
    __syncthreads();
    if (threadIdx.x < kWorkEvents) atomicAdd(g_workCounts + threadIdx.x, workSlots()[threadIdx.x]);
#endif
}

#ifndef CGS_TOP_LDS_KEYS
#define CGS_TOP_LDS_KEYS 512
#endif
static_assert(CGS_TOP_LDS_KEYS <= kTopKeyCap, "top-level LDS staging beyond the key map");

// Stages the top-level BVH's nodes with heap key < CGS_TOP_LDS_KEYS (its upper
// levels, visited by nearly every ray) in LDS: 64 B per key.  512 keys (32 KiB
// per workgroup) leave room for three resident workgroups per CU.
CGS_DEV NodeCache stageTopLevel(const SceneDev& sc, float4* lds)
{
    for (uint32_t k = threadIdx.x; k < 4 * CGS_TOP_LDS_KEYS; k += blockDim.x)
    {
        const uint32_t node = sc.topByKey[k >> 2];
        lds[k] = node != EMPTY ? sc.nodes[4 * (size_t)node + (k & 3)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    return NodeCache {(lds_float*)(float*)lds, CGS_TOP_LDS_KEYS};
}

// Sum of a per-lane count over the wave, added with one atomic.
CGS_DEV void addWave(unsigned long long* ctr, unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, v);
}

// Persistent launches: the grid holds exactly the resident waves; each wave
// takes chunks of 64 work items from a counter.  `lane` indexes the region
// stack slice of this lane.
struct Persist
{
    float2* stack;
    uint32_t lanes;
    unsigned int* counter;
};

CGS_DEV uint32_t nextChunk(unsigned int* counter)
{
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(counter, 64u);
    return __shfl(base, 0, 64);
}

#ifndef CGS_RENDER_WAVES
#define CGS_RENDER_WAVES 2
#endif
#ifndef CGS_TRACE_WAVES
#define CGS_TRACE_WAVES 3
#endif

template<int METHOD, bool POLISH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CGS_TRACE_WAVES))) void kTraceClosest(SceneDev sc, Persist ps, const cgs_ray* __restrict__ rays, uint32_t n,
                                                     cgs_hit* __restrict__ hits, int32_t* __restrict__ objects,
                                                     unsigned long long* __restrict__ overflowCount, TraceParams tp)
{
    __shared__ float4 topLds[4 * CGS_TOP_LDS_KEYS];
    __shared__ uint32_t ringLds[kRingWords];
    const NodeCache nc = stageTopLevel(sc, topLds);
    const Rings rg = laneRings(ringLds);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    unsigned long long nOverflow = 0;
    for (uint32_t base = nextChunk(ps.counter); base < n; base = nextChunk(ps.counter))
    {
        const uint32_t r = base + (threadIdx.x & 63);
        if (r >= n) continue;
        bool overflow = false;
        const cgs_ray ray = rays[r];
        ClosestHit h = sceneClosest<METHOD>(sc, nc, rg, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz}, ray.tmax,
                                            stk, overflow, tp);
        if constexpr (POLISH)
            if (h.object >= 0 && !overflow)
            {
                // the reported hit only: the traversal's decisions are the reference's
                const DevObject& ob = sc.objects[h.object];
                const DevBVH surf = sc.surfaces[ob.surface];
                polishHit64(sc.patches + 16 * (size_t)(surf.patchBase + h.patch), ob.w2l, {ray.ox, ray.oy, ray.oz},
                            {ray.dx, ray.dy, ray.dz}, h.t, h.u, h.v);
            }
        cgs_hit out;
        out.t = h.t;
        out.u = h.object < 0 ? 0.0f : h.u;
        out.v = h.object < 0 ? 0.0f : h.v;
        out.patch = overflow ? CGS_PATCH_OVERFLOW : (h.object < 0 ? EMPTY : h.patch);
        hits[r] = out;
        if (objects) objects[r] = h.object;
        nOverflow += overflow;
    }
    addWave(overflowCount, nOverflow);
}

template<int METHOD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CGS_TRACE_WAVES))) void kTraceAny(SceneDev sc, Persist ps, const cgs_ray* __restrict__ rays, uint32_t n,
                                                 uint32_t* __restrict__ occluded,
                                                 unsigned long long* __restrict__ overflowCount, TraceParams tp)
{
    __shared__ float4 topLds[4 * CGS_TOP_LDS_KEYS];
    __shared__ uint32_t ringLds[kRingWords];
    const NodeCache nc = stageTopLevel(sc, topLds);
    const Rings rg = laneRings(ringLds);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    unsigned long long nOverflow = 0;
    for (uint32_t base = nextChunk(ps.counter); base < n; base = nextChunk(ps.counter))
    {
        const uint32_t r = base + (threadIdx.x & 63);
        if (r >= n) continue;
        bool overflow = false;
        const cgs_ray ray = rays[r];
        bool occ = sceneAny<METHOD>(sc, nc, rg, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz}, ray.tmax, stk, overflow,
                                    tp);
        occluded[r] = overflow ? 2u : (occ ? 1u : 0u);
        nOverflow += overflow;
    }
    addWave(overflowCount, nOverflow);
}

// Batched ray queries over many resident scenes in one launch: rays of all
// jobs are numbered consecutively and waves take chunks of `chunk` rays (a
// multiple of 64) from one counter, so small and large jobs share the grid
// without a launch per scene.  Each lane finds its job by binary search over
// the jobs' first ray numbers and runs the same per-ray query as
// kTraceClosest / kTraceAny, with the same leaf methods and polish.
template<int METHOD, bool POLISH, bool BIG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CGS_TRACE_WAVES))) void kTraceBatch(
    const BatchJob* __restrict__ jobs, uint32_t njobs, uint32_t total, uint32_t chunk, Persist ps,
    unsigned long long* __restrict__ overflowCount, TraceParams tp)
{
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    const NodeCache none {nullptr, 0};
    __shared__ uint32_t ringLds[kRingWords];
    const Rings rg = laneRings(ringLds);
    unsigned long long nOverflow = 0;
    while (true)
    {
        uint32_t base = 0;
        if ((threadIdx.x & 63) == 0) base = atomicAdd(ps.counter, chunk);
        base = __shfl(base, 0, 64);
        if (base >= total) break;
        const uint32_t end = base + chunk < total ? base + chunk : total;
        for (uint32_t r0 = base; r0 < end; r0 += 64)
        {
            const uint32_t r = r0 + (threadIdx.x & 63);
            if (r >= end) continue;
            uint32_t lo = 0, hi = njobs;  // last job with rayBase <= r
            while (hi - lo > 1)
            {
                const uint32_t mid = (lo + hi) / 2;
                if (jobs[mid].rayBase <= r) lo = mid;
                else hi = mid;
            }
            const BatchJob& J = jobs[lo];
            const uint32_t k = r - J.rayBase;
            const cgs_ray ray = J.rays[k];
            bool overflow = false;
            if (J.any)
            {
                const bool occ = sceneAny<METHOD, BIG>(J.sc, none, rg, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz},
                                                  ray.tmax, stk, overflow, tp);
                J.occluded[k] = overflow ? 2u : (occ ? 1u : 0u);
            }
            else
            {
                ClosestHit h = sceneClosest<METHOD, BIG>(J.sc, none, rg, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz},
                                                    ray.tmax, stk, overflow, tp);
                if constexpr (POLISH)
                    if (h.object >= 0 && !overflow)
                    {
                        const DevObject& ob = J.sc.objects[h.object];
                        const DevBVH surf = J.sc.surfaces[ob.surface];
                        polishHit64(J.sc.patches + 16 * (size_t)(surf.patchBase + h.patch), ob.w2l,
                                    {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz}, h.t, h.u, h.v);
                    }
                cgs_hit out;
                out.t = h.t;
                out.u = h.object < 0 ? 0.0f : h.u;
                out.v = h.object < 0 ? 0.0f : h.v;
                out.patch = overflow ? CGS_PATCH_OVERFLOW : (h.object < 0 ? EMPTY : h.patch);
                J.hits[k] = out;
                if (J.objects) J.objects[k] = h.object;
            }
            nOverflow += overflow;
        }
    }
    addWave(overflowCount, nOverflow);
}

#if !CGS_SAMPLED_TU
// Batched doBezierClipping: ray r against patch r (Spline_BezierClipping.cu:46-114).
__global__ __launch_bounds__(256) void kClipBatch(Persist ps, const cgs_ray* __restrict__ rays,
                                                  const float4* __restrict__ patches, uint32_t n, float tol,
                                                  cgs_hit* __restrict__ hits)
{
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    for (uint32_t base = nextChunk(ps.counter); base < n; base = nextChunk(ps.counter))
    {
        const uint32_t r = base + (threadIdx.x & 63);
        if (r >= n) continue;
        const cgs_ray ray = rays[r];
        bool overflow = false;
        const ClipRay cr = makeClipRay({ray.ox, ray.oy, ray.oz}, versor3({ray.dx, ray.dy, ray.dz}));
        float t = ray.tmax, u = 0.0f, v = 0.0f;
        bool f = clipPatch<false>(patches + 16 * (size_t)r, cr, tol, t, u, v, stk, overflow);
        hits[r] = {t, u, v, overflow ? CGS_PATCH_OVERFLOW : (f ? 1u : EMPTY)};
    }
}

__global__ __launch_bounds__(256) void kSubdivBatch(const cgs_ray* __restrict__ rays, const float4* __restrict__ patches,
                                                    uint32_t n, float threshold, cgs_hit* __restrict__ hits)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const cgs_ray ray = rays[r];
    const V3 o {ray.ox, ray.oy, ray.oz};
    const V3 d = versor3({ray.dx, ray.dy, ray.dz});
    float best = ray.tmax, bu = 0.0f, bv = 0.0f;
    bool overflow = false;
    const bool found = subdividePatch<false>(patches + 16 * (size_t)r, o, d, threshold, best, bu, bv, overflow);
    hits[r] = {best, bu, bv, overflow ? CGS_PATCH_OVERFLOW : (found ? 1u : EMPTY)};
}

// ---------------------------------------------------------------------------
// Patch evaluation (Spline.h:165-254 de Casteljau, Spline.cu:10-27 normal,
// SplineMat.h:125-187 matrix form) and subpatch (Spline.h:384-396)
// ---------------------------------------------------------------------------

// deCasteljau (Spline.h:168-178) over 4 float4 points
CGS_DEV float4 lerp4(float4 a, float4 b, float t, float u)
{
    return {t * a.x + u * b.x, t * a.y + u * b.y, t * a.z + u * b.z, t * a.w + u * b.w};
}
CGS_DEV float4 deCasteljau4(float4 (&p)[4], float u)
{
    const float t = 1.0f - u;
#pragma unroll
    for (int k = 1; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < 4 - k; ++i) p[i] = lerp4(p[i], p[i + 1], t, u);
    return p[0];
}
// deCasteljauDx (Spline.h:183-195): the hodograph without its factor 3
CGS_DEV float4 deCasteljauDx4(float4 (&p)[4], float u)
{
    const float t = 1.0f - u;
#pragma unroll
    for (int i = 1; i < 4; ++i) p[i - 1] = {p[i].x - p[i - 1].x, p[i].y - p[i - 1].y, p[i].z - p[i - 1].z, p[i].w - p[i - 1].w};
#pragma unroll
    for (int k = 2; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < 4 - k; ++i) p[i] = lerp4(p[i], p[i + 1], t, u);
    return p[0];
}
// interpolate (Spline.h:197-214), derivativeU (:216-234), derivativeV (:236-254)
CGS_DEV float4 interpolateDC(const float4* __restrict__ S, float u, float v)
{
    float4 q[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
        float4 p[4] = {S[4 * j], S[4 * j + 1], S[4 * j + 2], S[4 * j + 3]};
        q[j] = deCasteljau4(p, u);
    }
    return deCasteljau4(q, v);
}
CGS_DEV float4 derivativeU_DC(const float4* __restrict__ S, float u, float v)
{
    float4 q[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
        float4 p[4] = {S[4 * j], S[4 * j + 1], S[4 * j + 2], S[4 * j + 3]};
        q[j] = deCasteljauDx4(p, u);
    }
    return deCasteljau4(q, v);
}
CGS_DEV float4 derivativeV_DC(const float4* __restrict__ S, float u, float v)
{
    float4 p[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        float4 q[4] = {S[i], S[4 + i], S[8 + i], S[12 + i]};
        p[i] = deCasteljauDx4(q, v);
    }
    return deCasteljau4(p, u);
}
// normal (Spline.cu:10-27): a vanishing derivative (a degenerate edge) is
// taken 1e-4 inside the patch instead
CGS_DEV V3 normalDC(const float4* __restrict__ S, float u, float v)
{
    const float e = 0.0001f;
    float4 dU = derivativeU_DC(S, u, v);
    float4 dV = derivativeV_DC(S, u, v);
    if (dU.x == 0.0f && dU.y == 0.0f && dU.z == 0.0f) dU = derivativeU_DC(S, u, __builtin_fabsf(v - e));
    if (dV.x == 0.0f && dV.y == 0.0f && dV.z == 0.0f) dV = derivativeV_DC(S, __builtin_fabsf(u - e), v);
    return cross3({dU.x, dU.y, dU.z}, {dV.x, dV.y, dV.z});
}
// mat::derivativeU (SplineMat.h:142-158) / mat::derivativeV (:160-176)
CGS_DEV float4 derivativeUMat(const float4* __restrict__ S, float u, float v)
{
    float4 x = dbasisMul(evalVecDt(u));
    float4 y = basisMul(evalVec(v));
    float4 a0 = comb4(S[0], S[1], S[2], S[3], x);
    float4 a1 = comb4(S[4], S[5], S[6], S[7], x);
    float4 a2 = comb4(S[8], S[9], S[10], S[11], x);
    float4 a3 = comb4(S[12], S[13], S[14], S[15], x);
    return comb4(a0, a1, a2, a3, y);
}
CGS_DEV float4 derivativeVMat(const float4* __restrict__ S, float u, float v)
{
    float4 xb = basisMul(evalVec(u));
    float4 yd = dbasisMul(evalVecDt(v));
    float4 b0 = comb4(S[0], S[4], S[8], S[12], yd);
    float4 b1 = comb4(S[1], S[5], S[9], S[13], yd);
    float4 b2 = comb4(S[2], S[6], S[10], S[14], yd);
    float4 b3 = comb4(S[3], S[7], S[11], S[15], yd);
    return comb4(b0, b1, b2, b3, xb);
}

// One evaluation per item: patch r at (u, v) = uv[r], by `kind` (CGS_EVAL_*).
__global__ __launch_bounds__(256) void kPatchEval(uint32_t kind, const float4* __restrict__ patches,
                                                  const float2* __restrict__ uv, uint32_t n, float4* __restrict__ out)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float4* S = patches + 16 * (size_t)r;
    const float2 t = uv[r];
    float4 o;
    switch (kind)
    {
    case CGS_EVAL_INTERPOLATE: o = interpolateDC(S, t.x, t.y); break;
    case CGS_EVAL_INTERPOLATE_MAT: o = interpolateMat(S, t.x, t.y); break;
    case CGS_EVAL_DERIVATIVE_U: o = derivativeU_DC(S, t.x, t.y); break;
    case CGS_EVAL_DERIVATIVE_V: o = derivativeV_DC(S, t.x, t.y); break;
    case CGS_EVAL_DERIVATIVE_U_MAT: o = derivativeUMat(S, t.x, t.y); break;
    case CGS_EVAL_DERIVATIVE_V_MAT: o = derivativeVMat(S, t.x, t.y); break;
    case CGS_EVAL_NORMAL:
    {
        const V3 nv = normalDC(S, t.x, t.y);
        o = {nv.x, nv.y, nv.z, 0.0f};
        break;
    }
    default:
    {
        const V3 nv = normalMat(S, t.x, t.y);
        o = {nv.x, nv.y, nv.z, 0.0f};
        break;
    }
    }
    out[r] = o;
}

// subpatch (Spline.h:384-396): subpatchV to [vmin, vmax], then subpatchU to
// [umin, umax]; ranges[r] = (umin, umax, vmin, vmax).
__global__ __launch_bounds__(256) void kSubpatch(const float4* __restrict__ patches, const float4* __restrict__ ranges,
                                                 uint32_t n, float4* __restrict__ out)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float4 p[16];
#pragma unroll
    for (int i = 0; i < 16; i++) p[i] = patches[16 * (size_t)r + i];
    const float4 g = ranges[r];
    subpatch4(p, true, g.z, g.w);
    subpatch4(p, false, g.x, g.y);
#pragma unroll
    for (int i = 0; i < 16; i++) out[16 * (size_t)r + i] = p[i];
}
#endif // !CGS_SAMPLED_TU

// ---------------------------------------------------------------------------
// cg::PatchBVH (PatchBVH.h:13-28, PatchBVH.cpp:84-121): rays against the
// patches of one surface, in the surface's own frame
// ---------------------------------------------------------------------------
//
// The ray is a Ray3f: its direction normalized once.  Closest (intersect(ray,
// hit)): hit.distance starts at ray.tmax and each leaf clips its patch
// against it.  Any (intersect(ray)): a leaf clips with hit.distance = inf
// (PatchBVH.cpp:93-94), so a patch whose box the traversal reaches within
// ray.tmax counts with a hit at any distance in front.
template<bool ANY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CGS_TRACE_WAVES))) void kPick(SceneDev sc, Persist ps, const cgs_ray* __restrict__ rays, uint32_t n,
                                                            cgs_hit* __restrict__ hits, uint32_t* __restrict__ occluded)
{
    __shared__ uint32_t ringLds[kRingWords];
    const Rings rg = laneRings(ringLds);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    const DevBVH surf = sc.surfaces[0];
    const BVHRef b = bvhRef(sc, surf);
    for (uint32_t base = nextChunk(ps.counter); base < n; base = nextChunk(ps.counter))
    {
        const uint32_t r = base + (threadIdx.x & 63);
        if (r >= n) continue;
        const cgs_ray ray = rays[r];
        const V3 o {ray.ox, ray.oy, ray.oz};
        const V3 d = versor3({ray.dx, ray.dy, ray.dz});
        const ClipRay cr = makeClipRay(o, d);
        bool overflow = false;
        if (ANY)
        {
            const bool f = hashIntersectAny<true>(b, o, d, ray.tmax, [&](uint32_t pid) -> bool {
                float best = __builtin_inff(), bu, bv;
                return clipPatch<true>(sc.patches + 16 * (size_t)(surf.patchBase + pid), cr, kClipTol, best, bu, bv, stk,
                                       overflow);
            }, NodeCache {nullptr, 0}, rg.bot);
            occluded[r] = overflow ? 2u : (f ? 1u : 0u);
        }
        else
        {
            float t = ray.tmax, u = 0.0f, v = 0.0f;
            uint32_t patch = EMPTY;
            hashIntersect<true>(b, o, d, t, [&](uint32_t pid) -> bool {
                const bool f = clipPatch<false>(sc.patches + 16 * (size_t)(surf.patchBase + pid), cr, kClipTol, t, u, v,
                                                stk, overflow);
                if (f) patch = pid;
                return f;
            }, NodeCache {nullptr, 0}, rg.bot);
            hits[r] = {t, patch == EMPTY ? 0.0f : u, patch == EMPTY ? 0.0f : v, overflow ? CGS_PATCH_OVERFLOW : patch};
        }
    }
}

// ---------------------------------------------------------------------------
// Renderer (RayTracer.cu:194-544, non-Monte-Carlo path)
// ---------------------------------------------------------------------------

struct CamDev
{
    float4 c[4];          // camera→world columns
    float px, py, pz;     // position
    float tlx, tly;       // top-left pixel centre
    float hdx, hdy;       // half pixel sizes
};

struct RenderDev
{
    float eps;
    int depth;
    uint32_t width, height;
    uint32_t stride;      // frame row pitch in pixels (>= width)
    uint32_t samples;     // camera rays per pixel of the sampled render (0: the plain render)
    uint32_t seed;        // the sampled render's random seed
};

// A render's camera, options and outputs, read by kRender from device memory
// where it needs them (once per tile, or per light for eps) rather than held
// in scalar registers for the whole launch: the traversal loops need those
// registers.
struct RenderArgs
{
    CamDev cam;
    RenderDev op;
    uint32_t* frame;
    float* color;
    unsigned long long* counts;
    const uint32_t* order;   // rank -> tile, or null for row order
    uint32_t* tileTicks;     // per tile: its wave's time in wall-clock ticks, or null
};

// PBR (rt/PBR.h, rt/PBR.cu)
CGS_DEV bool lightVector(float& d, V3& L, V3 P, const DevLight& lt)
{
    const float angle = lt.c.w;
    if (angle < 0.01f)
    {
        L = {-lt.c.x, -lt.c.y, -lt.c.z};
        d = __builtin_inff();
        return true;
    }
    L = sub3({lt.b.x, lt.b.y, lt.b.z}, P);
    d = len3(L);
    if (d < 1e-14f || (lt.b.w > 0.0f && d > lt.b.w)) return false;
    L = mul3(1.0f / d, L);
    if (angle > 1.57f) return true;
    float DL = dot3({lt.c.x, lt.c.y, lt.c.z}, L);
    return DL < 0.0f && angle > 2.0f * acosf(DL);
}
CGS_DEV V3 lightColor(float d, const DevLight& lt)
{
    const V3 col {lt.a.x, lt.a.y, lt.a.z};
    if (lt.c.w < 0.01f) return col;
    float range = lt.b.w, f;
    if (range == 0.0f)
    {
        f = 1.0f / d;
        f *= f;
    }
    else
    {
        f = d / range;
        f = 1.0f + f * (f - 2.0f);
    }
    return mul3(f, col);
}
CGS_DEV V3 schlick(V3 f0, float dotLH)
{
    float b = 1.0f - dotLH;
    float b2 = b * b;
    float s = b2 * b2 * b;
    return {f0.x + (1.0f - f0.x) * s, f0.y + (1.0f - f0.y) * s, f0.z + (1.0f - f0.z) * s};
}
CGS_DEV float G1(float dotNX, float k) { return dotNX / (dotNX * (1.0f - k) + k); }
CGS_DEV float Gsmith(float dotNL, float dotNV, float r)
{
    const float r1 = r + 1.0f;
    const float k = (r1 * r1) * 0.125f;
    return G1(dotNL, k) * G1(dotNV, k);
}
CGS_DEV float Dggx(float dotNH, float r)
{
    r = fmax_(r, 1e-3f);
    float a = r * r;
    float a2 = a * a;
    float d = dotNH * dotNH * (a2 - 1.0f) + 1.0f;
    return a2 / (0x1.921fb6p+1f * d * d);
}
CGS_DEV float microfacet(float dotNV, float dotNL, float dotNH, float r)
{
    return Gsmith(dotNL, dotNV, r) * Dggx(dotNH, r) / (4.0f * dotNL * dotNV);
}

// Ds-boundary pack_sRGB: sRGB encode, 8 bits per channel, 0xAABBGGRR.
CGS_DEV uint32_t packSRGB(float r, float g, float b)
{
    auto enc = [](float c) -> uint32_t {
        float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
        int q = (int)(e * 255.0f + 0.5f);
        q = q < 0 ? 0 : (q > 255 ? 255 : q);
        return (uint32_t)q;
    };
    return enc(r) | (enc(g) << 8) | (enc(b) << 16) | 0xFF000000u;
}

CGS_DEV float saturate(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// One pixel's random stream in the sampled render: the seed, the pixel's cell
// of the 256×256 grid and the draws taken so far.
struct SampleState
{
    uint32_t seed, stream, draw;
};

// The radiance carried back along one ray: the render kernel's iterative
// trace loop with Context::trace / miss / closestHit (RayTracer.cu:194-294,
// 385-544): closest hit, ambient + one shadow ray per light with the PBR
// terms, then the reflection bounce while depth remains and the attenuation
// exceeds 2^-8.  The first segment ends at rmax, bounces at infinity.  eps is
// read from device memory per use (the traversal loops need the registers a
// kernel argument would hold for the whole launch).  With RECORD, the first
// segment's hit is returned in `first`.
// SAMPLED: the Monte Carlo branches (RayTracer.cu:403-407,473-522): no ambient
// term, and the bounce mirrors about a GGX microfacet normal drawn from `rs`
// (microfacet(roughness², x, y), PBR.h:159-176) with weight mix(diffuse BRDF,
// F(M·V) G(N·R, N·V) (M·V) / (N·V N·M), metalness); the result adds to `acc`,
// the colour carried over from the pixel's earlier samples.
template<bool RECORD, bool BIG = true, bool SAMPLED = false>
CGS_DEV V3 rayRadiance(const SceneDev& sc, const NodeCache& nc, const Rings& rg, V3 ro, V3 rd, float rmax, int depth,
                       const float* eps, const ClipStack& stk, bool& overflow, uint32_t& nShadow,
                       uint32_t& nSecondary, ClosestHit& first, V3 acc = {0.0f, 0.0f, 0.0f},
                       SampleState* rs = nullptr)
{
    V3 att {1.0f, 1.0f, 1.0f};
    bool primary = true;
    while (true)
    {
        if (!primary) nSecondary++;
        ClosestHit h = sceneClosest<Clip, BIG>(sc, nc, rg, ro, rd, rmax, stk, overflow);
        if (RECORD && primary) first = h;
        primary = false;
        if (h.object < 0)
        {
            acc = add3(acc, mulv3({sc.background[0], sc.background[1], sc.background[2]}, att));
            break;
        }
        const DevObject& ob = sc.objects[h.object];
        const float4 m0 = ob.mat0, m1 = ob.mat1;
        const float metal = m0.w, rough = m1.w;
        const V3 spec {m1.x, m1.y, m1.z};
        const float invPi = 0x1.45f306p-2f; // std::numbers::inv_pi_v<float>
        const V3 diffuse = mul3(1.0f - metal, mul3(invPi, {m0.x, m0.y, m0.z}));
        V3 col = SAMPLED ? V3 {0.0f, 0.0f, 0.0f} : mulv3({sc.ambient[0], sc.ambient[1], sc.ambient[2]}, diffuse);
        V3 V = neg3(rd);
        const DevBVH surf = sc.surfaces[ob.surface];
        V3 N = versor3(normalMat(sc.patches + 16 * (size_t)(surf.patchBase + h.patch), h.u, h.v));
        M4 W;
        W.c[0] = ob.w2l[0]; W.c[1] = ob.w2l[1]; W.c[2] = ob.w2l[2]; W.c[3] = ob.w2l[3];
        N = versor3(m3tmul(W, N));
        float dotNV = dot3(N, V);
        if (dotNV < 0.0f)
        {
            N = neg3(N);
            dotNV = -dotNV;
        }
        V3 Pt = add3(ro, mul3(h.t, rd));
        for (uint32_t li = 0; li < sc.nlight; li++)
        {
            const DevLight lt = sc.lights[li];
            float d;
            V3 L;
            if (!lightVector(d, L, Pt, lt)) continue;
            const float dotNL = dot3(N, L);
            if (dotNL < 1e-14f) continue;
            nShadow++;
            // the hit's own patch last (OriginSkip): the answer is the same
            OriginSkip origin {h.object, h.patch, false};
            const V3 so = add3(Pt, mul3(*eps, L));
            if (sceneAny<Clip, BIG>(sc, nc, rg, so, L, d, stk, overflow, {kClipTol, 0x1p-4f, 0}, &origin)) continue;
            if (origin.seen && originAny(sc, origin, so, L, d, stk, overflow)) continue;
            V3 H = versor3(add3(L, V));
            float dotHN = dot3(H, N);
            float dotHL = dot3(H, L);
            V3 I = lightColor(d, lt);
            V3 bs = mul3(microfacet(dotNV, dotNL, dotHN, rough), schlick(spec, dotHL));
            col = add3(col, mulv3(mul3(dotNL, I), add3(diffuse, mul3(metal, bs))));
        }
        acc = add3(acc, mulv3(mul3(0x1.921fb6p+1f, col), att));
        if (depth <= 0) break;
        V3 R, brdf;
        if constexpr (SAMPLED)
        {
            float x, y;
            drawPair(rs->seed, rs->stream, rs->draw, x, y);
            rs->draw += 2;
            const float alpha = rough * rough;
            const float thetaM = atan2Quadrant(alpha * sqrtf(x), sqrtf(1.0f - x));
            const float phiM = 0x1.921fb6p+2f * y;   // 2π y
            const V3 ts = versor3(cross3(N, V));
            const V3 tp = versor3(cross3(N, ts));
            float sinT, cosT, sinP, cosP;
            sinCosAngle(thetaM, sinT, cosT);
            sinCosAngle(phiM, sinP, cosP);
            const V3 M = versor3(add3(add3(mul3(sinT * cosP, ts), mul3(sinT * sinP, tp)), mul3(cosT, N)));
            const float dotMV = dot3(M, V);
            const V3 refl = schlick(spec, dotMV);
            R = sub3(neg3(V), mul3(2.0f * -dotMV, M));
            const float dotMN = dot3(M, N);
            const float dotNR = dot3(N, R);
            if (dotNR < 1e-14f) break;
            const float g = Gsmith(dotNR, dotNV, rough) * dotMV / (dotNV * dotMN);
            brdf = add3(mul3(1.0f - metal, mul3(invPi, {m0.x, m0.y, m0.z})), mul3(metal, mul3(g, refl)));
        }
        else
        {
            V3 refl = schlick(spec, dotNV);
            R = sub3(neg3(V), mul3(2.0f * -dotNV, N));
            float g = Gsmith(dotNV, dotNV, rough);
            float rr = 1.0f - rough;
            brdf = mul3(rr * rr, mul3(dotNV, add3(diffuse, mul3(metal, mul3(g, refl)))));
        }
        V3 a = mulv3(att, brdf);
        if (max3(a) <= 0x1p-8f) break;
        ro = add3(Pt, mul3(*eps, R));
        rd = R;
        rmax = __builtin_inff();
        att = a;
        depth = depth - 1;
    }
    return acc;
}

// Occupancy target of the render megakernel (waves per SIMD): the compiler
// keeps VGPR + AGPR use within 512 / CGS_RENDER_WAVES.  With shadow rays
// clipping their origin patch last, two waves per SIMD (256 registers)
// render the 1M-patch frame in 9.0 ms against 9.4 ms with three (one wave:
// the same as two) and Test152 in 114 ms against 129 ms; the sampled render
// is the same either way.  (Before that change, when waves waited on
// grazing origin clips, three waves measured best.)
// BIG: the scene has a BVH hashed with the wide family or a surface with a
// cull tree (kernels for other scenes leave those paths out and keep their
// registers).
// SAMPLED: the Monte Carlo render (RayTracer.cu:212-246): op.samples camera
// rays per pixel through (i + s/n, j + vdc(s)), one running colour over them,
// scaled by 1/n; the pixel's stream is its cell of a 256×256 grid.
template<bool BIG, bool SAMPLED = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CGS_RENDER_WAVES))) void kRender(SceneDev sc, Persist ps, const RenderArgs* ra)
{
    __shared__ float4 topLds[4 * CGS_TOP_LDS_KEYS];
    __shared__ uint32_t ringLds[kRingWords];
    workBegin();
    const NodeCache nc = stageTopLevel(sc, topLds);
    const Rings rg = laneRings(ringLds);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t nItems;
    {
        const RenderDev op = ra->op;
        nItems = ((op.width + 7) / 8) * ((op.height + 7) / 8) * 64;
    }
    uint32_t nPrimary = 0, nShadow = 0, nSecondary = 0, nOverflow = 0;
    for (uint32_t base = nextChunk(ps.counter); base < nItems; base = nextChunk(ps.counter))
    {
        const RenderDev op = ra->op;
        const CamDev cam = ra->cam;
        const uint32_t tilesX = (op.width + 7) / 8;
        M4 C;
        C.c[0] = cam.c[0]; C.c[1] = cam.c[1]; C.c[2] = cam.c[2]; C.c[3] = cam.c[3];
        const V3 camPos {cam.px, cam.py, cam.pz};
        // one wave = one 8×8 pixel tile, taken in the order of ra->order
        const uint32_t rank = base / 64;
        const uint32_t tile = ra->order ? ra->order[rank] : rank;
        const unsigned long long tileStart = wall_clock64();
        const uint32_t i = (tile % tilesX) * 8 + (lane & 7), j = (tile / tilesX) * 8 + (lane >> 3);
        if (i >= op.width || j >= op.height) continue;
        bool overflow = false;
        ClosestHit unused;
        V3 acc;
        if constexpr (SAMPLED)
        {
            SampleState rs {op.seed, (i & 255u) + 256u * (j & 255u), 0u};
            const uint32_t n = op.samples;
            const float nInv = 1.0f / (float)n;
            acc = {0.0f, 0.0f, 0.0f};
            for (uint32_t smp = 0; smp < n; smp++)
            {
                nPrimary++;
                const CamDev cs = ra->cam;
                M4 Cs;
                Cs.c[0] = cs.c[0]; Cs.c[1] = cs.c[1]; Cs.c[2] = cs.c[2]; Cs.c[3] = cs.c[3];
                const V3 pos {cs.px, cs.py, cs.pz};
                const float x = (float)i + (float)smp * nInv, y = (float)j + vdc(smp);
                float4 P = make_float4(cs.tlx + cs.hdx * (2.0f * x), cs.tly - cs.hdy * (2.0f * y), -1.0f, 1.0f);
                const V3 rd = versor3(sub3(project4(m4mul(Cs, P)), pos));
                acc = rayRadiance<false, BIG, true>(sc, nc, rg, pos, rd, __builtin_inff(), ra->op.depth, &ra->op.eps, stk,
                                                    overflow, nShadow, nSecondary, unused, acc, &rs);
            }
            acc = mul3(nInv, acc);
        }
        else
        {
            nPrimary++;
            // pixelRayDirection (RayTracer.cu:59-71)
            float4 P = make_float4(cam.tlx + cam.hdx * (2.0f * (float)i), cam.tly - cam.hdy * (2.0f * (float)j), -1.0f, 1.0f);
            const V3 rd = versor3(sub3(project4(m4mul(C, P)), camPos));
            acc = rayRadiance<false, BIG>(sc, nc, rg, camPos, rd, __builtin_inff(), op.depth, &ra->op.eps, stk, overflow,
                                          nShadow, nSecondary, unused);
        }
        const uint32_t q = j * ra->op.width + i;
        if (float* color = ra->color)
        {
            color[3 * (size_t)q] = acc.x;
            color[3 * (size_t)q + 1] = acc.y;
            color[3 * (size_t)q + 2] = acc.z;
        }
        ra->frame[(size_t)j * ra->op.stride + i] = packSRGB(saturate(acc.x), saturate(acc.y), saturate(acc.z));
        nOverflow += overflow;
        // The tile's time, for ordering the next render of this view.  Lane 0
        // holds the tile's first pixel, which always lies inside the frame.
        if (uint32_t* ticks = ra->tileTicks)
        {
            const unsigned long long dt = wall_clock64() - tileStart;
            if (lane == 0) ticks[tile] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
        }
    }
    if (unsigned long long* counts = ra->counts)
    {
        addWave(counts + 0, nPrimary);
        addWave(counts + 1, nShadow);
        addWave(counts + 2, nSecondary);
        addWave(counts + 3, nOverflow);
    }
    workEnd();
}

// Shaded ray queries: each lane traces and shades one given world ray as the
// render does a camera ray (rayRadiance), over one or many jobs.  Rays of all
// jobs are numbered consecutively and waves take chunks of `chunk` rays from
// one counter; each lane finds its job by binary search over the jobs' first
// ray numbers.  With ONE_SCENE every job is on scene `sc0` (a kernel
// argument, like the render's) and the top-level BVH's upper levels are
// staged in LDS as in the render; otherwise each job's own scene is used.
template<bool ONE_SCENE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CGS_RENDER_WAVES))) void kShade(
    SceneDev sc0, const ShadeJob* __restrict__ jobs, uint32_t njobs, uint32_t total, uint32_t chunk, Persist ps,
    unsigned long long* __restrict__ counts)
{
    __shared__ uint32_t ringLds[kRingWords];
    NodeCache nc {nullptr, 0};
    if constexpr (ONE_SCENE)
    {
        __shared__ float4 topLds[4 * CGS_TOP_LDS_KEYS];
        nc = stageTopLevel(sc0, topLds);
    }
    const Rings rg = laneRings(ringLds);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(ps.stack, gl, ps.lanes);
    uint32_t nPrimary = 0, nShadow = 0, nSecondary = 0, nOverflow = 0;
    while (true)
    {
        uint32_t base = 0;
        if ((threadIdx.x & 63) == 0) base = atomicAdd(ps.counter, chunk);
        base = __shfl(base, 0, 64);
        if (base >= total) break;
        const uint32_t end = base + chunk < total ? base + chunk : total;
        for (uint32_t r0 = base; r0 < end; r0 += 64)
        {
            const uint32_t r = r0 + (threadIdx.x & 63);
            if (r >= end) continue;
            uint32_t lo = 0, hi = njobs;  // last job with rayBase <= r
            while (hi - lo > 1)
            {
                const uint32_t mid = (lo + hi) / 2;
                if (jobs[mid].rayBase <= r) lo = mid;
                else hi = mid;
            }
            const ShadeJob& J = jobs[lo];
            const uint32_t k = r - J.rayBase;
            const cgs_ray ray = J.rays[k];
            bool overflow = false;
            nPrimary++;
            ClosestHit h;
            V3 acc;
            if constexpr (ONE_SCENE)
                acc = rayRadiance<true>(sc0, nc, rg, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz}, ray.tmax,
                                        J.depth, &J.eps, stk, overflow, nShadow, nSecondary, h);
            else
                acc = rayRadiance<true>(J.sc, nc, rg, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz}, ray.tmax,
                                        J.depth, &J.eps, stk, overflow, nShadow, nSecondary, h);
            J.color[3 * (size_t)k] = acc.x;
            J.color[3 * (size_t)k + 1] = acc.y;
            J.color[3 * (size_t)k + 2] = acc.z;
            if (uint32_t* px = J.pixels) px[k] = packSRGB(saturate(acc.x), saturate(acc.y), saturate(acc.z));
            if (cgs_hit* hits = J.hits)
            {
                cgs_hit out;
                out.t = h.t;
                out.u = h.object < 0 ? 0.0f : h.u;
                out.v = h.object < 0 ? 0.0f : h.v;
                out.patch = overflow ? CGS_PATCH_OVERFLOW : (h.object < 0 ? EMPTY : h.patch);
                hits[k] = out;
            }
            if (int32_t* objects = J.objects) objects[k] = h.object;
            nOverflow += overflow;
        }
    }
    if (counts)
    {
        addWave(counts + 0, nPrimary);
        addWave(counts + 1, nShadow);
        addWave(counts + 2, nSecondary);
        addWave(counts + 3, nOverflow);
    }
}

} // namespace cgs

#if CGS_SAMPLED_TU
namespace cgs
{
hipError_t launchRenderSampled(const SceneDev& sc, const PersistArgs& pa, const void* argsDev, hipStream_t s)
{
    Persist ps {pa.stack, pa.lanes, pa.counter};
    hipLaunchKernelGGL((kRender<true, true>), dim3(pa.grid), dim3(256), 0, s, sc, ps, (const RenderArgs*)argsDev);
    return hipGetLastError();
}
} // namespace cgs
#else
// ---------------------------------------------------------------------------
// Launchers (called by the host library)
// ---------------------------------------------------------------------------
namespace cgs
{

static inline uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

// Resident workgroups per CU of each persistent kernel (occupancy API).
hipError_t persistentGrid(PersistKernel k, uint32_t& blocksPerCU)
{
    int n = 0;
    hipError_t e = hipSuccess;
    switch (k)
    {
    case PersistKernel::TraceClosest: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTraceClosest<Clip, false>, 256, 0); break;
    case PersistKernel::TraceAny: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTraceAny<Clip>, 256, 0); break;
    case PersistKernel::ClipBatch: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kClipBatch, 256, 0); break;
    case PersistKernel::Render: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kRender<true>, 256, 0); break;
    case PersistKernel::Batch: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTraceBatch<Clip, false, true>, 256, 0); break;
    case PersistKernel::Pick: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kPick<false>, 256, 0); break;
    case PersistKernel::Shade: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kShade<true>, 256, 0); break;
    }
    blocksPerCU = n > 0 ? (uint32_t)n : 1u;
    return e;
}

// Trace launches with an explicit leaf method.  The opt-in modes use the
// default kernels' grid (the same VGPR budget and LDS; they spill to scratch
// where the default kernels do not).  The default kernels clip at the fixed
// 2^-12, so a clipping run with another tolerance takes the capped kernels
// with a cap no run reaches (2^32 - 1 passes).
static TraceParams traceParams(const TraceMode& m)
{
    return {m.tol, m.threshold, m.maxIters ? m.maxIters : 0xFFFFFFFFu};
}

hipError_t launchTraceClosest(const SceneDev& sc, const PersistArgs& pa, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                              int32_t* objects, unsigned long long* overflow, hipStream_t s, const TraceMode& m)
{
    if (n == 0) return hipSuccess;
    Persist ps {pa.stack, pa.lanes, pa.counter};
    const TraceParams tp = traceParams(m);
#define CGS_LAUNCH_CLOSEST(M, P) \
    hipLaunchKernelGGL((kTraceClosest<M, P>), dim3(pa.grid), dim3(256), 0, s, sc, ps, rays, n, hits, objects, overflow, tp)
    if (m.method == CGS_METHOD_SUBDIVISION)
    {
        if (m.polish) CGS_LAUNCH_CLOSEST(Subdiv, true);
        else CGS_LAUNCH_CLOSEST(Subdiv, false);
    }
    else if (m.maxIters || m.tol != kClipTol)
    {
        if (m.polish) CGS_LAUNCH_CLOSEST(ClipCapped, true);
        else CGS_LAUNCH_CLOSEST(ClipCapped, false);
    }
    else if (m.polish) CGS_LAUNCH_CLOSEST(Clip, true);
    else CGS_LAUNCH_CLOSEST(Clip, false);
#undef CGS_LAUNCH_CLOSEST
    return hipGetLastError();
}

hipError_t launchTraceAny(const SceneDev& sc, const PersistArgs& pa, const cgs_ray* rays, uint32_t n, uint32_t* occluded,
                          unsigned long long* overflow, hipStream_t s, const TraceMode& m)
{
    if (n == 0) return hipSuccess;
    Persist ps {pa.stack, pa.lanes, pa.counter};
    const TraceParams tp = traceParams(m);
    if (m.method == CGS_METHOD_SUBDIVISION)
        hipLaunchKernelGGL((kTraceAny<Subdiv>), dim3(pa.grid), dim3(256), 0, s, sc, ps, rays, n, occluded, overflow, tp);
    else if (m.maxIters || m.tol != kClipTol)
        hipLaunchKernelGGL((kTraceAny<ClipCapped>), dim3(pa.grid), dim3(256), 0, s, sc, ps, rays, n, occluded, overflow, tp);
    else
        hipLaunchKernelGGL((kTraceAny<Clip>), dim3(pa.grid), dim3(256), 0, s, sc, ps, rays, n, occluded, overflow, tp);
    return hipGetLastError();
}

hipError_t launchTraceBatch(const BatchJob* jobs, uint32_t njobs, uint32_t total, uint32_t chunk, float2* stack,
                            uint32_t lanes, unsigned int* counter, uint32_t grid, unsigned long long* overflow, hipStream_t s,
                            const TraceMode& m, bool big)
{
    if (total == 0) return hipSuccess;
    Persist ps {stack, lanes, counter};
    const TraceParams tp = traceParams(m);
    // the default method has a kernel without the wide-hash and surface-cull paths
    if (!big && m.method != CGS_METHOD_SUBDIVISION && !m.maxIters && m.tol == kClipTol && !m.polish)
    {
        hipLaunchKernelGGL((kTraceBatch<Clip, false, false>), dim3(grid), dim3(256), 0, s, jobs, njobs, total, chunk, ps,
                           overflow, tp);
        return hipGetLastError();
    }
#define CGS_LAUNCH_BATCH(M, P) \
    hipLaunchKernelGGL((kTraceBatch<M, P, true>), dim3(grid), dim3(256), 0, s, jobs, njobs, total, chunk, ps, overflow, tp)
    if (m.method == CGS_METHOD_SUBDIVISION)
    {
        if (m.polish) CGS_LAUNCH_BATCH(Subdiv, true);
        else CGS_LAUNCH_BATCH(Subdiv, false);
    }
    else if (m.maxIters || m.tol != kClipTol)
    {
        if (m.polish) CGS_LAUNCH_BATCH(ClipCapped, true);
        else CGS_LAUNCH_BATCH(ClipCapped, false);
    }
    else if (m.polish) CGS_LAUNCH_BATCH(Clip, true);
    else CGS_LAUNCH_BATCH(Clip, false);
#undef CGS_LAUNCH_BATCH
    return hipGetLastError();
}

hipError_t launchClipBatch(const PersistArgs& pa, const cgs_ray* rays, const float* patches, uint32_t n, float tol,
                           cgs_hit* hits, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    Persist ps {pa.stack, pa.lanes, pa.counter};
    hipLaunchKernelGGL(kClipBatch, dim3(pa.grid), dim3(256), 0, s, ps, rays, (const float4*)patches, n, tol, hits);
    return hipGetLastError();
}

hipError_t launchSubdivBatch(const cgs_ray* rays, const float* patches, uint32_t n, float threshold, cgs_hit* hits,
                             hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kSubdivBatch, dim3(blocks(n, 256)), dim3(256), 0, s, rays, (const float4*)patches, n, threshold,
                       hits);
    return hipGetLastError();
}

hipError_t launchPatchEval(uint32_t kind, const float* patches, const float* uv, uint32_t n, float* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kPatchEval, dim3(blocks(n, 256)), dim3(256), 0, s, kind, (const float4*)patches,
                       (const float2*)uv, n, (float4*)out);
    return hipGetLastError();
}

hipError_t launchSubpatch(const float* patches, const float* ranges, uint32_t n, float* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kSubpatch, dim3(blocks(n, 256)), dim3(256), 0, s, (const float4*)patches, (const float4*)ranges,
                       n, (float4*)out);
    return hipGetLastError();
}

hipError_t launchShade(const SceneDev* one, const ShadeJob* jobs, uint32_t njobs, uint32_t total, uint32_t chunk,
                       float2* stack, uint32_t lanes, unsigned int* counter, uint32_t grid, unsigned long long* counts,
                       hipStream_t s)
{
    if (total == 0 || njobs == 0) return hipSuccess;
    Persist ps {stack, lanes, counter};
    if (one)
        hipLaunchKernelGGL(kShade<true>, dim3(grid), dim3(256), 0, s, *one, jobs, njobs, total, chunk, ps, counts);
    else
        hipLaunchKernelGGL(kShade<false>, dim3(grid), dim3(256), 0, s, SceneDev {}, jobs, njobs, total, chunk, ps,
                           counts);
    return hipGetLastError();
}

hipError_t launchPick(const SceneDev& sc, const PersistArgs& pa, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                      uint32_t* occluded, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    Persist ps {pa.stack, pa.lanes, pa.counter};
    if (occluded)
        hipLaunchKernelGGL(kPick<true>, dim3(pa.grid), dim3(256), 0, s, sc, ps, rays, n, hits, occluded);
    else
        hipLaunchKernelGGL(kPick<false>, dim3(pa.grid), dim3(256), 0, s, sc, ps, rays, n, hits, occluded);
    return hipGetLastError();
}

hipError_t launchRender(const SceneDev& sc, const PersistArgs& pa, const float* toWorld16, const float* pos3, float tlx,
                        float tly, float hdx, float hdy, float eps, int depth, uint32_t w, uint32_t h, uint32_t* frame,
                        float* color, unsigned long long* counts, uint32_t stride, const uint32_t* order,
                        uint32_t* tileTicks, void* argsDev, hipStream_t s, uint32_t samples, uint32_t seed)
{
    CamDev cam;
    for (int c = 0; c < 4; c++)
        cam.c[c] = make_float4(toWorld16[4 * c], toWorld16[4 * c + 1], toWorld16[4 * c + 2], toWorld16[4 * c + 3]);
    cam.px = pos3[0];
    cam.py = pos3[1];
    cam.pz = pos3[2];
    cam.tlx = tlx;
    cam.tly = tly;
    cam.hdx = hdx;
    cam.hdy = hdy;
    if ((uint64_t)w * h == 0) return hipSuccess;
    const RenderArgs ra {cam, RenderDev {eps, depth, w, h, stride, samples, seed}, frame, color, counts, order, tileTicks};
    static_assert(sizeof(RenderArgs) <= kRenderArgsBytes, "render argument block");
    // `ra` is pageable (the stack): HIP performs such a copy synchronously,
    // so it has been read when the call returns.
    hipError_t e = hipMemcpyAsync(argsDev, &ra, sizeof ra, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    Persist ps {pa.stack, pa.lanes, pa.counter};
    if (samples > 0) return launchRenderSampled(sc, pa, argsDev, s);
    if (sc.big)
        hipLaunchKernelGGL(kRender<true>, dim3(pa.grid), dim3(256), 0, s, sc, ps, (const RenderArgs*)argsDev);
    else
        hipLaunchKernelGGL(kRender<false>, dim3(pa.grid), dim3(256), 0, s, sc, ps, (const RenderArgs*)argsDev);
    return hipGetLastError();
}

// Tile order for the next render of a view: longest first, by a counting
// sort on the tile times (four buckets per octave; order within a bucket is
// arbitrary).  One workgroup; every tile gets exactly one rank.
constexpr int kOrderBuckets = 128;
CGS_DEV uint32_t timeBucket(uint32_t t)
{
    if (t < 2) return 0;
    const uint32_t l = 31 - __builtin_clz(t);
    return 4 * l + (l >= 2 ? (t >> (l - 2)) & 3 : (t << (2 - l)) & 3);
}

__global__ __launch_bounds__(1024) void kTileOrder(const uint32_t* __restrict__ ticks, uint32_t n,
                                                   uint32_t* __restrict__ order)
{
    __shared__ uint32_t count[kOrderBuckets];
    for (uint32_t b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) count[b] = 0;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) atomicAdd(&count[timeBucket(ticks[t])], 1u);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t first = 0;   // the longest bucket first
        for (int b = kOrderBuckets - 1; b >= 0; b--)
        {
            const uint32_t c = count[b];
            count[b] = first;
            first += c;
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) order[atomicAdd(&count[timeBucket(ticks[t])], 1u)] = t;
}

hipError_t launchTileOrder(const uint32_t* ticks, uint32_t n, uint32_t* order, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kTileOrder, dim3(1), dim3(1024), 0, s, ticks, n, order);
    return hipGetLastError();
}

// The work-event totals of the CGS_WORK_COUNTS build (zeros otherwise):
// copies kWorkEvents totals to `out` and clears them when `reset`.
hipError_t workCounts(unsigned long long* out, bool reset)
{
    hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(g_workCounts), sizeof(g_workCounts));
    if (e == hipSuccess && reset)
    {
        const unsigned long long zero[kWorkEvents] = {};
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_workCounts), zero, sizeof zero);
    }
    return e;
}

} // namespace cgs
#endif // CGS_SAMPLED_TU
