// SPDX-License-Identifier: MIT
//
// Staged (wavefront) ray tracer for gfx950.
//
// The reference runs each ray's whole ray tree in one thread: top-level
// hashed traversal → per-object local ray → per-object hashed traversal →
// Bézier clipping at each leaf (RayTracer.cu:268-618).  On a GPU that puts the
// clipper's working patch and hull in the same kernel as the traversal loops,
// and the kernel's register count (hence its occupancy) is set by the
// clipper although traversal is most of the work.  Here the same per-ray
// algorithm is cut at the leaf test:
//
//   kTraverse  resumes a ray's two-level traversal from its saved state and
//              runs until it needs a patch test (emits a clip request) or is
//              done.  Lean: no patch in registers.
//   kClip      runs doBezierClipping for each request and returns the ray to
//              the active list.
//
// The host alternates the two until no ray is active.  Shading, shadow rays
// and the reflection bounce are separate stages (kShade, kScatterShadow,
// kAccumulate, kPack).  Every operation and its order per ray is unchanged,
// so the results equal the single-thread formulation bit for bit.
//
#include <hip/hip_runtime.h>

#include "cgs_internal.hpp"
#include "cgs_scene_dev.hpp"
#include "cgs_stage.hpp"

namespace cgs
{

constexpr float kTol = 0x1p-12f; // Spline.h:493

// flag bits of a ray slot
constexpr uint32_t F_ANY = 1u, F_INOBJ = 2u, F_OBJHIT = 4u, F_DONE = 8u, F_OVERFLOW = 16u, F_OCCLUDED = 32u;

CGS_DEV uint32_t laneId() { return threadIdx.x & 63; }

CGS_DEV uint32_t nextChunk64(unsigned int* counter)
{
    uint32_t base = 0;
    if (laneId() == 0) base = atomicAdd(counter, 64u);
    return __shfl(base, 0, 64);
}

// Wave-aggregated append: returns this lane's slot (only meaningful if pred).
CGS_DEV uint32_t appendWave(unsigned int* counter, bool pred)
{
    unsigned long long m = __ballot(pred);
    uint32_t total = __popcll(m);
    uint32_t leader = m ? (uint32_t)__builtin_ctzll(m) : 0;
    uint32_t base = 0;
    if (m && laneId() == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader, 64);
    unsigned long long below = m & ((1ull << laneId()) - 1ull);
    return base + (uint32_t)__popcll(below);
}

// One traversal level's resumable state (rt/BVH.h:178-350): current node,
// heap key, postponed bits, position in the current leaf's element list
// (EMPTY = not inside a leaf).
struct Level
{
    uint32_t node, key, post, leafPos;
};

// Advances `L` in BVH `b` until the next leaf element (returned) or the end
// (EMPTY).  `limit` culls child boxes (L0 < limit).
CGS_DEV uint32_t stepToElement(const BVHRef& b, Level& L, V3 o, V3 Dinv, float limit)
{
    while (true)
    {
        if (L.leafPos != EMPTY)
        {
            const uint32_t e = b.indices[L.leafPos];
            if (e != EMPTY)
            {
                L.leafPos++;
                return e;
            }
            L.leafPos = EMPTY;
            // leaf done: ascend
        }
        else
        {
            NodeBoxes n = loadNode(b.nodes, L.node);
            if (n.left == EMPTY)
            {
                L.leafPos = n.right;
                continue;
            }
            float L0, R0;
            bool iL = nodeIntersect(L0, n.lmin, n.lmax, o, Dinv) && L0 < limit;
            bool iR = nodeIntersect(R0, n.rmin, n.rmax, o, Dinv) && R0 < limit;
            if (iL || iR) L.post <<= 1;
            if (iL)
            {
                if (iR)
                {
                    L.post |= 1;
                    if (L0 < R0) { L.key = L.key << 1; L.node = n.left; }
                    else { L.key = (L.key << 1) ^ 1u; L.node = n.right; }
                }
                else { L.key = L.key << 1; L.node = n.left; }
                continue;
            }
            else if (iR)
            {
                L.key = (L.key << 1) ^ 1u;
                L.node = n.right;
                continue;
            }
        }
        // going up
        if (!L.post) return EMPTY;
        int levels = __builtin_ctz(L.post);
        L.post >>= levels;
        L.key = (L.key >> levels) ^ 1u;
        L.post ^= 1;
        L.node = b.table[hashKey(b, L.key)];
        if (L.node >= b.nodeCount) return EMPTY; // a corrupt table ends the ray instead of faulting
    }
}

// ---------------------------------------------------------------------------
// traversal stage
// ---------------------------------------------------------------------------
// INLINE = false: stop at the next patch test and emit a clip request.
// INLINE = true : run the patch tests in place and finish the ray (used for
//                 the last few rays, whose long serial traversal would
//                 otherwise cost one staged round per patch test).
#ifndef CGS_FINISH_WAVES
#define CGS_FINISH_WAVES 2
#endif
template<bool INLINE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(INLINE ? CGS_FINISH_WAVES : 1))) void kTraverse(SceneDev sc, RayBatch rb, const uint32_t* __restrict__ active,
                                                 const unsigned int* __restrict__ nActive, ClipQueue cq,
                                                 unsigned int* __restrict__ work, float2* __restrict__ stack,
                                                 uint32_t lanes)
{
    const uint32_t n = *nActive;
    const ClipStack stk = clipStackFor(stack, blockIdx.x * blockDim.x + threadIdx.x, lanes);
    for (uint32_t base = nextChunk64(work); base < n; base = nextChunk64(work))
    {
        const uint32_t k = base + laneId();
        bool emit = false;
        uint32_t r = 0, patch = 0;
        if (k < n)
        {
            r = active[k];
            uint32_t flags = rb.flags[r];
            const bool any = flags & F_ANY;
            const float4 od0 = rb.od0[r], od1 = rb.od1[r];
            const V3 o {od0.x, od0.y, od0.z}, d {od1.x, od1.y, od1.z};
            const float tmax = od0.w;
            float bestT = od1.w;
            const uint4 tp = rb.top[r];
            Level top {tp.x, tp.y, tp.z, tp.w};
            const BVHRef tb = bvhRef(sc, sc.top);
            const V3 Dinv {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
            Level bot;
            float4 lr0, lr1, lh;
            uint32_t obj = rb.obj[r];
            if (flags & F_INOBJ)
            {
                uint4 bb = rb.bot[r];
                bot = {bb.x, bb.y, bb.z, bb.w};
                lr0 = rb.lr0[r];
                lr1 = rb.lr1[r];
                lh = rb.lh[r];
            }
            while (!(flags & F_DONE))
            {
                if (flags & F_INOBJ)
                {
                    const DevBVH surf = sc.surfaces[sc.objects[obj].surface];
                    const BVHRef b = bvhRef(sc, surf);
                    const V3 lo {lr0.x, lr0.y, lr0.z}, ld {lr1.x, lr1.y, lr1.z};
                    const V3 lDinv {1.0f / ld.x, 1.0f / ld.y, 1.0f / ld.z};
                    // closest: cull by the object's own best (lt); any: by lmax
                    const uint32_t e = stepToElement(b, bot, lo, lDinv, any ? lr0.w : lh.x);
                    if (e != EMPTY)
                    {
                        if constexpr (!INLINE)
                        {
                            emit = true;
                            patch = e;
                            break;
                        }
                        else
                        {
                            // the patch test of kClip, in place
                            const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + e);
                            const ClipRay cr = makeClipRay(lo, versor3(ld));
                            bool overflow = false;
                            if (any)
                            {
                                float best = lr0.w, bu, bv;
                                if (clipPatch<true>(S, cr, kTol, best, bu, bv, stk, overflow)) flags |= F_DONE | F_OCCLUDED;
                            }
                            else
                            {
                                float lt = lh.x, lu = lh.y, lv = lh.z;
                                if (clipPatch<false>(S, cr, kTol, lt, lu, lv, stk, overflow))
                                {
                                    lh = make_float4(lt, lu, lv, __uint_as_float(e));
                                    flags |= F_OBJHIT;
                                }
                            }
                            if (overflow) flags |= F_OVERFLOW;
                            continue;
                        }
                    }
                    // object finished (RayTracer.cu:403-415)
                    flags &= ~F_INOBJ;
                    if (!any && (flags & F_OBJHIT))
                    {
                        float t = lh.x * lr1.w;
                        if (t < bestT)
                        {
                            bestT = t;
                            rb.hit[r] = make_float4(lh.y, lh.z, __uint_as_float(__float_as_uint(lh.w)), __uint_as_float(obj));
                        }
                    }
                    flags &= ~F_OBJHIT;
                }
                else
                {
                    const uint32_t i = stepToElement(tb, top, o, Dinv, any ? tmax : bestT);
                    if (i == EMPTY)
                    {
                        flags |= F_DONE;
                        break;
                    }
                    // enter object i: local ray (RayTracer.cu:379-389); an object
                    // without geometry (its box is empty, so only a NaN ray gets
                    // here) has nothing to test
                    obj = i;
                    const DevObject& ob = sc.objects[i];
                    if (ob.surface == EMPTY) continue;
                    M4 M;
                    M.c[0] = ob.w2l[0]; M.c[1] = ob.w2l[1]; M.c[2] = ob.w2l[2]; M.c[3] = ob.w2l[3];
                    V3 lo = project4(m4mul(M, make_float4(o.x, o.y, o.z, 1.0f)));
                    V3 ld = m3mul(M, d);
                    float len = len3(ld);
                    float lmax = tmax * len;
                    float dinv = 1.0f / len;
                    ld = mul3(dinv, ld);
                    lr0 = make_float4(lo.x, lo.y, lo.z, lmax);
                    lr1 = make_float4(ld.x, ld.y, ld.z, dinv);
                    lh = make_float4(lmax, 0.0f, 0.0f, __uint_as_float(EMPTY));
                    const DevBVH surf = sc.surfaces[ob.surface];
                    bot = {surf.root, 1u, 0u, EMPTY};
                    flags |= F_INOBJ;
                }
            }
            // save
            rb.flags[r] = flags;
            rb.od1[r] = make_float4(od1.x, od1.y, od1.z, bestT);
            rb.top[r] = make_uint4(top.node, top.key, top.post, top.leafPos);
            rb.obj[r] = obj;
            if (flags & F_INOBJ)
            {
                rb.bot[r] = make_uint4(bot.node, bot.key, bot.post, bot.leafPos);
                rb.lr0[r] = lr0;
                rb.lr1[r] = lr1;
                rb.lh[r] = lh;
            }
        }
        if constexpr (!INLINE)
        {
            const uint32_t slot = appendWave(cq.count, emit);
            if (emit) cq.req[slot] = make_uint2(r, patch);
        }
    }
}

// ---------------------------------------------------------------------------
// clipping stage
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kClip(SceneDev sc, RayBatch rb, ClipQueue cq, float2* __restrict__ stack,
                                             uint32_t lanes, uint32_t* __restrict__ nextActive,
                                             unsigned int* __restrict__ nNext, unsigned int* __restrict__ work)
{
    const uint32_t n = *cq.count;
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
    const ClipStack stk = clipStackFor(stack, gl, lanes);
    for (uint32_t base = nextChunk64(work); base < n; base = nextChunk64(work))
    {
        const uint32_t k = base + laneId();
        bool keep = false;
        uint32_t r = 0;
        if (k < n)
        {
            const uint2 q = cq.req[k];
            r = q.x;
            const uint32_t pid = q.y;
            uint32_t flags = rb.flags[r];
            const bool any = flags & F_ANY;
            const float4 lr0 = rb.lr0[r], lr1 = rb.lr1[r];
            float4 lh = rb.lh[r];
            const uint32_t obj = rb.obj[r];
            const DevBVH surf = sc.surfaces[sc.objects[obj].surface];
            const float4* S = sc.patches + 16 * (size_t)(surf.patchBase + pid);
            // Ray3f normalizes the local direction (Primitive.cu:362)
            const ClipRay cr = makeClipRay({lr0.x, lr0.y, lr0.z}, versor3({lr1.x, lr1.y, lr1.z}));
            bool overflow = false;
            if (any)
            {
                float best = lr0.w, bu, bv;
                if (clipPatch<true>(S, cr, kTol, best, bu, bv, stk, overflow))
                    flags |= F_DONE | F_OCCLUDED;
            }
            else
            {
                float lt = lh.x, lu = lh.y, lv = lh.z;
                if (clipPatch<false>(S, cr, kTol, lt, lu, lv, stk, overflow))
                {
                    lh = make_float4(lt, lu, lv, __uint_as_float(pid));
                    flags |= F_OBJHIT;
                    rb.lh[r] = lh;
                }
            }
            if (overflow) flags |= F_OVERFLOW;
            rb.flags[r] = flags;
            keep = !(flags & F_DONE);
        }
        const uint32_t slot = appendWave(nNext, keep);
        if (keep) nextActive[slot] = r;
    }
}

// ---------------------------------------------------------------------------
// batch setup
// ---------------------------------------------------------------------------

CGS_DEV void initSlot(RayBatch& rb, uint32_t r, V3 o, V3 d, float tmax, bool any, const SceneDev& sc)
{
    rb.od0[r] = make_float4(o.x, o.y, o.z, tmax);
    rb.od1[r] = make_float4(d.x, d.y, d.z, tmax);
    rb.hit[r] = make_float4(0.0f, 0.0f, __uint_as_float(EMPTY), __uint_as_float(0xFFFFFFFFu));
    rb.top[r] = make_uint4(sc.top.root, 1u, 0u, EMPTY);
    rb.flags[r] = any ? F_ANY : 0u;
    rb.obj[r] = 0;
}

// User rays → slots 0..n-1 and the initial active list.
__global__ void kLoadRays(SceneDev sc, RayBatch rb, const cgs_ray* __restrict__ rays, uint32_t n, int any,
                          uint32_t* __restrict__ active)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const cgs_ray ray = rays[r];
    initSlot(rb, r, {ray.ox, ray.oy, ray.oz}, {ray.dx, ray.dy, ray.dz}, ray.tmax, any != 0, sc);
    active[r] = r;
}

__global__ void kStoreClosest(RayBatch rb, uint32_t n, cgs_hit* __restrict__ hits, int32_t* __restrict__ objects,
                              unsigned long long* __restrict__ overflowCount)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    bool ovf = false;
    if (r < n)
    {
        const float4 h = rb.hit[r];
        const uint32_t flags = rb.flags[r];
        const int32_t obj = (int32_t)__float_as_uint(h.w);
        ovf = flags & F_OVERFLOW;
        cgs_hit out;
        out.t = rb.od1[r].w;
        out.u = obj < 0 ? 0.0f : h.x;
        out.v = obj < 0 ? 0.0f : h.y;
        out.patch = ovf ? CGS_PATCH_OVERFLOW : (obj < 0 ? EMPTY : __float_as_uint(h.z));
        hits[r] = out;
        if (objects) objects[r] = obj;
    }
    unsigned long long m = __ballot(ovf);
    if (m && laneId() == (uint32_t)__builtin_ctzll(m)) atomicAdd(overflowCount, (unsigned long long)__popcll(m));
}

__global__ void kStoreAny(RayBatch rb, uint32_t n, uint32_t* __restrict__ occluded,
                          unsigned long long* __restrict__ overflowCount)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    bool ovf = false;
    if (r < n)
    {
        const uint32_t flags = rb.flags[r];
        ovf = flags & F_OVERFLOW;
        occluded[r] = ovf ? 2u : ((flags & F_OCCLUDED) ? 1u : 0u);
    }
    unsigned long long m = __ballot(ovf);
    if (m && laneId() == (uint32_t)__builtin_ctzll(m)) atomicAdd(overflowCount, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// render stages (RayTracer.cu:268-618, non-Monte-Carlo path)
// ---------------------------------------------------------------------------

// Primary rays through integer pixel centres (RayTracer.cu:133-145); slots
// in 8×8 tile order so a wave traces one tile.
__global__ void kGenPrimary(SceneDev sc, RayBatch rb, CamDevS cam, uint32_t w, uint32_t h, PixelState px,
                            uint32_t* __restrict__ active, unsigned int* __restrict__ nActive,
                            unsigned long long* __restrict__ counts)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tilesX = (w + 7) / 8, tilesY = (h + 7) / 8;
    const uint32_t tile = s / 64, l = s % 64;
    const uint32_t i = (tile % tilesX) * 8 + (l & 7), j = (tile / tilesX) * 8 + (l >> 3);
    const bool valid = s < tilesX * tilesY * 64 && i < w && j < h;
    if (counts)
    {
        unsigned long long m = __ballot(valid);
        if (m && laneId() == (uint32_t)__builtin_ctzll(m)) atomicAdd(counts, (unsigned long long)__popcll(m));
    }
    const uint32_t slot = appendWave(nActive, valid);
    if (!valid) return;
    M4 C;
    C.c[0] = cam.c[0]; C.c[1] = cam.c[1]; C.c[2] = cam.c[2]; C.c[3] = cam.c[3];
    const V3 camPos {cam.px, cam.py, cam.pz};
    float4 P = make_float4(cam.tlx + cam.hdx * (2.0f * (float)i), cam.tly - cam.hdy * (2.0f * (float)j), -1.0f, 1.0f);
    V3 rd = versor3(sub3(project4(m4mul(C, P)), camPos));
    const uint32_t pix = j * w + i;
    initSlot(rb, slot, camPos, rd, __builtin_inff(), false, sc);
    rb.pixel[slot] = pix;
    active[slot] = slot;
    px.acc[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    px.att[pix] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
}

CGS_DEV bool lightVectorS(float& d, V3& L, V3 P, const DevLight& lt)
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
CGS_DEV V3 lightColorS(float d, const DevLight& lt)
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
CGS_DEV V3 schlickS(V3 f0, float dotLH)
{
    float b = 1.0f - dotLH;
    float b2 = b * b;
    float s = b2 * b2 * b;
    return {f0.x + (1.0f - f0.x) * s, f0.y + (1.0f - f0.y) * s, f0.z + (1.0f - f0.z) * s};
}
CGS_DEV float G1S(float dotNX, float k) { return dotNX / (dotNX * (1.0f - k) + k); }
CGS_DEV float GsmithS(float dotNL, float dotNV, float r)
{
    const float r1 = r + 1.0f;
    const float k = (r1 * r1) * 0.125f;
    return G1S(dotNL, k) * G1S(dotNV, k);
}
CGS_DEV float DggxS(float dotNH, float r)
{
    r = fmax_(r, 1e-3f);
    float a = r * r;
    float a2 = a * a;
    float d = dotNH * dotNH * (a2 - 1.0f) + 1.0f;
    return a2 / (0x1.921fb6p+1f * d * d);
}
CGS_DEV float microfacetS(float dotNV, float dotNL, float dotNH, float r)
{
    return GsmithS(dotNL, dotNV, r) * DggxS(dotNH, r) / (4.0f * dotNL * dotNV);
}

// closestHit (RayTracer.cu:471-618) up to the shadow tests: the ambient term,
// one shadow ray and one pending light term per lit light, and the
// reflection ray of the next bounce.
__global__ __launch_bounds__(256) void kShade(SceneDev sc, RayBatch rb, uint32_t n, PixelState px, int depth, float eps,
                                              RayBatch next, uint32_t* __restrict__ nextActive,
                                              unsigned int* __restrict__ nNext, RayBatch shadow,
                                              uint32_t* __restrict__ shadowActive, unsigned int* __restrict__ nShadow,
                                              unsigned long long* __restrict__ counts)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    bool bounce = false;
    uint32_t pix = 0;
    V3 Pt {0, 0, 0}, R {0, 0, 0};
    uint32_t nShadowRays = 0;
    unsigned long long ovf = 0;
    if (r < n)
    {
        pix = rb.pixel[r];
        const uint32_t flags = rb.flags[r];
        ovf = (flags & F_OVERFLOW) ? 1 : 0;
        const float4 h = rb.hit[r];
        const int32_t object = (int32_t)__float_as_uint(h.w);
        const float4 att4 = px.att[pix];
        const V3 att {att4.x, att4.y, att4.z};
        if (object < 0)
        {
            float4 a = px.acc[pix];
            V3 acc = add3({a.x, a.y, a.z}, mulv3({sc.background[0], sc.background[1], sc.background[2]}, att));
            px.acc[pix] = make_float4(acc.x, acc.y, acc.z, 0.0f);
            px.state[pix] = 0; // no light terms pending
        }
        else
        {
            const float4 od0 = rb.od0[r], od1 = rb.od1[r];
            const V3 ro {od0.x, od0.y, od0.z}, rd {od1.x, od1.y, od1.z};
            const float t = od1.w;
            const DevObject& ob = sc.objects[object];
            const float4 m0 = ob.mat0, m1 = ob.mat1;
            const float metal = m0.w, rough = m1.w;
            const V3 spec {m1.x, m1.y, m1.z};
            const V3 diffuse = mul3(1.0f - metal, mul3(0x1.45f306p-2f, {m0.x, m0.y, m0.z}));
            const V3 col = mulv3({sc.ambient[0], sc.ambient[1], sc.ambient[2]}, diffuse);
            const V3 V = neg3(rd);
            const DevBVH surf = sc.surfaces[ob.surface];
            V3 N = versor3(normalMat(sc.patches + 16 * (size_t)(surf.patchBase + __float_as_uint(h.z)), h.x, h.y));
            M4 W;
            W.c[0] = ob.w2l[0]; W.c[1] = ob.w2l[1]; W.c[2] = ob.w2l[2]; W.c[3] = ob.w2l[3];
            N = versor3(m3tmul(W, N));
            float dotNV = dot3(N, V);
            if (dotNV < 0.0f)
            {
                N = neg3(N);
                dotNV = -dotNV;
            }
            Pt = add3(ro, mul3(t, rd));
            px.base[pix] = make_float4(col.x, col.y, col.z, 0.0f);
            uint32_t pending = 0;
            for (uint32_t li = 0; li < sc.nlight; li++)
            {
                const DevLight lt = sc.lights[li];
                float d;
                V3 L;
                if (!lightVectorS(d, L, Pt, lt)) continue;
                const float dotNL = dot3(N, L);
                if (dotNL < 1e-14f) continue;
                V3 H = versor3(add3(L, V));
                float dotHN = dot3(H, N);
                float dotHL = dot3(H, L);
                V3 I = lightColorS(d, lt);
                V3 bs = mul3(microfacetS(dotNV, dotNL, dotHN, rough), schlickS(spec, dotHL));
                V3 term = mulv3(mul3(dotNL, I), add3(diffuse, mul3(metal, bs)));
                px.term[(size_t)pix * px.nlight + li] = make_float4(term.x, term.y, term.z, 0.0f);
                pending |= 1u << li;
                // the shadow ray (RayTracer.cu:527-530)
                const uint32_t s = atomicAdd(nShadow, 1u);
                initSlot(shadow, s, add3(Pt, mul3(eps, L)), L, d, true, sc);
                shadow.pixel[s] = pix * 32u + li;
                shadowActive[s] = s;
                nShadowRays++;
            }
            px.state[pix] = pending | 0x80000000u; // hit: accumulate after the shadow pass
            if (depth > 0)
            {
                V3 refl = schlickS(spec, dotNV);
                R = sub3(neg3(V), mul3(2.0f * -dotNV, N));
                float g = GsmithS(dotNV, dotNV, rough);
                float rr = 1.0f - rough;
                V3 brdf = mul3(rr * rr, mul3(dotNV, add3(diffuse, mul3(metal, mul3(g, refl)))));
                V3 a = mulv3(att, brdf);
                if (!(max3(a) <= 0x1p-8f))
                {
                    bounce = true;
                    px.attNext[pix] = make_float4(a.x, a.y, a.z, 0.0f);
                }
            }
        }
    }
    const uint32_t slot = appendWave(nNext, bounce);
    if (bounce)
    {
        initSlot(next, slot, add3(Pt, mul3(eps, R)), R, __builtin_inff(), false, sc);
        next.pixel[slot] = pix;
        nextActive[slot] = slot;
    }
    if (counts)
    {
        unsigned long long v[3] = {nShadowRays, bounce ? 1ull : 0ull, ovf};
        int dst[3] = {1, 2, 3};
#pragma unroll
        for (int c = 0; c < 3; c++)
        {
            unsigned long long x = v[c];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if (laneId() == 0 && x) atomicAdd(counts + dst[c], x);
        }
    }
}

// Shadow-ray results → per (pixel, light) occlusion bits.
__global__ void kScatterShadow(RayBatch shadow, uint32_t n, uint32_t* __restrict__ occ,
                               unsigned long long* __restrict__ counts)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    bool ovf = false;
    if (s < n)
    {
        const uint32_t code = shadow.pixel[s];
        const uint32_t flags = shadow.flags[s];
        ovf = flags & F_OVERFLOW;
        if (flags & F_OCCLUDED) atomicOr(occ + code / 32u, 1u << (code % 32u));
    }
    unsigned long long m = __ballot(ovf);
    if (counts && m && laneId() == (uint32_t)__builtin_ctzll(m)) atomicAdd(counts + 3, (unsigned long long)__popcll(m));
}

// Sums the light terms in light order (skipping occluded lights) and adds
// (π·col)·att to the pixel (RayTracer.cu:513-542).  Pending bits were set by
// kShade; kScatterShadow sets an occluded light's bit in `occ`, so a term is
// added when it is pending and not occluded.
__global__ void kAccumulate(RayBatch rb, uint32_t n, PixelState px, const uint32_t* __restrict__ occ)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t pix = rb.pixel[r];
    const uint32_t st = px.state[pix];
    if (!(st & 0x80000000u)) return;
    const float4 b = px.base[pix];
    V3 col {b.x, b.y, b.z};
    const uint32_t o = occ[pix];
    for (uint32_t li = 0; li < px.nlight; li++)
        if (((st >> li) & 1u) && !((o >> li) & 1u))
        {
            const float4 t = px.term[(size_t)pix * px.nlight + li];
            col = add3(col, {t.x, t.y, t.z});
        }
    const float4 a4 = px.att[pix], c4 = px.acc[pix];
    V3 acc = add3({c4.x, c4.y, c4.z}, mulv3(mul3(0x1.921fb6p+1f, col), {a4.x, a4.y, a4.z}));
    px.acc[pix] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

// The bounce's attenuation becomes the pixel's attenuation.
__global__ void kAdvanceAtt(RayBatch next, uint32_t n, PixelState px)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t pix = next.pixel[r];
    px.att[pix] = px.attNext[pix];
}

CGS_DEV uint32_t packSRGBS(float r, float g, float b)
{
    auto enc = [](float c) -> uint32_t {
        float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
        int q = (int)(e * 255.0f + 0.5f);
        q = q < 0 ? 0 : (q > 255 ? 255 : q);
        return (uint32_t)q;
    };
    return enc(r) | (enc(g) << 8) | (enc(b) << 16) | 0xFF000000u;
}
CGS_DEV float sat(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

__global__ void kPack(PixelState px, uint32_t npix, uint32_t width, uint32_t stride, uint32_t* __restrict__ frame,
                      float* __restrict__ color)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float4 a = px.acc[p];
    frame[(size_t)(p / width) * stride + p % width] = packSRGBS(sat(a.x), sat(a.y), sat(a.z));
    if (color)
    {
        color[3 * (size_t)p] = a.x;
        color[3 * (size_t)p + 1] = a.y;
        color[3 * (size_t)p + 2] = a.z;
    }
}

__global__ void kZero(uint32_t* __restrict__ p, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline uint32_t nb(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

hipError_t stageGrid(int which, uint32_t& blocksPerCU)
{
    int n = 0;
    hipError_t e = which == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTraverse<false>, 256, 0)
                   : which == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kClip, 256, 0)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTraverse<true>, 256, 0);
    blocksPerCU = n > 0 ? (uint32_t)n : 1u;
    return e;
}

hipError_t launchTraverse(const SceneDev& sc, const RayBatch& rb, const uint32_t* active, const unsigned int* nActive,
                          const ClipQueue& cq, unsigned int* work, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(kTraverse<false>, dim3(grid), dim3(256), 0, s, sc, rb, active, nActive, cq, work, nullptr, 0u);
    return hipGetLastError();
}

hipError_t launchFinish(const SceneDev& sc, const RayBatch& rb, const uint32_t* active, const unsigned int* nActive,
                        unsigned int* work, float2* stack, uint32_t lanes, uint32_t grid, hipStream_t s)
{
    ClipQueue none {nullptr, nullptr};
    hipLaunchKernelGGL(kTraverse<true>, dim3(grid), dim3(256), 0, s, sc, rb, active, nActive, none, work, stack, lanes);
    return hipGetLastError();
}

hipError_t launchClip(const SceneDev& sc, const RayBatch& rb, const ClipQueue& cq, float2* stack, uint32_t lanes,
                      uint32_t* nextActive, unsigned int* nNext, unsigned int* work, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(kClip, dim3(grid), dim3(256), 0, s, sc, rb, cq, stack, lanes, nextActive, nNext, work);
    return hipGetLastError();
}

hipError_t launchLoadRays(const SceneDev& sc, const RayBatch& rb, const cgs_ray* rays, uint32_t n, int any,
                          uint32_t* active, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kLoadRays, dim3(nb(n, 256)), dim3(256), 0, s, sc, rb, rays, n, any, active);
    return hipGetLastError();
}

hipError_t launchStoreClosest(const RayBatch& rb, uint32_t n, cgs_hit* hits, int32_t* objects,
                              unsigned long long* overflow, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kStoreClosest, dim3(nb(n, 256)), dim3(256), 0, s, rb, n, hits, objects, overflow);
    return hipGetLastError();
}

hipError_t launchStoreAny(const RayBatch& rb, uint32_t n, uint32_t* occluded, unsigned long long* overflow, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kStoreAny, dim3(nb(n, 256)), dim3(256), 0, s, rb, n, occluded, overflow);
    return hipGetLastError();
}

hipError_t launchGenPrimary(const SceneDev& sc, const RayBatch& rb, const CamDevS& cam, uint32_t w, uint32_t h,
                            const PixelState& px, uint32_t* active, unsigned int* nActive, unsigned long long* counts,
                            hipStream_t s)
{
    uint64_t n = (uint64_t)((w + 7) / 8) * ((h + 7) / 8) * 64;
    hipLaunchKernelGGL(kGenPrimary, dim3(nb(n, 256)), dim3(256), 0, s, sc, rb, cam, w, h, px, active, nActive, counts);
    return hipGetLastError();
}

hipError_t launchShade(const SceneDev& sc, const RayBatch& rb, uint32_t n, const PixelState& px, int depth, float eps,
                       const RayBatch& next, uint32_t* nextActive, unsigned int* nNext, const RayBatch& shadow,
                       uint32_t* shadowActive, unsigned int* nShadow, unsigned long long* counts, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kShade, dim3(nb(n, 256)), dim3(256), 0, s, sc, rb, n, px, depth, eps, next, nextActive, nNext,
                       shadow, shadowActive, nShadow, counts);
    return hipGetLastError();
}

hipError_t launchScatterShadow(const RayBatch& shadow, uint32_t n, uint32_t* occ, unsigned long long* counts,
                               hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kScatterShadow, dim3(nb(n, 256)), dim3(256), 0, s, shadow, n, occ, counts);
    return hipGetLastError();
}

hipError_t launchAccumulate(const RayBatch& rb, uint32_t n, const PixelState& px, const uint32_t* occ, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kAccumulate, dim3(nb(n, 256)), dim3(256), 0, s, rb, n, px, occ);
    return hipGetLastError();
}

hipError_t launchAdvanceAtt(const RayBatch& next, uint32_t n, const PixelState& px, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdvanceAtt, dim3(nb(n, 256)), dim3(256), 0, s, next, n, px);
    return hipGetLastError();
}

hipError_t launchPack(const PixelState& px, uint32_t npix, uint32_t width, uint32_t stride, uint32_t* frame, float* color,
                      hipStream_t s)
{
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(kPack, dim3(nb(npix, 256)), dim3(256), 0, s, px, npix, width, stride, frame, color);
    return hipGetLastError();
}

hipError_t launchZero(uint32_t* p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kZero, dim3(nb(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

} // namespace cgs
