// SPDX-License-Identifier: MIT
//
// cgspline oracle — TEST INFRASTRUCTURE ONLY (see cgs_oracle.hpp).
//
// Scene level: BezierSurface (rt/Primitive.cu), two-level Scene (rt/Scene.cu),
// RayTracer shading and trace loop (rt/RayTracer.cu), CPURayTracer
// (rt/CPURayTracer.cpp) and PBR (rt/PBR.{h,cu}), restated on the CPU.
//
#pragma once

#include <atomic>
#include <deque>
#include <future>
#include <mutex>
#include <numbers>
#include <thread>

#include "cgs_oracle.hpp"
#include "cgs_oracle_sample.hpp"

namespace orc
{

using f3 = V3<float>;
using f4 = V4<float>;

// rt/Scene.h:79-95
struct Material
{
    f3 diffuse, specular;
    float metalness, roughness;
};

// rt/Scene.h:47-77
struct Light
{
    f3 color;
    float strength;
    f3 position;
    float range;
    f3 direction;
    float angle;
    bool isDirectional() const { return angle < 0.01f; }
    bool isPoint() const { return angle > 1.57f; }
};

// rt/Primitive.h:90-116 (one BezierSurface per surface) + its BVH and hash.
struct Surface
{
    std::vector<f4> vertices;
    std::vector<uint32_t> indices; // 16 per patch
    BVHData bvh;
    PHF hash;
    std::vector<uint32_t> table;
    uint32_t patchCount() const { return (uint32_t)(indices.size() / 16); }
    void patch(uint32_t p, f4 out[16]) const
    {
        for (int i = 0; i < 16; i++) out[i] = vertices[indices[16 * (size_t)p + i]];
    }
    BVHView view() const { return {bvh.nodes.data(), bvh.indices.data(), table.data(), &hash}; }
};

struct Object
{
    uint32_t surface;
    M4<float> local2world, world2local;
    Material material;
};

struct Scene
{
    std::vector<Surface> surfaces;
    std::vector<Object> objects;
    std::vector<Light> lights;
    BVHData top;
    PHF topHash;
    std::vector<uint32_t> topTable;
    f3 background {0.1f, 0.1f, 0.1f};
    f3 ambient {0.1f, 0.1f, 0.1f};
    BVHView topView() const { return {top.nodes.data(), top.indices.data(), topTable.data(), &topHash}; }
};

// Builds every BVH and hash table (MainWindow.cpp:363-492 convertScene order:
// per-surface BVHs, top-level BVH, then the hash tables).  Returns 0 or the
// first failing hash build's code.
// surface0 (optional): surface 0's BVH elements, as cg::PatchBVH gives a
// patch table's patches their boxes (PatchBVH.cpp:40-80).
inline int buildScene(Scene& s, uint32_t seed = 42, const std::vector<Element>* surface0 = nullptr)
{
    for (size_t k = 0; k < s.surfaces.size(); k++)
    {
        auto& surf = s.surfaces[k];
        surf.bvh = k == 0 && surface0 ? bvhBuild(*surface0, 1)
                                      : bvhBuild(patchElements(&surf.vertices[0].x, surf.indices.data(), surf.patchCount()), 1);
    }
    // rt/Scene.cu:39-53 — objects bounded by their world-space control points.
    // An object without a surface (EMPTY: the reference's Mesh and Sphere,
    // whose bounds() are empty, Primitive.cu:15-23,47-56) keeps the empty box:
    // its centroid is the box centre (0) and no ray ever enters its leaf.
    std::vector<Element> el(s.objects.size());
    for (size_t i = 0; i < s.objects.size(); i++)
    {
        const Object& o = s.objects[i];
        Element& e = el[i];
        e.index = (uint32_t)i;
        e.bounds = Box<float>();
        if (o.surface != EMPTY)
        {
            const Surface& surf = s.surfaces[o.surface];
            for (uint32_t idx : surf.indices)
            {
                f4 q = o.local2world * surf.vertices[idx];
                e.bounds.inflate(q.xyz());
            }
        }
        e.centroid = e.bounds.center();
    }
    s.top = bvhBuild(el, 1);
    // A BVH whose reference-family build fails (a duplicate: from 2^21 heap
    // keys on) is hashed with the wide family instead; a bucket overflow
    // fails both.
    auto hashOne = [&](PHF& h, const std::vector<uint32_t>& keys, std::vector<uint32_t>& table) -> int {
        int r = phfBuild<uint32_t>(h, keys, seed);
        if (r == 1) r = phfBuild<uint32_t>(h, keys, seed, true);
        if (r) return r;
        table = hashTable<uint32_t>(h, keys);
        return 0;
    };
    for (auto& surf : s.surfaces)
        if (int r = hashOne(surf.hash, surf.bvh.keys, surf.table)) return r;
    return hashOne(s.topHash, s.top.keys, s.topTable);
}

struct TraceHit
{
    float t = std::numeric_limits<float>::infinity();
    float u = 0, v = 0;
    uint32_t patch = EMPTY;
    int object = -1;
};

// Optional log of every closest-hit patch test (test instrumentation: the
// clipper's inputs — local origin, max, local direction — and the control
// points), set per thread by orc_scene_record_tests.
struct PatchTestLog
{
    std::vector<float> rays;     // 8 per test: o xyz, initial best t, d xyz, 0
    std::vector<float> patches;  // 64 per test
};
inline thread_local PatchTestLog* g_patchLog = nullptr;

// Leaf method of a query, as cgs_trace_options selects it: Bézier clipping at
// the reference's tolerance by default (Primitive.cu:234,260); subdivision
// (the commented-out alternative, Primitive.cu:236,263); the opt-in clip
// pass cap (no counterpart in the reference).  The opt-in Newton polish
// (polishHit64) refines the reported closest hit after the query, once per
// ray, so it changes no traversal decision.
struct TraceOpts
{
    int method = 0;          // 0 clipping, 1 subdivision
    float tol = 0x1p-12f;
    float threshold = 0x1p-4f;
    uint32_t maxIters = 0;
    bool polish = false;
};

inline bool leafTest(const TraceOpts& op, PatchHit<float>& ph, f3 origin, f3 dir, const f4 S[16], Work* w)
{
    return op.method == 1 ? doSubdivision<float>(ph, origin, dir, S, op.threshold)
                          : doBezierClipping<float>(ph, origin, dir, S, op.tol, w, op.maxIters);
}

// rt/Primitive.cu:222-248 — closest hit against one surface (local ray).
inline bool surfaceIntersect(const Surface& s, Hit& hit0, const Ray& ray0, Work* w, const TraceOpts& op = {})
{
    // Ds Ray3f{origin, direction} normalizes the direction; the result is the
    // same for every patch of this ray, so it is computed once here.
    const f3 dir = versor(ray0.direction);
    auto fn = [&](Hit& hit, uint32_t index) -> bool
    {
        PatchHit<float> ph {hit.t, {hit.coordinates.x, hit.coordinates.y}};
        f4 S[16];
        s.patch(index, S);
        if (g_patchLog)
        {
            const float r[8] {ray0.origin.x, ray0.origin.y, ray0.origin.z, hit.t,
                              ray0.direction.x, ray0.direction.y, ray0.direction.z, 0.0f};
            g_patchLog->rays.insert(g_patchLog->rays.end(), r, r + 8);
            for (int i = 0; i < 16; i++)
                g_patchLog->patches.insert(g_patchLog->patches.end(), {S[i].x, S[i].y, S[i].z, S[i].w});
        }
        if (leafTest(op, ph, ray0.origin, dir, S, w))
        {
            hit.index = index;
            hit.t = ph.distance;
            hit.coordinates = {ph.p.x, ph.p.y, 0.0f};
            return true;
        }
        return false;
    };
    return hashIntersect(s.view(), hit0, ray0, fn, w);
}

// rt/Primitive.cu:250-268 — any hit against one surface (local ray).
inline bool surfaceIntersectAny(const Surface& s, const Ray& ray0, Work* w, const TraceOpts& op = {})
{
    const f3 dir = versor(ray0.direction);
    auto fn = [&](uint32_t index) -> bool
    {
        PatchHit<float> ph {ray0.max, {0, 0}};
        f4 S[16];
        s.patch(index, S);
        return leafTest(op, ph, ray0.origin, dir, S, w);
    };
    return hashIntersectAny(s.view(), ray0, fn, w);
}

// World ray -> object-local ray (RayTracer.cu:305-314,333): t scales by |M d|.
inline Ray localRay(const Object& o, const Ray& ray, float& dinv)
{
    Ray lr;
    f4 O = o.world2local * f4 {ray.origin.x, ray.origin.y, ray.origin.z, 1.0f};
    lr.origin = project(O);
    lr.direction = M3<float>::from4(o.world2local) * ray.direction;
    float d = lr.direction.length();
    lr.max = ray.max * d;
    d = 1 / d;
    lr.direction = d * lr.direction;
    dinv = d;
    return lr;
}

// RayTracer.cu:296-350 — Context::intersect(hit, ray): returns the nearest object.
inline int sceneIntersect(const Scene& sc, Hit& hit0, const Ray& ray0, uint32_t* patchOut, Work* w,
                          const TraceOpts& op = {})
{
    int nearest = -1;
    uint64_t bottom = 0;
    const uint64_t before = w ? w->nodes : 0;
    auto fn = [&](Hit& hit, uint32_t i) -> bool
    {
        if (w) w->objects++;
        const Object& o = sc.objects[i];
        if (o.surface == EMPTY) return false;   // no geometry (never reached: empty box)
        float d;
        Ray lr = localRay(o, ray0, d);
        Hit lh;
        lh.t = lr.max;
        const uint64_t n0 = w ? w->nodes : 0;
        bool b = surfaceIntersect(sc.surfaces[o.surface], lh, lr, w, op);
        if (w) bottom += w->nodes - n0;
        if (b)
        {
            float t = CGS_ORACLE_PLANT == 7 ? lh.t : lh.t * d;
            if (t < hit.t)
            {
                hit.coordinates = lh.coordinates;
                hit.index = lh.index;
                hit.t = t;
                nearest = (int)i;
                return true;
            }
        }
        return false;
    };
    hashIntersect(sc.topView(), hit0, ray0, fn, w);
    if (w) w->topNodes += w->nodes - before - bottom;
    if (patchOut) *patchOut = hit0.index;
    return nearest;
}

// The same closest-hit query with the explicit-stack traversal of
// rt/BVH.cpp:126-189 at both levels (a cross-check of the hashed one).
inline int sceneIntersectStack(const Scene& sc, Hit& hit0, const Ray& ray0, uint32_t* patchOut)
{
    int nearest = -1;
    auto fn = [&](Hit& hit, uint32_t i) -> bool
    {
        const Object& o = sc.objects[i];
        float d;
        if (o.surface == EMPTY) return false;   // no geometry (never reached: empty box)
        Ray lr = localRay(o, ray0, d);
        Hit lh;
        lh.t = lr.max;
        const Surface& s = sc.surfaces[o.surface];
        const f3 dir = versor(lr.direction);
        auto leaf = [&](Hit& h, uint32_t index) -> bool
        {
            PatchHit<float> ph {h.t, {h.coordinates.x, h.coordinates.y}};
            f4 S[16];
            s.patch(index, S);
            if (doBezierClipping<float>(ph, lr.origin, dir, S))
            {
                h.index = index;
                h.t = ph.distance;
                h.coordinates = {ph.p.x, ph.p.y, 0.0f};
                return true;
            }
            return false;
        };
        bool b = stackIntersect(s.view(), lh, lr, leaf);
        if (b)
        {
            float t = lh.t * d;
            if (t < hit.t)
            {
                hit.coordinates = lh.coordinates;
                hit.index = lh.index;
                hit.t = t;
                nearest = (int)i;
                return true;
            }
        }
        return false;
    };
    stackIntersect(sc.topView(), hit0, ray0, fn);
    if (patchOut) *patchOut = hit0.index;
    return nearest;
}

// RayTracer.cu:352-383 — Context::intersect(ray): any hit.
inline bool sceneIntersectAny(const Scene& sc, const Ray& ray0, Work* w, const TraceOpts& op = {})
{
    uint64_t bottom = 0;
    const uint64_t before = w ? w->nodes : 0;
    auto fn = [&](uint32_t i) -> bool
    {
        if (w) w->objects++;
        const Object& o = sc.objects[i];
        if (o.surface == EMPTY) return false;   // no geometry (never reached: empty box)
        float d;
        Ray lr = localRay(o, ray0, d);
        const uint64_t n0 = w ? w->nodes : 0;
        const bool b = surfaceIntersectAny(sc.surfaces[o.surface], lr, w, op);
        if (w) bottom += w->nodes - n0;
        return b;
    };
    const bool any = hashIntersectAny(sc.topView(), ray0, fn, w);
    if (w) w->topNodes += w->nodes - before - bottom;
    return any;
}

// ---------------------------------------------------------------------------
// PBR (rt/PBR.h, rt/PBR.cu)
// ---------------------------------------------------------------------------

inline f3 splat(float s) { return {s, s, s}; }

// PBR.cu:21-42. A spot light (0.01 <= angle <= 1.57, Scene.h:73-76) never
// passes: DL < 0 puts acos(DL) above pi/2, so angle > 2 acos(DL) needs angle > pi.
inline bool lightVector(float& d, f3& L, f3 P, const Light& light)
{
    if (light.isDirectional())
    {
        L = -light.direction;
        d = std::numeric_limits<float>::infinity();
        return true;
    }
    L = light.position - P;
    d = L.length();
    if (d < 1e-14f || (light.range > 0 && d > light.range)) return false;
    L = (1 / d) * L;
    if (light.isPoint()) return true;
    float DL = light.direction.dot(L);
    return DL < 0 && light.angle > 2 * std::acos(DL);
}
// PBR.cu:44-64
inline f3 lightColor(float d, const Light& light)
{
    if (light.isDirectional()) return light.color;
    float range = light.range, f;
    if (range == 0)
    {
        f = 1 / d;
        f *= f;
    }
    else
    {
        f = d / range;
        f = 1 + f * (f - 2);
    }
    return f * light.color;
}
// schlick (PBR.h:53-61; vec3 and float instantiations)
inline f3 schlick(f3 f0, float dotLH)
{
    float b = 1 - dotLH;
    float b2 = b * b;
    return f0 + (splat(1) - f0) * splat(b2 * b2 * b);
}
// G1, G (PBR.h:72-85)
inline float G1(float dotNX, float k) { return dotNX / (dotNX * (1 - k) + k); }
inline float Gsmith(float dotNL, float dotNV, float r)
{
    const float r1 = r + 1;
    const float k = (float)((double)(r1 * r1) * 0.125);
    return G1(dotNL, k) * G1(dotNV, k);
}
// D (PBR.h:87-96)
inline float Dggx(float dotNH, float r)
{
    r = fmax_(r, 1e-3f);
    float a = r * r;
    float a2 = a * a;
    float d = dotNH * dotNH * (a2 - 1) + 1;
    return a2 / (std::numbers::pi_v<float> * d * d);
}
// BRDF_microfacet (PBR.h:98-107)
inline float microfacet(float dotNV, float dotNL, float dotNH, float r)
{
    return Gsmith(dotNL, dotNV, r) * Dggx(dotNH, r) / (4 * dotNL * dotNV);
}
// BRDF_diffuse (PBR.h:109-114)
inline f3 brdfDiffuse(const Material& m) { return std::numbers::inv_pi_v<float> * m.diffuse; }
// BRDF_specular (PBR.h:131-141)
inline f3 brdfSpecular(float dotHL, float dotHN, float dotNV, float dotNL, float r, f3 f0)
{
    return microfacet(dotNV, dotNL, dotHN, r) * schlick(f0, dotHL);
}
// reflect (PBR.h:34-38)
inline f3 reflect(f3 I, f3 N, float dotIN) { return I - (2 * dotIN) * N; }

// Primitive.cu:270-275 — shading normal from mat::normal, normalized.
inline f3 surfaceNormal(const Surface& s, uint32_t patch, float u, float v)
{
    f4 S[16];
    s.patch(patch, S);
    return versor(normalMat<float>(S, u, v));
}

// Ds-boundary: pack_sRGB — IEC 61966-2-1 sRGB encode of each clamped channel,
// rounded to 8 bits, packed as 0xAABBGGRR with alpha 255.
inline uint32_t packSRGB(float r, float g, float b)
{
    auto enc = [](float c) -> uint32_t
    {
        float e = c <= 0.0031308f ? 12.92f * c : 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
        int q = (int)(e * 255.0f + 0.5f);
        return (uint32_t)std::clamp(q, 0, 255);
    };
    return enc(r) | (enc(g) << 8) | (enc(b) << 16) | 0xFF000000u;
}

// Camera (RayTracer.cu:141-163): pixel centres at integer (i,j).
struct Camera
{
    M4<float> cameraToWorld;
    f3 position;
    float fovy;
    bool flipY = false;
};
struct CameraSetup
{
    M4<float> toWorld;
    f3 position;
    V2<float> topLeft;
    float half_dx, half_dy;
};
inline CameraSetup setupCamera(const Camera& c, uint32_t w, uint32_t h)
{
    CameraSetup s;
    s.toWorld = c.cameraToWorld;
    s.position = c.position;
    float halfHeight = std::tan(CGS_ORACLE_PLANT == 3 ? c.fovy : 0.5f * c.fovy);
    s.half_dy = halfHeight / (float)h;
    s.half_dx = s.half_dy;
    s.topLeft.x = -s.half_dx * (float)(w - 1);
    s.topLeft.y = halfHeight - s.half_dy;
    if (c.flipY)
    {
        s.topLeft.y = -s.topLeft.y;
        s.half_dy = -s.half_dy;
    }
    return s;
}
// Context::pixelRayDirection (RayTracer.cu:59-71)
inline f3 pixelRayDirection(const CameraSetup& c, float x, float y)
{
    if constexpr (CGS_ORACLE_PLANT == 4)
    {
        x += 0.5f;
        y += 0.5f;
    }
    f4 P {c.topLeft.x + c.half_dx * (2 * x), c.topLeft.y - c.half_dy * (2 * y), -1.0f, 1.0f};
    P = c.toWorld * P;
    return versor(project(P) - c.position);
}

struct RenderOptions
{
    float eps = 1e-4f;
    int recursionDepth = 6;
};

struct RayCounts
{
    uint64_t primary = 0, shadow = 0, secondary = 0;
};

// The first segment's closest hit of a shaded ray (t, u, v, patch, object).
struct FirstHit
{
    float t = 0, u = 0, v = 0;
    uint32_t patch = EMPTY;
    int object = -1;
};

// The direct light of Context::closestHit (RayTracer.cu:398-466): from
// `color`, for each light that is in range, faces the normal and whose shadow
// ray reaches it, the PBR diffuse + specular terms.
inline f3 directLight(const Scene& sc, const Material& m, f3 diffuse, f3 N, f3 V, float dotNV, f3 P, f3 color,
    const RenderOptions& op, Work* w, RayCounts* rc)
{
    for (const Light& light : sc.lights)
    {
        float d;
        f3 L;
        if (!lightVector(d, L, P, light)) continue;
        const float dotNL = N.dot(L);
        if (dotNL < 1e-14f) continue;
        Ray r1 {P + op.eps * L, L, d};
        if (rc) rc->shadow++;
        if (sceneIntersectAny(sc, r1, w)) continue;
        f3 H = versor(L + V);
        float dotHN = H.dot(N);
        float dotHL = H.dot(L);
        f3 I = lightColor(d, light);
        color = color + (dotNL * I) * (diffuse + m.metalness * brdfSpecular(dotHL, dotHN, dotNV, dotNL, m.roughness, m.specular));
    }
    return color;
}

// The shading normal of closestHit (RayTracer.cu:398-431): the patch normal
// to world space by the transposed world→local matrix, normalized, turned to
// face V; dotNV = N·V >= 0.
inline f3 shadingNormal(const Scene& sc, const Object& o, const Hit& hit, f3 V, float& dotNV)
{
    f3 N = surfaceNormal(sc.surfaces[o.surface], hit.index, hit.coordinates.x, hit.coordinates.y);
    N = versor(M3<float>::from4(o.world2local).transposed() * N);
    dotNV = N.dot(V);
    if (dotNV < 0)
    {
        N = -N;
        dotNV = -dotNV;
    }
    return N;
}

// render kernel and Context::trace/miss/closestHit (RayTracer.cu:194-294,385-544, non-Monte-Carlo path): the GPU renderer's
// iterative trace loop from one ray (the render starts it at a camera ray with max = infinity).  Returns the
// unclamped colour; `first` (nullable) receives the first segment's hit.
inline f3 shadeRay(const Scene& sc, const Ray& ray, const RenderOptions& op, Work* w, RayCounts* rc,
    FirstHit* firstHit = nullptr)
{
    struct Payload
    {
        Ray ray;
        f3 color, attenuation;
        int depth;
    } pl;
    pl.color = splat(0);
    pl.ray = ray;
    pl.attenuation = splat(1);
    pl.depth = op.recursionDepth;
    bool first = true;
    while (true)
    {
        if (rc) (first ? rc->primary : rc->secondary)++;
        Hit hit;
        hit.t = pl.ray.max;
        uint32_t patch = EMPTY;
        int obj = sceneIntersect(sc, hit, pl.ray, &patch, w);
        if (first && firstHit)
            *firstHit = obj < 0 ? FirstHit {hit.t, 0, 0, EMPTY, -1}
                                : FirstHit {hit.t, hit.coordinates.x, hit.coordinates.y, patch, obj};
        first = false;
        if (obj < 0)
        {
            pl.color = pl.color + sc.background * pl.attenuation;
            break;
        }
        // closestHit (RayTracer.cu:398-544)
        const Object& o = sc.objects[obj];
        const Material& m = o.material;
        const f3 diffuse = (1 - m.metalness) * brdfDiffuse(m);
        f3 V = -pl.ray.direction;
        float dotNV;
        const f3 N = shadingNormal(sc, o, hit, V, dotNV);
        f3 P = pl.ray.origin + hit.t * pl.ray.direction;
        const f3 color = directLight(sc, m, diffuse, N, V, dotNV, P, sc.ambient * diffuse, op, w, rc);
        pl.color = pl.color + ((CGS_ORACLE_PLANT == 5 ? 1.0f : std::numbers::pi_v<float>) * color) * pl.attenuation;
        if (pl.depth <= 0) break;
        f3 reflectance = schlick(m.specular, dotNV);
        f3 R = reflect(-V, N, -dotNV);
        float g = Gsmith(dotNV, dotNV, m.roughness);
        float rr = CGS_ORACLE_PLANT == 6 ? 1.0f : 1 - m.roughness;
        f3 brdf = (rr * rr) * (dotNV * (diffuse + m.metalness * (g * reflectance)));
        f3 a = pl.attenuation * brdf;
        if (a.max() <= 0x1p-8f) break;
        pl.ray = {P + op.eps * R, R, std::numeric_limits<float>::infinity()};
        pl.attenuation = a;
        pl.depth = pl.depth - 1;
    }
    return pl.color;
}

// The packed frame pixel of a colour (RayTracer.cu:262-266): __saturatef
// (NaN -> 0) on each channel, then pack_sRGB.
inline uint32_t packColor(f3 c)
{
    auto sat = [](float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; };
    return packSRGB(sat(c.x), sat(c.y), sat(c.z));
}

// One pixel of the render kernel (RayTracer.cu:194-294): the camera ray
// through pixel (i, j) with max = infinity, shaded by shadeRay.  Writes the
// unclamped colour and the packed pixel.
inline void renderPixel(const Scene& sc, const CameraSetup& cam, const RenderOptions& op,
    uint32_t i, uint32_t j, f3& colorOut, uint32_t& pixel, Work* w, RayCounts* rc)
{
    const Ray ray {cam.position, pixelRayDirection(cam, (float)i, (float)j), std::numeric_limits<float>::infinity()};
    colorOut = shadeRay(sc, ray, op, w, rc);
    pixel = packColor(colorOut);
}

// ---------------------------------------------------------------------------
// Monte Carlo render (USE_MONTECARLO_SAMPLING: RayTracer.cu:212-246 and the
// sampled branches of closestHit, 403-407 and 473-522).  The draws, sample
// positions and angle functions are this build's (cgs_oracle_sample.hpp).
// ---------------------------------------------------------------------------

// One pixel's random stream: the seed, the reference's grid cell and the
// number of draws taken so far (the reference's copied curand state).
struct SampleStream
{
    uint32_t seed, stream, draw;
};

// The sampled reflection of closestHit (RayTracer.cu:398-522): a microfacet
// normal M from (x, y) by microfacet(roughness², x, y) (PBR.h:159-176) in the
// frame t_s = N×V, t_p = N×t_s, N; R = V mirrored about M; the path ends where
// N·R < 1e-14.  brdf = mix(diffuse BRDF, F(M·V) G(N·R, N·V) (M·V) / (N·V N·M),
// metalness) (PBR.h:27-31).
inline bool sampledBounce(f3 N, f3 V, float dotNV, const Material& m, float x, float y, f3& R, f3& brdf)
{
    const float alpha = CGS_ORACLE_PLANT == 11 ? m.roughness : m.roughness * m.roughness;
    const float thetaM = atan2Quadrant(alpha * std::sqrt(x), std::sqrt(1 - x));
    const float phiM = 2 * std::numbers::pi_v<float> * y;
    const f3 ts = versor(N.cross(V));
    const f3 tp = versor(N.cross(ts));
    float sinT, cosT, sinP, cosP;
    sinCosAngle(thetaM, sinT, cosT);
    sinCosAngle(phiM, sinP, cosP);
    f3 M = (sinT * cosP) * ts + (sinT * sinP) * tp + cosT * N;
    M = versor(M);
    const float dotMV = M.dot(V);
    const f3 reflectance = schlick(m.specular, dotMV);
    R = reflect(-V, M, -dotMV);
    const float dotMN = M.dot(N);
    const float dotNR = N.dot(R);
    if (dotNR < 1e-14f) return false;
    const float G = Gsmith(dotNR, dotNV, m.roughness);
    const float g = CGS_ORACLE_PLANT == 12 ? G : G * dotMV / (dotNV * dotMN);
    brdf = (1 - m.metalness) * brdfDiffuse(m) + m.metalness * (g * reflectance);
    return true;
}

// The sampled trace loop from one ray (RayTracer.cu:232-246 with the sampled
// closestHit): the payload colour `color` carries over from earlier samples
// and takes each hit's direct light (no ambient term) and the background at a
// miss, times the attenuation; while depth remains, a bounce takes two draws.
inline void shadeRaySampled(const Scene& sc, const Ray& ray, const RenderOptions& op, SampleStream& rs, f3& color,
    Work* w, RayCounts* rc)
{
    Ray r = ray;
    f3 att = splat(1);
    int depth = op.recursionDepth;
    bool first = true;
    while (true)
    {
        if (rc) (first ? rc->primary : rc->secondary)++;
        first = false;
        Hit hit;
        hit.t = r.max;
        uint32_t patch = EMPTY;
        const int obj = sceneIntersect(sc, hit, r, &patch, w);
        if (obj < 0)
        {
            color = color + sc.background * att;
            break;
        }
        const Object& o = sc.objects[obj];
        const Material& m = o.material;
        const f3 diffuse = (1 - m.metalness) * brdfDiffuse(m);
        const f3 V = -r.direction;
        float dotNV;
        const f3 N = shadingNormal(sc, o, hit, V, dotNV);
        const f3 P = r.origin + hit.t * r.direction;
        const f3 direct = directLight(sc, m, diffuse, N, V, dotNV, P, splat(0), op, w, rc);
        color = color + (std::numbers::pi_v<float> * direct) * att;
        if (depth <= 0) break;
        const uint32_t m0 = rs.draw;
        rs.draw += 2;
        const float x = unitFloat(sampleWord(rs.seed, rs.stream, m0));
        const float y = unitFloat(sampleWord(rs.seed, rs.stream, m0 + 1));
        f3 R, brdf;
        if (!sampledBounce(N, V, dotNV, m, x, y, R, brdf)) break;
        const f3 a = att * brdf;
        if (a.max() <= 0x1p-8f) break;
        r = {P + op.eps * R, R, std::numeric_limits<float>::infinity()};
        att = a;
        depth = depth - 1;
    }
}

// One pixel of the sampled render (RayTracer.cu:212-246): `samples` camera
// rays through (i + s/n, j + vdc(s)), one running colour, scaled by 1/n; the
// stream is the pixel's cell of the 256×256 grid (RayTracer.cu:213-218).
inline void renderPixelSampled(const Scene& sc, const CameraSetup& cam, const RenderOptions& op, uint32_t samples,
    uint32_t seed, uint32_t i, uint32_t j, f3& colorOut, uint32_t& pixel, Work* w, RayCounts* rc)
{
    SampleStream rs {seed, (i & 255u) + 256u * (j & 255u), 0u};
    const float nInv = 1.0f / (float)samples;
    f3 color = splat(0);
    for (uint32_t s = 0; s < samples; s++)
    {
        const float s0 = (float)s * nInv;
        const float s1 = CGS_ORACLE_PLANT == 10 ? s0 : vdc(s);
        const Ray ray {cam.position, pixelRayDirection(cam, (float)i + s0, (float)j + s1),
                       std::numeric_limits<float>::infinity()};
        shadeRaySampled(sc, ray, op, rs, color, w, rc);
    }
    colorOut = nInv * color;
    pixel = packColor(colorOut);
}

// ---------------------------------------------------------------------------
// CPURayTracer (rt/CPURayTracer.cpp:55-440), non-Monte-Carlo path: the CPU
// baseline.  Recursive trace, 32×32 tiles pulled from a mutex-guarded deque
// by `threads` std::async workers.
// ---------------------------------------------------------------------------
class CPURayTracer
{
public:
    struct Options
    {
        float eps = 1e-4f;
        bool flipYAxis = false;
        int recursionDepth = 8;
        int threads = (int)std::thread::hardware_concurrency();
        int tileSize = 32;
    } options;

    std::atomic<uint64_t> rays {0}, shadowRays {0};

    // Renders rows row0, row0 + stride, ... below row1 (a bounded sample of
    // the frame for timing; stride 1 = every row) and returns the wall time
    // from entry to the last tile, in seconds (CPURayTracer.cpp:72,137).
    // Tiles are 32×32 over the sampled rows.
    double render(const Scene* scene, const Camera& camera, uint32_t w, uint32_t h,
        float* color, uint32_t* frame, uint32_t row0 = 0, uint32_t row1 = 0xFFFFFFFFu, uint32_t stride = 1)
    {
        _stride = std::max(1u, stride);
        auto t0 = std::chrono::steady_clock::now();
        options.threads = std::clamp(options.threads, 1, 0x1000);
        options.tileSize = std::max(options.tileSize, 8);
        _scene = scene;
        _w = w;
        _h = h;
        _color = color;
        _frame = frame;
        row1 = std::min(row1, h);
        _row0 = row0;
        _row1 = row1;
        _cam = setupCamera(Camera {camera.cameraToWorld, camera.position, camera.fovy, options.flipYAxis}, w, h);
        uint32_t countX = (w + options.tileSize - 1) / options.tileSize;
        _nrows = (row1 - row0 + _stride - 1) / _stride;
        uint32_t countY = (_nrows + options.tileSize - 1) / options.tileSize;
        _work.clear();
        for (int j = (int)countY - 1; j >= 0; j--)
            for (int i = (int)countX - 1; i >= 0; i--)
                _work.push_back({uint32_t(i * options.tileSize), uint32_t(j * options.tileSize)});
        std::vector<std::future<void>> workers;
        for (int i = 0; i < options.threads; i++)
            workers.emplace_back(std::async(std::launch::async, &CPURayTracer::work, this));
        for (auto& f : workers) f.wait();
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

private:
    struct Tile { uint32_t x, y; };
    const Scene* _scene = nullptr;
    uint32_t _w = 0, _h = 0, _row0 = 0, _row1 = 0, _stride = 1, _nrows = 0;
    float* _color = nullptr;
    uint32_t* _frame = nullptr;
    CameraSetup _cam;
    std::deque<Tile> _work;
    std::mutex _mutex;

    void work()
    {
        while (true)
        {
            Tile t;
            {
                std::lock_guard lck(_mutex);
                if (_work.empty()) return;
                t = _work.back();
                _work.pop_back();
            }
            uint32_t x1 = std::min(_w, t.x + options.tileSize);
            uint32_t y1 = std::min(_nrows, t.y + options.tileSize);
            for (uint32_t jj = t.y; jj < y1; jj++)
                for (uint32_t i = t.x; i < x1; i++)
                {
                    const uint32_t j = _row0 + jj * _stride;
                    Ray r {_cam.position, pixelRayDirection(_cam, (float)i, (float)j), std::numeric_limits<float>::infinity()};
                    f3 c = trace(r, splat(1), options.recursionDepth);
                    size_t k = (size_t)j * _w + i;
                    if (_color)
                    {
                        _color[3 * k] = c.x;
                        _color[3 * k + 1] = c.y;
                        _color[3 * k + 2] = c.z;
                    }
                    auto cl = [](float x) { return std::clamp(x, 0.0f, 1.0f); };
                    if (_frame) _frame[k] = packSRGB(cl(c.x), cl(c.y), cl(c.z));
                }
        }
    }

    // CPURayTracer::trace (CPURayTracer.cpp:199-219)
    f3 trace(const Ray& ray, f3 attenuation, int depth)
    {
        rays.fetch_add(1, std::memory_order_relaxed);
        Hit hit;
        hit.t = ray.max;
        int obj = sceneIntersect(*_scene, hit, ray, nullptr, nullptr);
        if (obj < 0) return attenuation * _scene->background;
        return closestHit(hit, ray, attenuation, (uint32_t)obj, depth);
    }

    // CPURayTracer::closestHit (CPURayTracer.cpp:291-440, non-MC branch)
    f3 closestHit(const Hit& hit, const Ray& ray, f3 attenuation, uint32_t object, int depth)
    {
        const Object& o = _scene->objects[object];
        const Material& m = o.material;
        const f3 diffuse = (1 - m.metalness) * brdfDiffuse(m);
        f3 color = _scene->ambient * diffuse;
        f3 V = -ray.direction;
        f3 N = versor(M3<float>::from4(o.world2local).transposed() *
            surfaceNormal(_scene->surfaces[o.surface], hit.index, hit.coordinates.x, hit.coordinates.y));
        float dotNV = N.dot(V);
        if (dotNV < 0)
        {
            N = -N;
            dotNV = -dotNV;
        }
        f3 P = ray.origin + hit.t * ray.direction;
        for (const Light& light : _scene->lights)
        {
            float d;
            f3 L;
            if (!lightVector(d, L, P, light)) continue;
            const float dotNL = N.dot(L);
            if (dotNL < 1e-14f) continue;
            Ray r1 {P + options.eps * L, L, d};
            shadowRays.fetch_add(1, std::memory_order_relaxed);
            if (sceneIntersectAny(*_scene, r1, nullptr)) continue;
            f3 H = versor(L + V);
            float dotHN = H.dot(N);
            float dotHL = H.dot(L);
            f3 I = lightColor(d, light);
            color = color + (dotNL * I) * (diffuse + m.metalness * brdfSpecular(dotHL, dotHN, dotNV, dotNL, m.roughness, m.specular));
        }
        color = color * (std::numbers::pi_v<float> * attenuation);
        if (depth <= 0) return color;
        f3 reflectance = schlick(m.specular, dotNV);
        f3 R = reflect(-V, N, -dotNV);
        float g = Gsmith(dotNV, dotNV, m.roughness);
        float rr = CGS_ORACLE_PLANT == 6 ? 1.0f : 1 - m.roughness;
        f3 brdf = (rr * rr) * (dotNV * (diffuse + m.metalness * (g * reflectance)));
        f3 a = attenuation * brdf;
        if (a.max() > 0x1p-8f)
        {
            Ray r {P + options.eps * R, R, std::numeric_limits<float>::infinity()};
            color = color + trace(r, a, depth - 1);
        }
        return color;
    }
};

} // namespace orc
