// SPDX-License-Identifier: MIT
//
// cgspline oracle — TEST INFRASTRUCTURE ONLY (see cgs_oracle.hpp).
//
// Plain C entry points over the CPU restatement, loaded with ctypes by the
// tests and by bench.py's cpu_baseline leg.  Nothing in the product loads
// this library.
//
#include <chrono>
#include <cstdio>
#include <memory>

#include "cgs_oracle_scene.hpp"

using namespace orc;

template<class F> static void parallelFor(uint64_t n, int threads, F&& f)
{
    if (threads <= 1 || n < 64)
    {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + threads - 1) / threads;
    for (int k = 0; k < threads; k++)
    {
        uint64_t a = k * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&, a, b] { f(a, b); });
    }
    for (auto& t : th) t.join();
}

extern "C" {

struct orc_work
{
    uint64_t nodes, leaves, hashProbes, patchTests, clipIters, subdivs, prunes, converged, onHits, maxStack, objects,
        maxPostponed, overflows, topNodes;
};

static void putWork(orc_work* o, const Work& w)
{
    if (!o) return;
    *o = {w.nodes, w.leaves, w.hashProbes, w.patchTests, w.clipIters, w.subdivs, w.prunes, w.converged, w.onHits, w.maxStack, w.objects,
          w.maxPostponed, w.overflows, w.topNodes};
}

int orc_version(void) { return 1; }

// Caps the clip's region stack of the calling thread (0: unbounded, the
// reference's semantics); see clipStackCap.
void orc_set_clip_stack_cap(int cap) { clipStackCap = cap; }

// --- per-patch intersection ------------------------------------------------
// rays: n × 8 floats (ox, oy, oz, tmax, dx, dy, dz, pad); the direction is
// normalized as the Ds Ray3f constructor does.  patches: n × 64 floats
// (16 control points × (x,y,z,w), point(i,j) at [4j+i]).  The hit starts at
// distance tmax.  out: n × 4 (t, u, v, found) in the chosen precision.
int orc_clip_batch(int fp64, const float* rays, const float* patches, uint64_t n, float tol,
    float* out32, double* out64, orc_work* work, int method, float threshold)
{
    Work w;
    for (uint64_t r = 0; r < n; r++)
    {
        const float* ry = rays + 8 * r;
        const float* pt = patches + 64 * r;
        if (!fp64)
        {
            V3<float> o {ry[0], ry[1], ry[2]}, d {ry[4], ry[5], ry[6]};
            d = versor(d);
            V4<float> S[16];
            for (int i = 0; i < 16; i++) S[i] = {pt[4 * i], pt[4 * i + 1], pt[4 * i + 2], pt[4 * i + 3]};
            PatchHit<float> h {ry[3], {0, 0}};
            bool f = method == 0 ? doBezierClipping<float>(h, o, d, S, tol, &w)
                                 : doSubdivision<float>(h, o, d, S, threshold);
            out32[4 * r] = h.distance;
            out32[4 * r + 1] = h.p.x;
            out32[4 * r + 2] = h.p.y;
            out32[4 * r + 3] = f ? 1.0f : 0.0f;
        }
        else
        {
            V3<double> o {ry[0], ry[1], ry[2]}, d {ry[4], ry[5], ry[6]};
            d = versor(d);
            V4<double> S[16];
            for (int i = 0; i < 16; i++) S[i] = {pt[4 * i], pt[4 * i + 1], pt[4 * i + 2], pt[4 * i + 3]};
            PatchHit<double> h {(double)ry[3], {0, 0}};
            bool f = method == 0 ? doBezierClipping<double>(h, o, d, S, (double)tol, nullptr)
                                 : doSubdivision<double>(h, o, d, S, (double)threshold);
            out64[4 * r] = h.distance;
            out64[4 * r + 1] = h.p.x;
            out64[4 * r + 2] = h.p.y;
            out64[4 * r + 3] = f ? 1.0 : 0.0;
        }
    }
    putWork(work, w);
    return 0;
}

// 2D clipping on an already projected patch (16 × (x,y)); records every
// onHit candidate (u,v) in call order (up to cap) and returns the count.
int orc_clip2d_candidates(const float* p2d, float tol, float* uv_out, int cap, orc_work* work)
{
    V2<float> P[16];
    for (int i = 0; i < 16; i++) P[i] = {p2d[2 * i], p2d[2 * i + 1]};
    int count = 0;
    Work w;
    auto onHit = [&](V2<float> e) -> bool
    {
        if (count < cap)
        {
            uv_out[2 * count] = e.x;
            uv_out[2 * count + 1] = e.y;
        }
        count++;
        return true;
    };
    clip2D<float>(onHit, P, tol, &w);
    putWork(work, w);
    return count;
}

// --- primitive math --------------------------------------------------------
// kind: 0 interpolate (de Casteljau), 1 interpolate (matrix), 2 dU (dC),
// 3 dV (dC), 4 dU (matrix), 5 dV (matrix), 6 normal (dC, degenerate fix),
// 7 normal (matrix).  fp64 evaluates in double.  out: 4 values.
int orc_eval(int fp64, const float* patch, double u, double v, int kind, double* out)
{
    auto run = [&](auto zero) -> int
    {
        using T = decltype(zero);
        V4<T> S[16];
        for (int i = 0; i < 16; i++) S[i] = {(T)patch[4 * i], (T)patch[4 * i + 1], (T)patch[4 * i + 2], (T)patch[4 * i + 3]};
        V4<T> r {0, 0, 0, 0};
        T uu = (T)u, vv = (T)v;
        switch (kind)
        {
        case 0: r = interpolateDC<T>(S, uu, vv); break;
        case 1: r = interpolateMat<T>(S, uu, vv); break;
        case 2: r = derivativeU_DC<T>(S, uu, vv); break;
        case 3: r = derivativeV_DC<T>(S, uu, vv); break;
        case 4: r = derivativeUMat<T>(S, uu, vv); break;
        case 5: r = derivativeVMat<T>(S, uu, vv); break;
        case 6: { V3<T> n = normalDC<T>(S, uu, vv); r = {n.x, n.y, n.z, 0}; break; }
        case 7: { V3<T> n = normalMat<T>(S, uu, vv); r = {n.x, n.y, n.z, 0}; break; }
        default: return 1;
        }
        out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
        return 0;
    };
    return fp64 ? run(0.0) : run(0.0f);
}

// Spline.h:384-396 — subpatch of a 16-point (x,y,z,w) patch in place:
// subpatchV to [vmin, vmax], then subpatchU to [umin, umax].
void orc_subpatch(float* patch64, float umin, float umax, float vmin, float vmax)
{
    V4<float> P[16];
    for (int i = 0; i < 16; i++) P[i] = {patch64[4 * i], patch64[4 * i + 1], patch64[4 * i + 2], patch64[4 * i + 3]};
    subpatchV(P, vmin, vmax);
    subpatchU(P, umin, umax);
    for (int i = 0; i < 16; i++)
    {
        patch64[4 * i] = P[i].x;
        patch64[4 * i + 1] = P[i].y;
        patch64[4 * i + 2] = P[i].z;
        patch64[4 * i + 3] = P[i].w;
    }
}

// Slices the 16 (x,y) points of a patch: side 0 = U (rows), 1 = V (columns).
void orc_subpatch2d(float* p2d, int side, float a, float b)
{
    V2<float> P[16];
    for (int i = 0; i < 16; i++) P[i] = {p2d[2 * i], p2d[2 * i + 1]};
    if (side == 0) subpatchU(P, a, b); else subpatchV(P, a, b);
    for (int i = 0; i < 16; i++) { p2d[2 * i] = P[i].x; p2d[2 * i + 1] = P[i].y; }
}

// Matrix-form slice of a 4-point curve with (x,y) points (SplineMat.h:189-206).
void orc_slice_matrix(const float* a_b, float* m16)
{
    M4<float> M = sliceMatrix<float>(a_b[0], a_b[1]);
    for (int c = 0; c < 4; c++)
    {
        m16[4 * c] = M.c[c].x; m16[4 * c + 1] = M.c[c].y; m16[4 * c + 2] = M.c[c].z; m16[4 * c + 3] = M.c[c].w;
    }
}

int orc_x_axis_intersection(const float* A, const float* B, float* out2)
{
    V2<float> r {0, 0};
    bool ok = xAxisIntersection<float>(r, {A[0], A[1]}, {B[0], B[1]});
    out2[0] = r.x; out2[1] = r.y;
    return ok;
}

int orc_triangle_origin(const float* P, const float* A, const float* B, float* out2)
{
    V2<float> c {0, 0};
    bool ok = triangleOriginIntersection<float>(c, {P[0], P[1]}, {A[0], A[1]}, {B[0], B[1]});
    out2[0] = c.x; out2[1] = c.y;
    return ok;
}

// PBR building blocks (rt/PBR.{h,cu}).  kind 0: schlick(f0 rgb, dotLH) → rgb;
// 1: G(dotNL, dotNV, r); 2: D(dotNH, r); 3: microfacet(dotNV, dotNL, dotNH, r);
// 4: lightColor(d, light[12]) → rgb; 5: lightVector(P xyz, light[12]) →
// (ok, d, L xyz).
int orc_pbr(int kind, const float* a, float* out)
{
    auto light = [&](const float* l) {
        return Light {{l[0], l[1], l[2]}, l[3], {l[4], l[5], l[6]}, l[7], {l[8], l[9], l[10]}, l[11]};
    };
    switch (kind)
    {
    case 0: { f3 r = schlick({a[0], a[1], a[2]}, a[3]); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
    case 1: out[0] = Gsmith(a[0], a[1], a[2]); return 0;
    case 2: out[0] = Dggx(a[0], a[1]); return 0;
    case 3: out[0] = microfacet(a[0], a[1], a[2], a[3]); return 0;
    case 4: { f3 r = lightColor(a[0], light(a + 1)); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
    case 5:
    {
        float d = 0;
        f3 L {0, 0, 0};
        bool ok = lightVector(d, L, {a[0], a[1], a[2]}, light(a + 3));
        out[0] = ok ? 1.0f : 0.0f; out[1] = d; out[2] = L.x; out[3] = L.y; out[4] = L.z;
        return 0;
    }
    default: return 1;
    }
}

// The oracle's restatements of the absent Ds library's boundary functions,
// for the tests that pin them.  kind 0: versor(v xyz) -> xyz;
// 1: triangleIntersect(o, d, p0, p1, p2) -> ok, t, b1, b2, b0;
// 2: packSRGB(r, g, b) -> the packed pixel (bits in out[0]);
// 3: boxIntersect(mn, mx, o, d) -> ok, tmin, tmax.
int orc_ds(int kind, const float* a, float* out)
{
    auto v3 = [&](int i) { return f3 {a[i], a[i + 1], a[i + 2]}; };
    switch (kind)
    {
    case 0: { f3 r = versor(v3(0)); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
    case 1:
    {
        f3 uvw {0, 0, 0};
        float t = 0;
        const bool ok = triangleIntersect(v3(0), v3(3), v3(6), v3(9), v3(12), uvw, t);
        out[0] = ok ? 1.0f : 0.0f; out[1] = t; out[2] = uvw.x; out[3] = uvw.y; out[4] = uvw.z;
        return 0;
    }
    case 2: { const uint32_t q = packSRGB(a[0], a[1], a[2]); std::memcpy(out, &q, 4); return 0; }
    case 3:
    {
        Box<float> b;
        b.mn = v3(0);
        b.mx = v3(3);
        float tmin = 0, tmax = 0;
        const bool ok = boxIntersect(b, v3(6), v3(9), tmin, tmax);
        out[0] = ok ? 1.0f : 0.0f; out[1] = tmin; out[2] = tmax;
        return 0;
    }
    }
    return -1;
}

float orc_fminf(float a, float b) { return fmin_(a, b); }
float orc_fmaxf(float a, float b) { return fmax_(a, b); }

// Newton polish of S(u,v) - o - t·d = 0 in double from (u,v,t); returns 1 on
// convergence.  d is normalized first.
int orc_newton(const float* patch, const double* o, const double* d_in, double* uvt, int iters)
{
    V4<double> S[16];
    for (int i = 0; i < 16; i++) S[i] = {patch[4 * i], patch[4 * i + 1], patch[4 * i + 2], patch[4 * i + 3]};
    V3<double> O {o[0], o[1], o[2]}, D = versor(V3<double> {d_in[0], d_in[1], d_in[2]});
    double u = uvt[0], v = uvt[1], t = uvt[2];
    for (int k = 0; k < iters; k++)
    {
        V3<double> F = project(interpolateDC<double>(S, u, v)) - O - t * D;
        // deCasteljauDx returns the hodograph without its factor 3 (Spline.h:183-195)
        V3<double> Su = 3.0 * derivativeU_DC<double>(S, u, v).xyz(), Sv = 3.0 * derivativeV_DC<double>(S, u, v).xyz();
        // J = [Su Sv -D]; solve J·delta = -F (Cramer)
        V3<double> c0 = Su, c1 = Sv, c2 = -D;
        double det = c0.dot(c1.cross(c2));
        if (std::fabs(det) < 1e-300) return 0;
        V3<double> mF = -F;
        double du = mF.dot(c1.cross(c2)) / det;
        double dv = c0.dot(mF.cross(c2)) / det;
        double dt = c0.dot(c1.cross(mF)) / det;
        u += du; v += dv; t += dt;
        if (std::fabs(du) + std::fabs(dv) < 1e-15 && std::fabs(dt) < 1e-13 * std::max(1.0, std::fabs(t))) break;
    }
    uvt[0] = u; uvt[1] = v; uvt[2] = t;
    V3<double> F = project(interpolateDC<double>(S, u, v)) - O - t * D;
    return F.length() < 1e-9 * std::max(1.0, std::fabs(t)) ? 1 : 0;
}

// FP64 clipping of FP64 rays (n × 8: o, tmax, d, pad; d normalized here) —
// the same algorithm as orc_clip_batch's FP64 instantiation without rounding
// the ray to float first.  out: n × 4 (t, u, v, found).
int orc_clip_batch64(const double* rays, const float* patches, uint64_t n, double tol, double* out)
{
    for (uint64_t r = 0; r < n; r++)
    {
        const double* ry = rays + 8 * r;
        const float* pt = patches + 64 * r;
        V3<double> o {ry[0], ry[1], ry[2]}, d {ry[4], ry[5], ry[6]};
        d = versor(d);
        V4<double> S[16];
        for (int i = 0; i < 16; i++) S[i] = {pt[4 * i], pt[4 * i + 1], pt[4 * i + 2], pt[4 * i + 3]};
        PatchHit<double> h {ry[3], {0, 0}};
        bool f = doBezierClipping<double>(h, o, d, S, tol, nullptr);
        out[4 * r] = h.distance;
        out[4 * r + 1] = h.p.x;
        out[4 * r + 2] = h.p.y;
        out[4 * r + 3] = f ? 1.0 : 0.0;
    }
    return 0;
}

// Batched orc_newton from (u,v,t) seeds (n × 3, updated in place), plus the
// cosine |N̂·d̂| of the surface normal with the ray at the polished root
// (grazing-ray flag) and a convergence flag per ray.
int orc_newton_batch(const float* patches, const double* rays, uint64_t n, double* uvt, double* cosnd, int* ok,
                     int iters)
{
    for (uint64_t r = 0; r < n; r++)
    {
        const float* pt = patches + 64 * r;
        const double* ry = rays + 8 * r;
        ok[r] = orc_newton(pt, ry, ry + 4, uvt + 3 * r, iters);
        V4<double> S[16];
        for (int i = 0; i < 16; i++) S[i] = {pt[4 * i], pt[4 * i + 1], pt[4 * i + 2], pt[4 * i + 3]};
        const double u = uvt[3 * r], v = uvt[3 * r + 1];
        V3<double> N = derivativeU_DC<double>(S, u, v).xyz().cross(derivativeV_DC<double>(S, u, v).xyz());
        V3<double> D = versor(V3<double> {ry[4], ry[5], ry[6]});
        const double ln = N.length();
        cosnd[r] = ln > 0 ? std::fabs(N.dot(D)) / ln : 0.0;
    }
    return 0;
}

// --- BVH ---------------------------------------------------------------------
struct orc_bvh
{
    BVHData b;
};

// Node words: lmin xyz, lmax xyz, rmin xyz, rmax xyz (as float bits), left, right, 0, 0.
static void nodeWords(const Node& n, uint32_t* w)
{
    float f[12] = {n.leftBox.mn.x, n.leftBox.mn.y, n.leftBox.mn.z, n.leftBox.mx.x, n.leftBox.mx.y, n.leftBox.mx.z,
                   n.rightBox.mn.x, n.rightBox.mn.y, n.rightBox.mn.z, n.rightBox.mx.x, n.rightBox.mx.y, n.rightBox.mx.z};
    std::memcpy(w, f, 48);
    w[12] = n.left;
    w[13] = n.right;
    w[14] = 0;
    w[15] = 0;
}

// Elements: n × 8 floats (min xyz, max xyz... ) are derived from patches.
orc_bvh* orc_bvh_build_patches(const float* vertices4, const uint32_t* indices, uint32_t npatch)
{
    try
    {
        auto* h = new orc_bvh;
        h->b = bvhBuild(patchElements(vertices4, indices, npatch), 1);
        return h;
    }
    catch (...)
    {
        return nullptr;
    }
}

// Generic elements: n × 10 floats (bmin xyz, bmax xyz, centroid xyz, index as
// u32 bits).
orc_bvh* orc_bvh_build_elements(const float* e10, uint32_t n, uint32_t elementsPerNode)
{
    try
    {
        std::vector<Element> el(n);
        for (uint32_t i = 0; i < n; i++)
        {
            const float* e = e10 + 10 * (size_t)i;
            el[i].bounds.mn = {e[0], e[1], e[2]};
            el[i].bounds.mx = {e[3], e[4], e[5]};
            el[i].centroid = {e[6], e[7], e[8]};
            std::memcpy(&el[i].index, e + 9, 4);
        }
        auto* h = new orc_bvh;
        h->b = bvhBuild(el, elementsPerNode);
        return h;
    }
    catch (...)
    {
        return nullptr;
    }
}

void orc_bvh_sizes(const orc_bvh* h, uint64_t* nnodes, uint64_t* nindices)
{
    *nnodes = h->b.nodes.size();
    *nindices = h->b.indices.size();
}

void orc_bvh_copy(const orc_bvh* h, uint32_t* nodes16, uint32_t* keys, uint32_t* indices)
{
    for (size_t i = 0; i < h->b.nodes.size(); i++) nodeWords(h->b.nodes[i], nodes16 + 16 * i);
    std::memcpy(keys, h->b.keys.data(), 4 * h->b.keys.size());
    std::memcpy(indices, h->b.indices.data(), 4 * h->b.indices.size());
}

void orc_bvh_free(orc_bvh* h) { delete h; }

// --- perfect hash --------------------------------------------------------------
uint32_t orc_phf_displacement_size(uint64_t n)
{
    int s = int(n - 1) >> 1;
    if (s < 2) s = 2;
    return std::bit_floor((unsigned)s);
}

// params out: A, B, mask, tableSize, attempts.  D: mask+1 words.
int orc_phf_build(const uint32_t* keys, uint64_t n, uint32_t seed, uint32_t* params5, uint32_t* D)
{
    PHF h;
    int r = phfBuild(h, std::span<const uint32_t>(keys, n), seed);
    params5[0] = h.A; params5[1] = h.B; params5[2] = h.mask; params5[3] = h.tableSize; params5[4] = (uint32_t)h.attempts;
    if (D && !h.D.empty()) std::memcpy(D, h.D.data(), 4 * h.D.size());
    return r;
}

// The wide family over 64-bit keys (no counterpart in the reference).
int orc_phf_build_wide(const uint64_t* keys, uint64_t n, uint32_t seed, uint32_t* params5, uint32_t* D)
{
    PHF h;
    int r = phfBuild<uint64_t>(h, std::span<const uint64_t>(keys, n), seed, true);
    params5[0] = h.A; params5[1] = h.B; params5[2] = h.mask; params5[3] = h.tableSize; params5[4] = (uint32_t)h.attempts;
    if (D && !h.D.empty()) std::memcpy(D, h.D.data(), 4 * h.D.size());
    return r;
}

// h(key) of the wide family (params: A, B, mask, tableSize) for n keys.
void orc_phf_hash_wide(const uint32_t* params4, const uint32_t* D, const uint64_t* keys, uint64_t n, uint32_t* out)
{
    PHF h;
    h.A = params4[0]; h.B = params4[1]; h.mask = params4[2]; h.tableSize = params4[3];
    h.D.assign(D, D + (size_t)h.mask + 1);
    h.wide = true;
    for (uint64_t i = 0; i < n; i++) out[i] = h.hash(keys[i]);
}

// --- scenes --------------------------------------------------------------------
// Scene description (host arrays), identical in meaning to the product's
// cgs_scene_desc: vertices (x,y,z,w) of all surfaces concatenated, patch
// indices relative to the owning surface's first vertex.
struct orc_scene_desc
{
    const float* vertices;
    const uint32_t* indices;
    const uint32_t* surf_vbase;  // nsurf + 1 (prefix over vertices)
    const uint32_t* surf_pbase;  // nsurf + 1 (prefix over patches)
    uint32_t nsurf;
    const uint32_t* obj_surface; // nobj
    const float* obj_local2world; // nobj × 16, column-major
    const float* obj_world2local; // nobj × 16
    const float* obj_material;    // nobj × 8: diffuse rgb, specular rgb, metalness, roughness
    uint32_t nobj;
    const float* lights;          // nlight × 12: color rgb, strength, position xyz, range, direction xyz, angle
    uint32_t nlight;
    float background[3];
    float ambient[3];
    uint32_t seed;
};

struct orc_scene
{
    Scene s;
};

static M4<float> mat16(const float* m)
{
    M4<float> M;
    for (int c = 0; c < 4; c++) M.c[c] = {m[4 * c], m[4 * c + 1], m[4 * c + 2], m[4 * c + 3]};
    return M;
}

static orc_scene* sceneCreate(const orc_scene_desc* d, int* status, const float* elements10);

orc_scene* orc_scene_create(const orc_scene_desc* d, int* status) { return sceneCreate(d, status, nullptr); }

// The scene with surface 0's BVH over the given elements (10 floats each:
// min xyz, max xyz, centroid xyz, index bits), one per patch of surface 0.
orc_scene* orc_scene_create_elements(const orc_scene_desc* d, const float* elements10, int* status)
{
    return sceneCreate(d, status, elements10);
}

static orc_scene* sceneCreate(const orc_scene_desc* d, int* status, const float* elements10)
{
    auto h = std::make_unique<orc_scene>();
    Scene& s = h->s;
    try
    {
        s.surfaces.resize(d->nsurf);
        for (uint32_t k = 0; k < d->nsurf; k++)
        {
            Surface& sf = s.surfaces[k];
            uint32_t v0 = d->surf_vbase[k], v1 = d->surf_vbase[k + 1];
            uint32_t p0 = d->surf_pbase[k], p1 = d->surf_pbase[k + 1];
            sf.vertices.resize(v1 - v0);
            for (uint32_t i = v0; i < v1; i++)
                sf.vertices[i - v0] = {d->vertices[4 * (size_t)i], d->vertices[4 * (size_t)i + 1],
                                       d->vertices[4 * (size_t)i + 2], d->vertices[4 * (size_t)i + 3]};
            sf.indices.assign(d->indices + 16 * (size_t)p0, d->indices + 16 * (size_t)p1);
        }
        s.objects.resize(d->nobj);
        for (uint32_t i = 0; i < d->nobj; i++)
        {
            Object& o = s.objects[i];
            o.surface = d->obj_surface[i];
            o.local2world = mat16(d->obj_local2world + 16 * (size_t)i);
            o.world2local = mat16(d->obj_world2local + 16 * (size_t)i);
            const float* m = d->obj_material + 8 * (size_t)i;
            o.material = {{m[0], m[1], m[2]}, {m[3], m[4], m[5]}, m[6], m[7]};
        }
        s.lights.resize(d->nlight);
        for (uint32_t i = 0; i < d->nlight; i++)
        {
            const float* l = d->lights + 12 * (size_t)i;
            s.lights[i] = {{l[0], l[1], l[2]}, l[3], {l[4], l[5], l[6]}, l[7], {l[8], l[9], l[10]}, l[11]};
        }
        s.background = {d->background[0], d->background[1], d->background[2]};
        s.ambient = {d->ambient[0], d->ambient[1], d->ambient[2]};
        std::vector<Element> el;
        if (elements10)
        {
            el.resize(d->surf_pbase[1] - d->surf_pbase[0]);
            for (size_t i = 0; i < el.size(); i++)
            {
                const float* q = elements10 + 10 * i;
                el[i].bounds.mn = {q[0], q[1], q[2]};
                el[i].bounds.mx = {q[3], q[4], q[5]};
                el[i].centroid = {q[6], q[7], q[8]};
                std::memcpy(&el[i].index, q + 9, 4);
            }
        }
        int r = buildScene(s, d->seed, elements10 ? &el : nullptr);
        *status = r;
        if (r) return nullptr;
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "orc_scene_create: %s\n", e.what());
        *status = 100;
        return nullptr;
    }
    return h.release();
}

void orc_scene_destroy(orc_scene* h) { delete h; }

// surface >= 0: that surface's BVH/hash; -1: the top-level BVH.
int orc_scene_bvh_sizes(const orc_scene* h, int surface, uint64_t* nnodes, uint64_t* nindices, uint64_t* ndispl, uint64_t* ntable)
{
    const BVHData& b = surface < 0 ? h->s.top : h->s.surfaces.at(surface).bvh;
    const PHF& p = surface < 0 ? h->s.topHash : h->s.surfaces.at(surface).hash;
    *nnodes = b.nodes.size();
    *nindices = b.indices.size();
    *ndispl = p.D.size();
    *ntable = p.tableSize;
    return 0;
}

// cg::PatchBVH's queries (PatchBVH.cpp:84-121 under BVHBase::intersect) on
// surface 0 of a scene: the ray is a Ray3f (its direction normalized once);
// closest: the hit distance starts at ray.max; any: each leaf clips from
// distance infinity.  hits3 (t,u,v) / patch, or occluded (1/0).
int orc_scene_pick(const orc_scene* h, const float* rays, uint64_t n, int any, float* hits3, uint32_t* patch,
    uint32_t* occluded, int threads)
{
    const Surface& s = h->s.surfaces.at(0);
    parallelFor(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t r = a; r < b; r++)
        {
            const float* ry = rays + 8 * r;
            const Ray ray {{ry[0], ry[1], ry[2]}, versor(V3<float> {ry[4], ry[5], ry[6]}), ry[3]};
            if (any)
            {
                occluded[r] = hashIntersectAny(s.view(), ray, [&](uint32_t idx) -> bool {
                    PatchHit<float> ph {std::numeric_limits<float>::infinity(), {0, 0}};
                    f4 S[16];
                    s.patch(idx, S);
                    return doBezierClipping<float>(ph, ray.origin, ray.direction, S, 0x1p-12f, nullptr);
                }, nullptr) ? 1u : 0u;
                continue;
            }
            Hit hit;
            hit.t = ray.max;
            uint32_t pid = EMPTY;
            hashIntersect(s.view(), hit, ray, [&](Hit& hh, uint32_t idx) -> bool {
                PatchHit<float> ph {hh.t, {hh.coordinates.x, hh.coordinates.y}};
                f4 S[16];
                s.patch(idx, S);
                if (!doBezierClipping<float>(ph, ray.origin, ray.direction, S, 0x1p-12f, nullptr)) return false;
                hh.index = idx;
                hh.t = ph.distance;
                hh.coordinates = {ph.p.x, ph.p.y, 0.0f};
                pid = idx;
                return true;
            }, nullptr);
            hits3[3 * r] = hit.t;
            hits3[3 * r + 1] = pid == EMPTY ? 0.0f : hit.coordinates.x;
            hits3[3 * r + 2] = pid == EMPTY ? 0.0f : hit.coordinates.y;
            patch[r] = pid;
        }
    });
    return 0;
}

// 1 when that BVH is hashed with the wide family.
int orc_scene_bvh_wide(const orc_scene* h, int surface)
{
    return (surface < 0 ? h->s.topHash : h->s.surfaces.at(surface).hash).wide ? 1 : 0;
}

int orc_scene_bvh_copy(const orc_scene* h, int surface, uint32_t* nodes16, uint32_t* keys, uint32_t* indices,
    uint32_t* params5, uint32_t* D, uint32_t* table)
{
    const BVHData& b = surface < 0 ? h->s.top : h->s.surfaces.at(surface).bvh;
    const PHF& p = surface < 0 ? h->s.topHash : h->s.surfaces.at(surface).hash;
    const auto& t = surface < 0 ? h->s.topTable : h->s.surfaces.at(surface).table;
    for (size_t i = 0; i < b.nodes.size(); i++) nodeWords(b.nodes[i], nodes16 + 16 * i);
    if (keys) std::memcpy(keys, b.keys.data(), 4 * b.keys.size());
    if (indices) std::memcpy(indices, b.indices.data(), 4 * b.indices.size());
    if (params5)
    {
        params5[0] = p.A; params5[1] = p.B; params5[2] = p.mask; params5[3] = p.tableSize; params5[4] = (uint32_t)p.attempts;
    }
    if (D) std::memcpy(D, p.D.data(), 4 * p.D.size());
    if (table) std::memcpy(table, t.data(), 4 * t.size());
    return 0;
}

// World-space rays (n × 8: o, tmax, d, pad; d is used as given, as the
// RayTracer's Ray).  any = 0: closest hit → hits (t,u,v) + patch + object;
// any = 1: occlusion → object[r] = 1 when occluded; any = 2: closest hit with
// the explicit-stack traversal of rt/BVH.cpp:126-189.
// Leaf method of a trace (cgs_trace_options's meaning): method 0 clipping,
// 1 subdivision; clip tolerance; subdivision threshold; clip pass cap (0:
// none); Newton polish of closest hits.
struct orc_trace_opts
{
    int method;
    float tol;
    float threshold;
    uint32_t maxIters;
    int polish;
};

int orc_scene_trace_ex(const orc_scene* h, const float* rays, uint64_t n, int any, float* hits3,
    uint32_t* patch, int32_t* object, orc_work* work, int threads, const orc_trace_opts* opts)
{
    TraceOpts op;
    if (opts)
    {
        op.method = opts->method;
        op.tol = opts->tol;
        op.threshold = opts->threshold;
        op.maxIters = opts->maxIters;
        op.polish = opts->polish != 0;
    }
    std::mutex mu;
    Work total;
    parallelFor(n, threads, [&](uint64_t a, uint64_t b) {
        Work w;
        for (uint64_t r = a; r < b; r++)
        {
            const float* ry = rays + 8 * r;
            Ray ray {{ry[0], ry[1], ry[2]}, {ry[4], ry[5], ry[6]}, ry[3]};
            if (any == 1)
            {
                object[r] = sceneIntersectAny(h->s, ray, &w, op) ? 1 : 0;
                continue;
            }
            Hit hit;
            hit.t = ray.max;
            uint32_t pid = EMPTY;
            const uint64_t overflowsBefore = w.overflows;
            int o = any == 2 ? sceneIntersectStack(h->s, hit, ray, &pid) : sceneIntersect(h->s, hit, ray, &pid, &w, op);
            if (op.polish && o >= 0 && w.overflows == overflowsBefore)
            {
                const Object& ob = h->s.objects[o];
                f4 S[16];
                h->s.surfaces[ob.surface].patch(pid, S);
                polishHit64(S, ob.world2local, ray.origin, ray.direction, hit.t, hit.coordinates.x, hit.coordinates.y);
            }
            hits3[3 * r] = o < 0 ? ray.max : hit.t;
            hits3[3 * r + 1] = o < 0 ? 0.0f : hit.coordinates.x;
            hits3[3 * r + 2] = o < 0 ? 0.0f : hit.coordinates.y;
            patch[r] = o < 0 ? EMPTY : pid;
            object[r] = o;
        }
        std::lock_guard lk(mu);
        total.add(w);
    });
    putWork(work, total);
    return 0;
}

int orc_scene_trace(const orc_scene* h, const float* rays, uint64_t n, int any, float* hits3,
    uint32_t* patch, int32_t* object, orc_work* work, int threads)
{
    return orc_scene_trace_ex(h, rays, n, any, hits3, patch, object, work, threads, nullptr);
}

// Closest-hit traces of n world rays, recording every patch test in call
// order (single thread): up to cap tests into rays_out (8 floats each: local
// o, initial best t, local d, 0) and patches_out (64 floats each).  Returns
// the number of tests.
uint64_t orc_scene_record_tests(const orc_scene* h, const float* rays, uint64_t n, float* rays_out,
                                float* patches_out, uint64_t cap)
{
    PatchTestLog log;
    g_patchLog = &log;
    for (uint64_t r = 0; r < n; r++)
    {
        const float* ry = rays + 8 * r;
        Ray ray {{ry[0], ry[1], ry[2]}, {ry[4], ry[5], ry[6]}, ry[3]};
        Hit hit;
        hit.t = ray.max;
        uint32_t pid;
        sceneIntersect(h->s, hit, ray, &pid, nullptr);
    }
    g_patchLog = nullptr;
    const uint64_t k = log.rays.size() / 8, m = std::min(k, cap);
    std::memcpy(rays_out, log.rays.data(), m * 32);
    std::memcpy(patches_out, log.patches.data(), m * 256);
    return k;
}

struct orc_camera
{
    float toWorld[16];
    float position[3];
    float fovy;
    int flipY;
};

// Context::pixelRayDirection (RayTracer.cu:59-71) after the render's camera
// setup (RayTracer.cu:141-163) for a w x h frame, at n fractional pixel
// coordinates xy (2 floats each); dirs3: n unit directions.
int orc_pixel_rays(const orc_camera* c, uint32_t w, uint32_t hgt, const float* xy, uint64_t n, float* dirs3)
{
    Camera cam {mat16(c->toWorld), {c->position[0], c->position[1], c->position[2]}, c->fovy, c->flipY != 0};
    CameraSetup cs = setupCamera(cam, w, hgt);
    for (uint64_t k = 0; k < n; k++)
    {
        f3 d = pixelRayDirection(cs, xy[2 * k], xy[2 * k + 1]);
        dirs3[3 * k] = d.x;
        dirs3[3 * k + 1] = d.y;
        dirs3[3 * k + 2] = d.z;
    }
    return 0;
}

// The GPU renderer's semantics (RayTracer.cu), one pixel per iteration.
// color: w*h*3 floats (unclamped), frame: w*h packed pixels (stride w).
int orc_scene_render(const orc_scene* h, const orc_camera* c, float eps, int depth, uint32_t w, uint32_t hgt,
    float* color, uint32_t* frame, orc_work* work, uint64_t* raycounts3, int threads, uint32_t row0, uint32_t row1)
{
    Camera cam {mat16(c->toWorld), {c->position[0], c->position[1], c->position[2]}, c->fovy, c->flipY != 0};
    CameraSetup cs = setupCamera(cam, w, hgt);
    RenderOptions op {eps, depth};
    row1 = std::min(row1, hgt);
    std::mutex mu;
    Work total;
    RayCounts rc;
    uint64_t npx = (uint64_t)w * (row1 - row0);
    parallelFor(npx, threads, [&](uint64_t a, uint64_t b) {
        Work wk;
        RayCounts r;
        for (uint64_t k = a; k < b; k++)
        {
            uint32_t i = (uint32_t)(k % w), j = row0 + (uint32_t)(k / w);
            f3 col;
            uint32_t px;
            renderPixel(h->s, cs, op, i, j, col, px, &wk, &r);
            size_t q = (size_t)j * w + i;
            if (color) { color[3 * q] = col.x; color[3 * q + 1] = col.y; color[3 * q + 2] = col.z; }
            if (frame) frame[q] = px;
        }
        std::lock_guard lk(mu);
        total.add(wk);
        rc.primary += r.primary; rc.shadow += r.shadow; rc.secondary += r.secondary;
    });
    putWork(work, total);
    if (raycounts3) { raycounts3[0] = rc.primary; raycounts3[1] = rc.shadow; raycounts3[2] = rc.secondary; }
    return 0;
}

// The render's trace loop from n given world rays (8 floats each, as
// orc_scene_trace): color3 unclamped colours, pixel packed sRGB (nullable),
// hits3 / patch / object the first segment's closest hit (nullable).
int orc_scene_shade(const orc_scene* h, const float* rays, uint64_t n, float eps, int depth, float* color3,
    uint32_t* pixel, float* hits3, uint32_t* patch, int32_t* object, uint64_t* raycounts3, int threads)
{
    RenderOptions op {eps, depth};
    std::mutex mu;
    RayCounts rc;
    parallelFor(n, threads, [&](uint64_t a, uint64_t b) {
        RayCounts r;
        for (uint64_t k = a; k < b; k++)
        {
            const float* ry = rays + 8 * k;
            const Ray ray {{ry[0], ry[1], ry[2]}, {ry[4], ry[5], ry[6]}, ry[3]};
            FirstHit fh;
            const f3 col = shadeRay(h->s, ray, op, nullptr, &r, &fh);
            color3[3 * k] = col.x;
            color3[3 * k + 1] = col.y;
            color3[3 * k + 2] = col.z;
            if (pixel) pixel[k] = packColor(col);
            if (hits3)
            {
                hits3[3 * k] = fh.t;
                hits3[3 * k + 1] = fh.u;
                hits3[3 * k + 2] = fh.v;
            }
            if (patch) patch[k] = fh.patch;
            if (object) object[k] = fh.object;
        }
        std::lock_guard lk(mu);
        rc.primary += r.primary; rc.shadow += r.shadow; rc.secondary += r.secondary;
    });
    if (raycounts3) { raycounts3[0] = rc.primary; raycounts3[1] = rc.shadow; raycounts3[2] = rc.secondary; }
    return 0;
}

// The sampled render (USE_MONTECARLO_SAMPLING): rows [row0, row1) as
// orc_scene_render, `samples` camera rays per pixel, draws from `seed`.
int orc_scene_render_sampled(const orc_scene* h, const orc_camera* c, float eps, int depth, uint32_t samples,
    uint32_t seed, uint32_t w, uint32_t hgt, float* color, uint32_t* frame, uint64_t* raycounts3, int threads,
    uint32_t row0, uint32_t row1)
{
    if (samples == 0) return 1;
    Camera cam {mat16(c->toWorld), {c->position[0], c->position[1], c->position[2]}, c->fovy, c->flipY != 0};
    CameraSetup cs = setupCamera(cam, w, hgt);
    RenderOptions op {eps, depth};
    row1 = std::min(row1, hgt);
    std::mutex mu;
    RayCounts rc;
    uint64_t npx = (uint64_t)w * (row1 - row0);
    parallelFor(npx, threads, [&](uint64_t a, uint64_t b) {
        RayCounts r;
        for (uint64_t k = a; k < b; k++)
        {
            uint32_t i = (uint32_t)(k % w), j = row0 + (uint32_t)(k / w);
            f3 col;
            uint32_t px;
            renderPixelSampled(h->s, cs, op, samples, seed, i, j, col, px, nullptr, &r);
            size_t q = (size_t)j * w + i;
            if (color) { color[3 * q] = col.x; color[3 * q + 1] = col.y; color[3 * q + 2] = col.z; }
            if (frame) frame[q] = px;
        }
        std::lock_guard lk(mu);
        rc.primary += r.primary; rc.shadow += r.shadow; rc.secondary += r.secondary;
    });
    if (raycounts3) { raycounts3[0] = rc.primary; raycounts3[1] = rc.shadow; raycounts3[2] = rc.secondary; }
    return 0;
}

// The sampling pieces, for their pins: Philox4x32-10 blocks (4 counter words
// and 2 key words per block in, 4 words out), draws as floats, vdc, the
// angle functions, and one sampled bounce (N, V: 3 floats each; material:
// diffuse3, specular3, metalness, roughness; out: R3, brdf3, 1 or 0).
void orc_philox(const uint32_t* ctr4, const uint32_t* key2, uint64_t n, uint32_t* out4)
{
    for (uint64_t k = 0; k < n; k++) philox4x32_10(ctr4 + 4 * k, key2 + 2 * k, out4 + 4 * k);
}
void orc_sample_words(uint32_t seed, uint32_t stream, uint32_t first, uint64_t n, uint32_t* words, float* units)
{
    for (uint64_t k = 0; k < n; k++)
    {
        const uint32_t wv = sampleWord(seed, stream, first + (uint32_t)k);
        if (words) words[k] = wv;
        if (units) units[k] = unitFloat(wv);
    }
}
void orc_vdc(const uint32_t* i, uint64_t n, float* out)
{
    for (uint64_t k = 0; k < n; k++) out[k] = vdc(i[k]);
}
void orc_sincos(const float* a, uint64_t n, float* s, float* c)
{
    for (uint64_t k = 0; k < n; k++) sinCosAngle(a[k], s[k], c[k]);
}
void orc_atan2q(const float* y, const float* x, uint64_t n, float* out)
{
    for (uint64_t k = 0; k < n; k++) out[k] = atan2Quadrant(y[k], x[k]);
}
void orc_sampled_bounce(const float* N3, const float* V3, const float* material8, const float* xy, uint64_t n,
    float* out7)
{
    for (uint64_t k = 0; k < n; k++)
    {
        const float* Np = N3 + 3 * k;
        const float* Vp = V3 + 3 * k;
        const float* mp = material8 + 8 * k;
        Material m;
        m.diffuse = {mp[0], mp[1], mp[2]};
        m.specular = {mp[3], mp[4], mp[5]};
        m.metalness = mp[6];
        m.roughness = mp[7];
        const f3 N {Np[0], Np[1], Np[2]}, V {Vp[0], Vp[1], Vp[2]};
        f3 R {0, 0, 0}, brdf {0, 0, 0};
        const bool ok = sampledBounce(N, V, N.dot(V), m, xy[2 * k], xy[2 * k + 1], R, brdf);
        float* o = out7 + 7 * k;
        o[0] = R.x; o[1] = R.y; o[2] = R.z;
        o[3] = brdf.x; o[4] = brdf.y; o[5] = brdf.z;
        o[6] = ok ? 1.0f : 0.0f;
    }
}

// CPURayTracer restatement (the CPU baseline): renders rows [row0,row1) with
// `threads` workers over 32×32 tiles (rows row0, row0+stride, ...).
// Returns wall seconds; ray counts out.
double orc_cpu_render(const orc_scene* h, const orc_camera* c, float eps, int depth, uint32_t w, uint32_t hgt,
    float* color, uint32_t* frame, int threads, uint32_t row0, uint32_t row1, uint32_t stride, uint64_t* raycounts2)
{
    CPURayTracer rt;
    rt.options.eps = eps;
    rt.options.recursionDepth = depth;
    rt.options.threads = threads;
    rt.options.flipYAxis = c->flipY != 0;
    Camera cam {mat16(c->toWorld), {c->position[0], c->position[1], c->position[2]}, c->fovy, c->flipY != 0};
    double sec = rt.render(&h->s, cam, w, hgt, color, frame, row0, row1, stride);
    if (raycounts2) { raycounts2[0] = rt.rays.load(); raycounts2[1] = rt.shadowRays.load(); }
    return sec;
}

} // extern "C"
