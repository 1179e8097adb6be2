// SPDX-License-Identifier: MIT
//
// cgspline oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of MachSilva/cgspline's ray–bicubic-patch path, written to
// check the HIP product path against.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; the product library never links it.
//
// Every function follows the reference's host semantics and cites the file:line
// it restates (paths relative to the reference repository root).  Everything is
// templated on the scalar type so the same algorithm can run in FP32 (parity,
// bit-exact against the GPU) and FP64 (accuracy reference).  Build with
// -ffp-contract=off: the reference's host path is compiled without FMA
// contraction (the .cu host TUs do not get -march=native, CMakeLists.txt:83).
//
// Ds boundary.  The reference's vector/bounds/ray types come from the absent
// Ds submodule.  This file defines them from first principles (see the
// "Ds-boundary" comments), and tests/test_oracle_ds.py pins them by
// mathematics: versor, the slab test, Möller–Trumbore, the sRGB encode.
// Parity unpinned: the empty box's convention (+max, -max), which only shows
// where a box stays empty (the top-level element of an object without
// geometry: centroid 0 here, NaN under ±infinity).
//
// Planted mistakes.  CGS_ORACLE_PLANT selects one deliberate error (0, the
// default and the only value the Makefile builds, selects none).  The tests
// build copies with a nonzero value to show that the mathematical pins of
// the oracle fail on a plausible mistake (tests/planted.py):
//   1 subdivision reports a region's corner instead of its centre
//   2 subdivision compares a region's sides with <= instead of < threshold
//   3 camera: tan(fovy) instead of tan(fovy / 2)
//   4 camera: pixel centres at i + 1/2 instead of the integer i
//   5 shading: the factor pi on a hit's colour dropped
//   6 shading: the reflection attenuation without (1 - roughness)^2
//   7 scene: an object's local t not rescaled by 1 / |M d|
//   8 polish: the residual S - o + t d instead of S - o - t d
//   9 polish: the u derivative without the quotient rule's -p w_u term
//  10 sampled render: sample s at (i + s/n, j + s/n) instead of (i + s/n, j + vdc(s))
//  11 sampled bounce: the microfacet distribution's alpha = roughness instead of roughness^2
//  12 sampled bounce: the weight G without the factor (M.V) / (N.V N.M)
//
#pragma once

#ifndef CGS_ORACLE_PLANT
#define CGS_ORACLE_PLANT 0
#endif

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <span>
#include <stdexcept>
#include <vector>

namespace orc
{

// ---------------------------------------------------------------------------
// Scalar helpers
// ---------------------------------------------------------------------------

// glibc x86-64 fminf/fmaxf semantics (minss/maxss + NaN handling): on ties
// the second operand is returned; a NaN operand yields the other operand.
// The reference calls fmin/fmax on its host path (Spline_BezierClipping.cu:
// 151,290-291,336-337); these helpers reproduce that behaviour exactly,
// including signed zeros.  Pinned against libm in tests/test_oracle_math.py.
template<class T> inline T fmin_(T a, T b) { return (a < b || b != b) ? a : b; }
template<class T> inline T fmax_(T a, T b) { return (a > b || b != b) ? a : b; }

template<class T> struct C
{
    static constexpr T eps = std::numeric_limits<T>::epsilon();
};

// ---------------------------------------------------------------------------
// Vector types (Ds-boundary: Vector<float,N> of the Ds library)
// ---------------------------------------------------------------------------

template<class T> struct V2
{
    T x, y;
    V2 operator+(V2 b) const { return {x + b.x, y + b.y}; }
    V2 operator-(V2 b) const { return {x - b.x, y - b.y}; }
    V2 operator-() const { return {-x, -y}; }
    V2 operator*(V2 b) const { return {x * b.x, y * b.y}; } // component-wise
    T dot(V2 b) const { return x * b.x + y * b.y; }
    T length() const { return std::sqrt(x * x + y * y); }
};
template<class T> inline V2<T> operator*(T s, V2<T> v) { return {s * v.x, s * v.y}; }

template<class T> struct V3
{
    T x, y, z;
    T& operator[](int i) { return (&x)[i]; }
    T operator[](int i) const { return (&x)[i]; }
    V3 operator+(V3 b) const { return {x + b.x, y + b.y, z + b.z}; }
    V3 operator-(V3 b) const { return {x - b.x, y - b.y, z - b.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 operator*(V3 b) const { return {x * b.x, y * b.y, z * b.z}; }
    T dot(V3 b) const { return x * b.x + y * b.y + z * b.z; }
    V3 cross(V3 b) const
    {
        return {y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x};
    }
    T squaredNorm() const { return x * x + y * y + z * z; }
    T length() const { return std::sqrt(squaredNorm()); }
    T max() const { return fmax_(x, fmax_(y, z)); }
    V3 inverse() const { return {T(1) / x, T(1) / y, T(1) / z}; }
};
template<class T> inline V3<T> operator*(T s, V3<T> v) { return {s * v.x, s * v.y, s * v.z}; }

template<class T> struct V4
{
    T x, y, z, w;
    V4 operator+(V4 b) const { return {x + b.x, y + b.y, z + b.z, w + b.w}; }
    V4 operator-(V4 b) const { return {x - b.x, y - b.y, z - b.z, w - b.w}; }
    V3<T> xyz() const { return {x, y, z}; }
};
template<class T> inline V4<T> operator*(T s, V4<T> v) { return {s * v.x, s * v.y, s * v.z, s * v.w}; }

// Ds-boundary: Vector::normalize() scales by the reciprocal length unless the
// length is within machine epsilon of zero, in which case the vector is left
// unchanged.  versor() is the normalized copy.
template<class T> inline V2<T> normalized(V2<T> v)
{
    T len = v.length();
    if (!(std::fabs(len) <= C<T>::eps))
    {
        T inv = T(1) / len;
        v = {v.x * inv, v.y * inv};
    }
    return v;
}
template<class T> inline V3<T> versor(V3<T> v)
{
    T len = v.length();
    if (!(std::fabs(len) <= C<T>::eps))
    {
        T inv = T(1) / len;
        v = {v.x * inv, v.y * inv, v.z * inv};
    }
    return v;
}

// AlignedTypes.h:63-66
template<class T> inline V3<T> project(V4<T> p) { return {p.x / p.w, p.y / p.w, p.z / p.w}; }

// 4x4 column-major matrix; AlignedTypes.h:156-216
template<class T> struct M4
{
    V4<T> c[4];
    // AlignedTypes.h:185-188: v.x*c0 + v.y*c1 + v.z*c2 + v.w*c3
    V4<T> operator*(V4<T> v) const
    {
        return ((v.x * c[0] + v.y * c[1]) + v.z * c[2]) + v.w * c[3];
    }
};

// 3x3 column-major (AlignedTypes.h:83-154)
template<class T> struct M3
{
    V3<T> c[3];
    static M3 from4(const M4<T>& m)
    {
        return {{m.c[0].xyz(), m.c[1].xyz(), m.c[2].xyz()}};
    }
    V3<T> operator*(V3<T> v) const { return (v.x * c[0] + v.y * c[1]) + v.z * c[2]; }
    M3 transposed() const
    {
        return {{{c[0].x, c[1].x, c[2].x}, {c[0].y, c[1].y, c[2].y}, {c[0].z, c[1].z, c[2].z}}};
    }
};

// Ds-boundary: Bounds3<float>.  Empty box = (+max, -max); inflate by point is
// component-wise min/max.
template<class T> struct Box
{
    V3<T> mn {std::numeric_limits<T>::max(), std::numeric_limits<T>::max(), std::numeric_limits<T>::max()};
    V3<T> mx {-std::numeric_limits<T>::max(), -std::numeric_limits<T>::max(), -std::numeric_limits<T>::max()};
    void inflate(V3<T> p)
    {
        for (int i = 0; i < 3; i++)
        {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
    void inflate(const Box& b) { inflate(b.mn); inflate(b.mx); }
    V3<T> size() const { return mx - mn; }
    V3<T> center() const { return T(0.5) * (mn + mx); }
};

// ---------------------------------------------------------------------------
// Instrumentation: algorithmic work counters (per call, summed by callers)
// ---------------------------------------------------------------------------
struct Work
{
    uint64_t nodes = 0;       // inner nodes visited (two child boxes tested)
    uint64_t leaves = 0;      // leaves visited
    uint64_t hashProbes = 0;  // sibling lookups through the perfect hash
    uint64_t patchTests = 0;  // doBezierClipping calls
    uint64_t clipIters = 0;   // clip iterations (hull computed)
    uint64_t subdivs = 0;     // subdivisions (delta > 0.8)
    uint64_t prunes = 0;      // prunes (delta < 0)
    uint64_t converged = 0;   // regions under tolerance
    uint64_t onHits = 0;      // onHit evaluations
    uint64_t maxStack = 0;    // deepest clipping stack seen
    uint64_t objects = 0;     // top-level leaf element (object) tests
    uint64_t maxPostponed = 0; // most siblings postponed at once in one traversal
    uint64_t overflows = 0;   // clips stopped by the optional stack cap (clipStackCap)
    uint64_t topNodes = 0;    // the part of `nodes` visited in scenes' top-level BVHs
    void add(const Work& o)
    {
        overflows += o.overflows;
        topNodes += o.topNodes;
        objects += o.objects;
        maxPostponed = std::max(maxPostponed, o.maxPostponed);
        nodes += o.nodes; leaves += o.leaves; hashProbes += o.hashProbes;
        patchTests += o.patchTests; clipIters += o.clipIters; subdivs += o.subdivs;
        prunes += o.prunes; converged += o.converged; onHits += o.onHits;
        maxStack = std::max(maxStack, o.maxStack);
    }
};

// ---------------------------------------------------------------------------
// Curves and patches (Spline.h, SplineMat.h)
// ---------------------------------------------------------------------------

// Spline.h:324-364 — restriction of a cubic to [u,v]: split1 at u, then
// split0 at (v-u)/(1-u).  `stride` walks a row (1) or a column (4).
template<class T, class P>
inline void slice(P* curve, int stride, T u, T v)
{
    P& C0 = curve[0];
    P& C1 = curve[stride];
    P& C2 = curve[2 * stride];
    P& C3 = curve[3 * stride];
    T iu = T(1) - u;
    T vv = (v - u) / iu;
    T ivv = T(1) - vv;
    P P5, P6, P8;
    P6 = u * C1 + iu * C0;
    P5 = u * C2 + iu * C1;
    C2 = u * C3 + iu * C2;
    P8 = u * P5 + iu * P6;
    C1 = u * C2 + iu * P5;
    C0 = u * C1 + iu * P8;

    P6 = ivv * C2 + vv * C3;
    P5 = ivv * C1 + vv * C2;
    C1 = ivv * C0 + vv * C1;
    P8 = ivv * P5 + vv * P6;
    C2 = ivv * C1 + vv * P5;
    C3 = ivv * C2 + vv * P8;
}

// Spline.h:366-382: point(i,j) = [4j+i]; subpatchU slices each row j.
template<class T, class P> inline void subpatchU(P* p, T a, T b)
{
    for (int j = 0; j < 4; j++) slice(p + 4 * j, 1, a, b);
}
template<class T, class P> inline void subpatchV(P* p, T a, T b)
{
    for (int i = 0; i < 4; i++) slice(p + i, 4, a, b);
}

// Spline.h:168-178 — de Casteljau (NURBS Book p.24).
template<class T, class P> inline P deCasteljau(P p[4], T u)
{
    const T t = T(1) - u;
    for (int k = 1; k < 4; ++k)
        for (int i = 0; i < 4 - k; ++i)
            p[i] = t * p[i] + u * p[i + 1];
    return p[0];
}
// Spline.h:183-195
template<class T, class P> inline P deCasteljauDx(P p[4], T u)
{
    const T t = T(1) - u;
    for (int i = 1; i < 4; ++i) p[i - 1] = p[i] - p[i - 1];
    for (int k = 2; k < 4; ++k)
        for (int i = 0; i < 4 - k; ++i)
            p[i] = t * p[i] + u * p[i + 1];
    return p[0];
}
// Spline.h:197-214
template<class T> inline V4<T> interpolateDC(const V4<T> s[16], T u, T v)
{
    V4<T> p[4], q[4];
    for (int j = 0; j < 4; j++)
    {
        for (int i = 0; i < 4; i++) p[i] = s[4 * j + i];
        q[j] = deCasteljau(p, u);
    }
    return deCasteljau(q, v);
}
// Spline.h:216-234
template<class T> inline V4<T> derivativeU_DC(const V4<T> s[16], T u, T v)
{
    V4<T> p[4], q[4];
    for (int j = 0; j < 4; j++)
    {
        for (int i = 0; i < 4; i++) p[i] = s[4 * j + i];
        q[j] = deCasteljauDx(p, u);
    }
    return deCasteljau(q, v);
}
// Spline.h:236-254
template<class T> inline V4<T> derivativeV_DC(const V4<T> s[16], T u, T v)
{
    V4<T> p[4], q[4];
    for (int i = 0; i < 4; i++)
    {
        for (int j = 0; j < 4; j++) q[j] = s[4 * j + i];
        p[i] = deCasteljauDx(q, v);
    }
    return deCasteljau(p, u);
}
// Spline.cu:10-27 — normal with the degenerate-derivative fix (eps = 1e-4).
template<class T> inline V3<T> normalDC(const V4<T> s[16], T u, T v)
{
    const T e = T(0.0001f);
    V4<T> dU = derivativeU_DC(s, u, v);
    V4<T> dV = derivativeV_DC(s, u, v);
    if (dU.x == 0 && dU.y == 0 && dU.z == 0) dU = derivativeU_DC(s, u, std::fabs(v - e));
    if (dV.x == 0 && dV.y == 0 && dV.z == 0) dV = derivativeV_DC(s, std::fabs(u - e), v);
    return dU.xyz().cross(dV.xyz());
}

// SplineMat.h:22-44 — power-basis matrices (columns).
template<class T> inline M4<T> cubicBase()
{
    return {{{1, 0, 0, 0}, {-3, 3, 0, 0}, {3, -6, 3, 0}, {-1, 3, -3, 1}}};
}
template<class T> inline M4<T> cubicDerivativeBase()
{
    return {{{-3, 3, 0, 0}, {6, -12, 6, 0}, {-3, 9, -9, 3}, {0, 0, 0, 0}}};
}
// SplineMat.h:101-110
template<class T> inline V4<T> evalVec(T t) { T t2 = t * t; return {1, t, t2, t2 * t}; }
template<class T> inline V4<T> evalVecDt(T t) { return {1, t, t * t, 0}; }

template<class T> inline M4<T> cols(V4<T> a, V4<T> b, V4<T> c, V4<T> d) { return {{a, b, c, d}}; }

// SplineMat.h:125-140 — mat::interpolate (used by doBezierClipping's onHit).
template<class T> inline V4<T> interpolateMat(const V4<T> s[16], T u, T v)
{
    const M4<T> B = cubicBase<T>();
    V4<T> x = B * evalVec(u);
    V4<T> y = B * evalVec(v);
    M4<T> A = cols(cols(s[0], s[1], s[2], s[3]) * x,
                   cols(s[4], s[5], s[6], s[7]) * x,
                   cols(s[8], s[9], s[10], s[11]) * x,
                   cols(s[12], s[13], s[14], s[15]) * x);
    return A * y;
}
// SplineMat.h:142-158
template<class T> inline V4<T> derivativeUMat(const V4<T> s[16], T u, T v)
{
    const M4<T> B = cubicBase<T>();
    const M4<T> Cd = cubicDerivativeBase<T>();
    V4<T> x = Cd * evalVecDt(u);
    V4<T> y = B * evalVec(v);
    M4<T> A = cols(cols(s[0], s[1], s[2], s[3]) * x,
                   cols(s[4], s[5], s[6], s[7]) * x,
                   cols(s[8], s[9], s[10], s[11]) * x,
                   cols(s[12], s[13], s[14], s[15]) * x);
    return A * y;
}
// SplineMat.h:160-176
template<class T> inline V4<T> derivativeVMat(const V4<T> s[16], T u, T v)
{
    const M4<T> B = cubicBase<T>();
    const M4<T> Cd = cubicDerivativeBase<T>();
    V4<T> x = B * evalVec(u);
    V4<T> y = Cd * evalVecDt(v);
    M4<T> A = cols(cols(s[0], s[4], s[8], s[12]) * y,
                   cols(s[1], s[5], s[9], s[13]) * y,
                   cols(s[2], s[6], s[10], s[14]) * y,
                   cols(s[3], s[7], s[11], s[15]) * y);
    return A * x;
}
// SplineMat.h:178-187 — no degenerate fix.
template<class T> inline V3<T> normalMat(const V4<T> s[16], T u, T v)
{
    V4<T> dU = derivativeUMat(s, u, v);
    V4<T> dV = derivativeVMat(s, u, v);
    return dU.xyz().cross(dV.xyz());
}

// SplineMat.h:64-82 — closed-form slice matrix (columns).
template<class T> inline M4<T> sliceMatrix(T a, T b)
{
    const T a2 = a * a, b2 = b * b, ia = 1 - a, ib = 1 - b, ia2 = ia * ia, ib2 = ib * ib,
            a_2b_3ab = a + 2 * b - 3 * a * b, _2a_b_3ab = 2 * a + b - 3 * a * b;
    return {{{ia2 * ia, 3 * a * ia2, 3 * a2 * ia, a2 * a},
             {ia2 * ib, ia * _2a_b_3ab, a * a_2b_3ab, a2 * b},
             {ia * ib2, ib * a_2b_3ab, b * _2a_b_3ab, a * b2},
             {ib2 * ib, 3 * b * ib2, 3 * b2 * ib, b2 * b}}};
}

// ---------------------------------------------------------------------------
// Bézier clipping, host semantics (Spline_BezierClipping.cu:174-445)
// ---------------------------------------------------------------------------

// Spline_BezierClipping.cu:127-156
template<class T> inline bool xAxisIntersection(V2<T>& result, V2<T> A, V2<T> B)
{
    const T eps = C<T>::eps;
    T q = A.y - B.y;
    if (std::fabs(q) > eps)
    {
        T t = A.y / q;
        if (T(0) <= t && t <= T(1))
        {
            T s = (T(1) - t) * A.x + t * B.x;
            result = {s, s};
            return true;
        }
    }
    else if (std::fabs(A.y) <= eps)
    {
        result = {fmin_(A.x, B.x), fmax_(A.x, B.x)};
        return true;
    }
    return false;
}

// Spline_BezierClipping.cu:161-165
template<class T> inline T isCCW(V2<T> P, V2<T> Q) { return P.x * Q.y - P.y * Q.x; }

// Spline.h:709-737
template<class T> inline bool triangleOriginIntersection(V2<T>& coord, V2<T> P, V2<T> A, V2<T> B)
{
    V2<T> p = A - P;
    V2<T> q = B - P;
    T det = p.x * q.y - q.x * p.y;
    if (std::fabs(det) < T(0x1p-80f)) return false;
    det = T(1) / det;
    V2<T> m0 {q.y, -q.x};
    V2<T> m1 {-p.y, p.x};
    T u = det * m0.dot(-P);
    if (u < 0 || u > 1) return false;
    T v = det * m1.dot(-P);
    if (v < 0 || u + v > 1) return false;
    coord = {u, v};
    return true;
}

// Spline.h:739-765
template<class T, class F>
inline bool approximateIntersection(F&& onHit, const V2<T> p[16], V2<T> pmin, V2<T> pmax, V2<T> psize)
{
    bool found = false;
    V2<T> p00 = p[0], p10 = p[3], p01 = p[12], p11 = p[15];
    V2<T> coord;
    if (triangleOriginIntersection(coord, p00, p10, p01))
        found |= onHit(pmin + coord * psize);
    if (triangleOriginIntersection(coord, p11, p01, p10))
        found |= onHit(pmax - coord * psize);
    return found;
}

// Optional cap on the clip's region stack (0: unbounded, the reference's
// host semantics).  With a cap the clip stops, as the GPU clipper does, when
// a split would store more than `cap` regions below the current one, and
// counts an overflow; test infrastructure for finding such rays without
// running an unbounded clip.
inline thread_local int clipStackCap = 0;

template<class T> struct ClipConst
{
    static constexpr T over3[4] {T(0.0f), T(1.0f / 3.0f), T(2.0f / 3.0f), T(1.0f)};
};
template<> struct ClipConst<double>
{
    static constexpr double over3[4] {0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0};
};

// Spline_BezierClipping.cu:174-445 — the host variant, with its heap stack of
// full states (the working patch is stored per entry, so a popped region is
// exactly the patch that was split, not a recomputation).
template<class T, class F>
bool clip2D(F&& onHit, const V2<T> patch[16], T tol, Work* w = nullptr, uint32_t maxIters = 0)
{
    struct State
    {
        V2<T> p[16];
        V2<T> mn, mx, sz;
        int side; // 0 = U, 1 = V
    };
    bool found = false;
    std::vector<State> S;
    S.push_back({});
    S.back().mn = {0, 0};
    S.back().mx = {1, 1};
    S.back().sz = {1, 1};
    S.back().side = 0;
    for (int i = 0; i < 16; i++) S.back().p[i] = patch[i];
    if (w) w->maxStack = std::max<uint64_t>(w->maxStack, 1);

    uint32_t iters = 0;
    do
    {
        // opt-in cap on the loop passes (no counterpart in the reference)
        if (maxIters && iters++ >= maxIters) break;
        State& e = S.back();
        V2<T>* p = e.p;
        // Spline_BezierClipping.cu:240-248 — tolerance on the two diagonals
        T d0 = (p[3] - p[12]).length();
        T d1 = (p[15] - p[0]).length();
        if (d0 < tol && d1 < tol)
        {
            if (w) w->converged++;
            found |= approximateIntersection<T>(onHit, p, e.mn, e.mx, e.sz);
            S.pop_back();
            continue;
        }
        if (w) w->clipIters++;

        T lower = 1, upper = 0;
        T ytop[4], ybot[4];
        {
            // :262-279 — line L through the patch, normal N, distance patch
            V2<T> L = p[15] - p[0];
            V2<T> D = p[12] - p[3];
            if (e.side == 0) L = L + D; else L = L - D;
            L = normalized(L);
            const V2<T> N {-L.y, L.x};
            T d[16];
            for (int i = 0; i < 16; i++) d[i] = N.dot(p[i]);
            // :281-307 — envelope per column (U) or row (V)
            if (e.side == 0)
            {
                for (int i = 0; i < 4; i++)
                {
                    ybot[i] = ytop[i] = d[i];
                    for (int k = 1; k < 4; k++)
                    {
                        T y = d[4 * k + i];
                        ytop[i] = fmax_(ytop[i], y);
                        ybot[i] = fmin_(ybot[i], y);
                    }
                }
            }
            else
            {
                for (int i = 0; i < 4; i++)
                {
                    ybot[i] = ytop[i] = d[4 * i];
                    for (int k = 1; k < 4; k++)
                    {
                        T y = d[4 * i + k];
                        ytop[i] = fmax_(ytop[i], y);
                        ybot[i] = fmin_(ybot[i], y);
                    }
                }
            }
        }
        {
            // :312-348 — fixed-order hull walk
            const T* o3 = ClipConst<T>::over3;
            const V2<T> hull[7] {
                {o3[2], ybot[2]}, {o3[3], ybot[3]}, {o3[3], ytop[3]}, {o3[2], ytop[2]},
                {o3[1], ytop[1]}, {o3[0], ytop[0]}, {o3[0], ybot[0]}};
            V2<T> P1 {o3[0], ybot[0]};
            V2<T> P2 {o3[1], ybot[1]};
            for (const V2<T>& P3 : hull)
            {
                if (isCCW(P2 - P1, P3 - P1) >= 0)
                {
                    V2<T> s;
                    if (xAxisIntersection(s, P1, P2))
                    {
                        lower = fmin_(lower, s.x);
                        upper = fmax_(upper, s.y);
                    }
                    P1 = P2;
                }
                P2 = P3;
            }
            V2<T> s;
            if (xAxisIntersection(s, P1, P2))
            {
                lower = fmin_(lower, s.x);
                upper = fmax_(upper, s.y);
            }
        }
        // :351-353 — widening against long, narrow subpatches
        lower = lower * T(0x0FFp-8f);
        upper = upper + (T(1) - upper) * T(0x001p-8f);
        T delta = upper - lower;

        if (delta < 0)
        {
            if (w) w->prunes++;
            S.pop_back();
        }
        else if (delta > T(0.8f))
        {
            // :373-409 — split in half; the upper half is pushed (processed first)
            if (clipStackCap > 0 && S.size() - 1 >= (size_t)clipStackCap)
            {
                if (w) w->overflows++;
                return found;
            }
            if (w) w->subdivs++;
            State s1;
            for (int i = 0; i < 16; i++) s1.p[i] = e.p[i];
            if (e.side == 0)
            {
                T hs = T(0.5f) * e.sz.x;
                T hx = e.mn.x + hs;
                s1.mx = e.mx;
                s1.mn = {hx, e.mn.y};
                e.mx = {hx, e.mx.y};
                s1.sz = e.sz = {hs, e.sz.y};
                s1.side = e.side = 1;
                subpatchU(e.p, T(0), T(0.5f));
                subpatchU(s1.p, T(0.5f), T(1));
            }
            else
            {
                T hs = T(0.5f) * e.sz.y;
                T hy = e.mn.y + hs;
                s1.mx = e.mx;
                s1.mn = {e.mn.x, hy};
                e.mx = {e.mx.x, hy};
                s1.sz = e.sz = {e.sz.x, hs};
                s1.side = e.side = 0;
                subpatchV(e.p, T(0), T(0.5f));
                subpatchV(s1.p, T(0.5f), T(1));
            }
            S.push_back(s1);
            if (w) w->maxStack = std::max<uint64_t>(w->maxStack, S.size());
        }
        else
        {
            // :410-432 — clip to [lower, upper]
            if (e.side == 0)
            {
                T u0 = e.mn.x + e.sz.x * lower;
                T u1 = e.mx.x - e.sz.x * (T(1) - upper);
                subpatchU(e.p, lower, upper);
                e.mn.x = u0;
                e.mx.x = u1;
                e.sz.x = u1 - u0;
                e.side = 1;
            }
            else
            {
                T v0 = e.mn.y + e.sz.y * lower;
                T v1 = e.mx.y - e.sz.y * (T(1) - upper);
                subpatchV(e.p, lower, upper);
                e.mn.y = v0;
                e.mx.y = v1;
                e.sz.y = v1 - v0;
                e.side = 0;
            }
        }
    } while (!S.empty());
    return found;
}

// Ds-boundary: cg::Intersection {distance, p} and Ray3f {origin, direction}
// (the Ray3f constructor normalizes the direction).
template<class T> struct PatchHit
{
    T distance;
    V2<T> p;
};

// Spline_BezierClipping.cu:46-114 — doBezierClipping.  `dir` must already be
// the Ray3f direction (versor of the caller's direction).
template<class T>
bool doBezierClipping(PatchHit<T>& hit, V3<T> origin, V3<T> dir, const V4<T> S[16],
    T tol = T(0x1p-12f), Work* w = nullptr, uint32_t maxIters = 0)
{
    if (w) w->patchTests++;
    V2<T> p2D[16];
    const V3<T> axis0 = versor(dir.cross(dir + V3<T>{T(0), T(0), T(8)}));
    const V3<T> axis1 = versor(dir.cross(axis0));
    for (int i = 0; i < 16; i++)
    {
        const V3<T> Q {S[i].x, S[i].y, S[i].z};
        p2D[i] = {axis0.dot(Q - origin), axis1.dot(Q - origin)};
    }
    auto onHit = [&](V2<T> e) -> bool
    {
        if (w) w->onHits++;
        V3<T> V = project(interpolateMat(S, e.x, e.y)) - origin;
        T t = V.length();
        if (t < hit.distance && V.dot(dir) > 0)
        {
            hit.distance = t;
            hit.p = e;
            return true;
        }
        return false;
    };
    return clip2D<T>(onHit, p2D, tol, w, maxIters);
}

// Newton polish of a closest hit (an opt-in refinement with no counterpart
// in the reference; the library's polishHit64, cgs_device.hpp, with the same
// FP64 operations in the same order).  The object-local ray is rebuilt in
// FP64 from the FP32 world ray and world→local matrix, o_l + t M3 d with t
// the world parameter; up to six Newton steps on S(u,v) - o_l - t M3 d = 0
// (Cramer's rule, the quotient rule for w) refine (u, v, t), kept when the
// last step is below 1e-9 in u and v and lands inside the parameter square
// with t > 0.
namespace polish
{
inline void comb4(const double* a, const double* b, const double* c, const double* d, const double* x, double* r)
{
    for (int k = 0; k < 4; k++) r[k] = ((x[0] * a[k] + x[1] * b[k]) + x[2] * c[k]) + x[3] * d[k];
}

// the cubic Bernstein polynomials and their derivatives in power form
inline void basis(double t, double* b, double* db)
{
    const double t2 = t * t, t3 = t2 * t;
    b[0] = ((1.0 - 3.0 * t) + 3.0 * t2) - t3;
    b[1] = (3.0 * t - 6.0 * t2) + 3.0 * t3;
    b[2] = 3.0 * t2 - 3.0 * t3;
    b[3] = t3;
    db[0] = (-3.0 + 6.0 * t) - 3.0 * t2;
    db[1] = (3.0 - 12.0 * t) + 9.0 * t2;
    db[2] = 6.0 * t - 9.0 * t2;
    db[3] = 3.0 * t2;
}
} // namespace polish

inline bool polishHit64(const V4<float> S[16], const M4<float>& w2l, V3<float> ow, V3<float> dw, float& t, float& u,
                        float& v)
{
    using V3d = V3<double>;
    double M[4][4];
    for (int c = 0; c < 4; c++)
    {
        M[c][0] = w2l.c[c].x;
        M[c][1] = w2l.c[c].y;
        M[c][2] = w2l.c[c].z;
        M[c][3] = w2l.c[c].w;
    }
    const double ox = ow.x, oy = ow.y, oz = ow.z, dx = dw.x, dy = dw.y, dz = dw.z;
    double O4[4];
    for (int k = 0; k < 4; k++) O4[k] = ((ox * M[0][k] + oy * M[1][k]) + oz * M[2][k]) + M[3][k];
    const V3d o {O4[0] / O4[3], O4[1] / O4[3], O4[2] / O4[3]};
    const V3d D {(dx * M[0][0] + dy * M[1][0]) + dz * M[2][0], (dx * M[0][1] + dy * M[1][1]) + dz * M[2][1],
                 (dx * M[0][2] + dy * M[1][2]) + dz * M[2][2]};
    double pu = u, pv = v, pt = t;
    bool converged = false;
    for (int it = 0; it < 6; it++)
    {
        double bu[4], dbu[4], bv[4], dbv[4];
        polish::basis(pu, bu, dbu);
        polish::basis(pv, bv, dbv);
        double a[4][4], ad[4][4];
        for (int j = 0; j < 4; j++)
        {
            double q[4][4];
            for (int i = 0; i < 4; i++)
            {
                const V4<float>& c = S[4 * j + i];
                q[i][0] = c.x;
                q[i][1] = c.y;
                q[i][2] = c.z;
                q[i][3] = c.w;
            }
            polish::comb4(q[0], q[1], q[2], q[3], bu, a[j]);
            polish::comb4(q[0], q[1], q[2], q[3], dbu, ad[j]);
        }
        double X[4], Xu[4], Xv[4];
        polish::comb4(a[0], a[1], a[2], a[3], bv, X);
        polish::comb4(ad[0], ad[1], ad[2], ad[3], bv, Xu);
        polish::comb4(a[0], a[1], a[2], a[3], dbv, Xv);
        const double iw = 1.0 / X[3];
        const V3d p {X[0] * iw, X[1] * iw, X[2] * iw};
        const V3d Pu = CGS_ORACLE_PLANT == 9
                           ? V3d {Xu[0] * iw, Xu[1] * iw, Xu[2] * iw}
                           : V3d {(Xu[0] - p.x * Xu[3]) * iw, (Xu[1] - p.y * Xu[3]) * iw, (Xu[2] - p.z * Xu[3]) * iw};
        const V3d Pv {(Xv[0] - p.x * Xv[3]) * iw, (Xv[1] - p.y * Xv[3]) * iw, (Xv[2] - p.z * Xv[3]) * iw};
        const double sgn = CGS_ORACLE_PLANT == 8 ? -1.0 : 1.0;
        const V3d F {(p.x - o.x) - sgn * pt * D.x, (p.y - o.y) - sgn * pt * D.y, (p.z - o.z) - sgn * pt * D.z};
        const V3d mD {-D.x, -D.y, -D.z};
        auto dot = [](V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
        auto cross = [](V3d a, V3d b) -> V3d
        { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
        const double det = dot(Pu, cross(Pv, mD));
        if (!(det > 1e-300 || det < -1e-300)) return false;
        const double idet = 1.0 / det;
        const V3d nF {-F.x, -F.y, -F.z};
        const double du = dot(nF, cross(Pv, mD)) * idet;
        const double dv = dot(Pu, cross(nF, mD)) * idet;
        const double dt = dot(Pu, cross(Pv, nF)) * idet;
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

// ---------------------------------------------------------------------------
// Subdivision intersection, host semantics (Spline_Subdivision.cu:168-288)
// ---------------------------------------------------------------------------

// Ds-boundary: Bounds3::intersect(ray, tmin, tmax) — slab test over the ray's
// [tMin, tMax] (tMin = 0, tMax = +inf for the Ray3f built by the callers).
template<class T> inline bool boxIntersect(const Box<T>& b, V3<T> o, V3<T> d, T& tmin, T& tmax)
{
    tmin = 0;
    tmax = std::numeric_limits<T>::infinity();
    for (int i = 0; i < 3; i++)
    {
        T inv = T(1) / d[i];
        T t0 = (b.mn[i] - o[i]) * inv;
        T t1 = (b.mx[i] - o[i]) * inv;
        if (inv < 0) std::swap(t0, t1);
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
        if (tmin > tmax) return false;
    }
    return true;
}

// Ds-boundary: triangle::intersect (Möller–Trumbore), writes barycentrics
// (u,v) of the hit relative to (p0,p1,p2) and t; accepts t in (0, +inf).
template<class T>
inline bool triangleIntersect(V3<T> o, V3<T> d, V3<T> p0, V3<T> p1, V3<T> p2, V3<T>& uvw, T& t)
{
    V3<T> e1 = p1 - p0, e2 = p2 - p0;
    V3<T> s1 = d.cross(e2);
    T det = s1.dot(e1);
    if (std::fabs(det) < std::numeric_limits<T>::min()) return false;
    T inv = T(1) / det;
    V3<T> s = o - p0;
    T b1 = s.dot(s1) * inv;
    if (b1 < 0 || b1 > 1) return false;
    V3<T> s2 = s.cross(e1);
    T b2 = d.dot(s2) * inv;
    if (b2 < 0 || b1 + b2 > 1) return false;
    T tt = e2.dot(s2) * inv;
    if (!(tt > 0)) return false;
    t = tt;
    uvw = {b1, b2, T(1) - b1 - b2};
    return true;
}

template<class T>
bool doSubdivision(PatchHit<T>& hit, V3<T> o, V3<T> d, const V4<T> S[16], T threshold)
{
    struct State { V2<T> mn, mx; bool side; };
    std::vector<State> st;
    T tmin, tmax;
    Box<T> pb;
    for (int i = 0; i < 16; i++) pb.inflate(project(S[i]));
    if (!boxIntersect(pb, o, d, tmin, tmax)) return false;
    st.push_back({{0, 0}, {T(0.5f), 1}, false});
    st.push_back({{T(0.5f), 0}, {1, 1}, false});
    bool found = false;
    V4<T> sub[16];
    while (!st.empty())
    {
        State e = st.back();
        st.pop_back();
        for (int i = 0; i < 16; i++) sub[i] = S[i];
        // Spline.h:462-478 — subpatchBoundingbox (V then U)
        subpatchV(sub, e.mn.y, e.mx.y);
        subpatchU(sub, e.mn.x, e.mx.x);
        pb = Box<T>();
        for (int i = 0; i < 16; i++) pb.inflate(project(sub[i]));
        if (!boxIntersect(pb, o, d, tmin, tmax)) continue;
        T x = T(0.5f) * (e.mn.x + e.mx.x);
        T y = T(0.5f) * (e.mn.y + e.mx.y);
        const bool small = CGS_ORACLE_PLANT == 2
                               ? (e.mx.x - e.mn.x) <= threshold && (e.mx.y - e.mn.y) <= threshold
                               : (e.mx.x - e.mn.x) < threshold && (e.mx.y - e.mn.y) < threshold;
        if (small)
        {
            V3<T> p00 = project(sub[0]), p10 = project(sub[3]), p01 = project(sub[12]), p11 = project(sub[15]);
            V3<T> uvw, uvw1;
            T t = 0, t1 = 0;
            bool b0 = triangleIntersect(o, d, p10, p01, p00, uvw, t);
            bool b1 = triangleIntersect(o, d, p01, p10, p11, uvw1, t1);
            if (b0)
            {
                if (b1 && t1 < t) t = t1;
            }
            else if (b1)
            {
                t = t1;
                b0 = true;
            }
            if (b0 && t < hit.distance)
            {
                hit.distance = t;
                hit.p = CGS_ORACLE_PLANT == 1 ? V2<T> {e.mn.x, e.mn.y} : V2<T> {x, y};
                found = true;
            }
        }
        else if (e.side)
        {
            st.push_back({{e.mn.x, e.mn.y}, {x, e.mx.y}, !e.side});
            st.push_back({{x, e.mn.y}, {e.mx.x, e.mx.y}, !e.side});
        }
        else
        {
            st.push_back({{e.mn.x, e.mn.y}, {e.mx.x, y}, !e.side});
            st.push_back({{e.mn.x, y}, {e.mx.x, e.mx.y}, !e.side});
        }
    }
    return found;
}

// ---------------------------------------------------------------------------
// BVH (rt/BVH.h, rt/BVH.cpp)
// ---------------------------------------------------------------------------

constexpr uint32_t EMPTY = 0xFFFFFFFFu;
constexpr uint32_t ROOT_KEY = 1;

// rt/BVH.h:23-39: two child boxes + left/right (or first) offsets.
struct Node
{
    Box<float> leftBox, rightBox;
    uint32_t left = 0, right = 0; // right doubles as `first` for leaves
    bool isLeaf() const { return left == EMPTY; }
};

// rt/BVH.h:41-46
struct Element
{
    Box<float> bounds;
    V3<float> centroid;
    uint32_t index;
};

struct BVHData
{
    std::vector<Node> nodes;
    std::vector<uint32_t> keys;
    std::vector<uint32_t> indices;
    uint32_t elementsPerNode = 1;
};

// rt/BVH.cpp:82-108
inline uint32_t bvhWrap(BVHData& b, std::span<Element> el, uint32_t key)
{
    if (el.empty()) throw std::logic_error("invalid");
    Box<float> box;
    const uint32_t idx = (uint32_t)b.indices.size();
    b.indices.resize(idx + el.size() + 1);
    for (size_t i = 0; i < el.size(); i++)
    {
        b.indices[idx + i] = el[i].index;
        box.inflate(el[i].bounds);
    }
    b.indices[idx + el.size()] = EMPTY;
    b.keys.push_back(key);
    Node n;
    n.leftBox = box;
    n.rightBox = Box<float>();
    n.left = EMPTY;
    n.right = idx;
    b.nodes.push_back(n);
    return (uint32_t)b.nodes.size() - 1;
}

// rt/BVH.cpp:40-80 — the split axis comes from a box that is never inflated,
// so every size component is equal and the axis is always Y (dim = 1).
inline uint32_t bvhSplit(BVHData& b, std::span<Element> el, uint32_t key)
{
    size_t count = el.size();
    if (count <= b.elementsPerNode || (key & (1u << 31)))
        return bvhWrap(b, el, key);
    Box<float> box;
    V3<float> v = box.size();
    int dim = v.z > v.y ? (v.x > v.z ? 0 : 2) : (v.x > v.y ? 0 : 1);
    size_t middle = count / 2;
    std::nth_element(el.begin(), el.begin() + middle, el.end(),
        [=](const Element& a, const Element& c) { return a.centroid[dim] < c.centroid[dim]; });
    b.keys.push_back(key);
    const uint32_t idx = (uint32_t)b.nodes.size();
    b.nodes.emplace_back();
    uint32_t l = bvhSplit(b, el.subspan(0, middle), key << 1);
    uint32_t r = bvhSplit(b, el.subspan(middle), (key << 1) ^ 1u);
    Node& node = b.nodes[idx];
    node.left = l;
    node.right = r;
    const Node& L = b.nodes[l];
    const Node& R = b.nodes[r];
    node.leftBox = L.leftBox;
    node.rightBox = R.leftBox;
    if (!L.isLeaf()) node.leftBox.inflate(L.rightBox);
    if (!R.isLeaf()) node.rightBox.inflate(R.rightBox);
    return idx;
}

// rt/BVH.cpp:9-22
inline BVHData bvhBuild(std::vector<Element> el, uint32_t elementsPerNode = 1)
{
    BVHData b;
    b.elementsPerNode = elementsPerNode;
    b.nodes.reserve(2 * el.size());
    b.keys.reserve(2 * el.size());
    b.indices.reserve(el.size());
    bvhSplit(b, std::span<Element>(el), ROOT_KEY);
    return b;
}

// BezierSurface::buildBVH (Primitive.cu:199-220) — per-patch element data: bounds and centroid of the
// projected control points (centroid accumulated as += (1/16)·p).
inline std::vector<Element> patchElements(const float* vertices4, const uint32_t* indices, uint32_t patchCount)
{
    std::vector<Element> el(patchCount);
    for (uint32_t i = 0; i < patchCount; i++)
    {
        Element& e = el[i];
        e.bounds = Box<float>();
        e.centroid = {0, 0, 0};
        e.index = i;
        const float inv = 1.0f / 16.0f;
        for (int j = 0; j < 16; j++)
        {
            const float* w = vertices4 + 4 * (size_t)indices[16 * (size_t)i + j];
            V3<float> p = project(V4<float>{w[0], w[1], w[2], w[3]});
            e.bounds.inflate(p);
            e.centroid = e.centroid + inv * p;
        }
    }
    return el;
}

// ---------------------------------------------------------------------------
// Perfect hash function (rt/PerfectHashFunction.{h,cu}) — deterministic build
//
// The reference's GPU build is race-dependent (slot order within a bucket,
// bucket order within a size class and contention winners all come from
// atomics), so its output is not reproducible even by itself.  This is the
// deterministic restatement both this oracle and the HIP build implement:
//  * hash family, table sizing, growth, (A,B) draws and attempt budget as in
//    the reference (PerfectHashFunction.h:45-63, .cu:229-484);
//  * buckets = key & mask; keys kept in input order inside a bucket;
//  * size classes 4..1; inside a class buckets are ranked g by ascending id;
//  * rounds: every pending bucket proposes d = a·(g+1) (a = its attempt
//    count); a proposal is valid when its slots are pairwise distinct and
//    free; a slot claimed by several valid proposals goes to the lowest g; a
//    bucket that wins all of its slots is placed, every other pending bucket
//    advances a; a bucket reaching 32 attempts fails the (A,B) attempt.
// ---------------------------------------------------------------------------

// The wide family (no counterpart in the reference) is the same function
// evaluated in exact integer arithmetic on keys of up to 64 bits:
// h(k) = (A·(k mod T) + B + D[k & mask]) mod T, equal to (A·k + B + D) mod T
// without the reference's 32-bit wrap of A·k, which breaks the function's
// near-bijection on dense heap keys once A·k passes 2^32 (at 2^21 keys every
// reference attempt fails).  Sizes, draws and the build are the reference's.
struct PHF
{
    uint32_t A = 31, B = 19;
    uint32_t mask = 0;
    uint32_t tableSize = 0;
    std::vector<uint32_t> D; // displacement, size mask+1
    int attempts = 0;        // (A,B) pairs tried
    bool wide = false;       // the wide family
    uint32_t hash(uint64_t key) const { return slot(key, D[(uint32_t)(key & mask)], tableSize); }
    // PerfectHashFunction.h:58-63 (32-bit arithmetic), or the wide family
    uint32_t slot(uint64_t key, uint32_t d, uint32_t T) const
    {
        if (wide) return (uint32_t)(((uint64_t)A * (key % T) + B + d) % T);
        return (A * (uint32_t)key + B + d) % T;
    }
};

// Returns 0 on success, 1 when the final function is not injective
// (reference: "DUPLICATE FOUND!", PerfectHashFunction.cu:486-502), 2 when a
// bucket holds more than 4 keys (reference: __trap in init_and_bucketsort).
// wide: build the wide family (keys of up to 64 bits).
template<class Key = uint32_t>
inline int phfBuild(PHF& h, std::span<const Key> keys, uint32_t seed = 42, bool wide = false)
{
    h.wide = wide;
    constexpr uint32_t kCap = 4, kMaxAttempts = 32, kMaxFunctions = 16;
    const size_t n = keys.size();
    int s = int(n - 1) >> 1;
    if (s < 2) s = 2;
    const uint32_t displSize = std::bit_floor((unsigned)s);
    h.mask = displSize - 1u;
    h.tableSize = (uint32_t)n;
    h.D.assign(displSize, 0);
    const uint32_t step = std::max<uint32_t>(4u, (uint32_t)(0.05f * (float)n));
    std::mt19937 mt(seed);
    std::uniform_int_distribution<uint32_t> dist(0, (uint32_t)n);

    // buckets (input order inside each bucket)
    std::vector<uint32_t> bsize(displSize, 0);
    std::vector<Key> bkeys(size_t(displSize) * kCap, 0);
    for (size_t i = 0; i < n; i++)
    {
        uint32_t j = (uint32_t)(keys[i] & h.mask);
        uint32_t k = bsize[j]++;
        if (k >= kCap) return 2;
        bkeys[size_t(j) * kCap + k] = keys[i];
    }
    // classes: ascending bucket id inside each size
    std::vector<std::vector<uint32_t>> cls(kCap + 1);
    for (uint32_t b = 0; b < displSize; b++)
        if (bsize[b] > 0) cls[bsize[b]].push_back(b);

    uint32_t T = (uint32_t)n;
    std::vector<uint8_t> occ;
    std::vector<uint32_t> claim; // lowest g claiming a slot this round
    h.attempts = 0;
    for (uint32_t fn = 0; fn < kMaxFunctions; fn++)
    {
        do
        {
            h.A = dist(mt);
            if (h.A == 0) h.A = 31;
            h.B = dist(mt);
        } while (std::gcd((size_t)n, (size_t)h.A) > 1);
        h.attempts++;
        h.tableSize = T;
        std::fill(h.D.begin(), h.D.end(), 0u);
        occ.assign(T, 0);
        claim.assign(T, EMPTY);
        uint64_t failures = 0;
        for (int sz = kCap; sz >= 1; sz--)
        {
            const auto& list = cls[sz];
            std::vector<uint32_t> att(list.size(), 0);
            std::vector<uint32_t> pend(list.size());
            std::iota(pend.begin(), pend.end(), 0u);
            std::vector<uint32_t> slots(list.size() * kCap);
            std::vector<uint8_t> valid(list.size());
            while (!pend.empty())
            {
                // propose
                for (uint32_t g : pend)
                {
                    uint32_t b = list[g];
                    uint32_t d = att[g] * (g + 1u);
                    bool ok = true;
                    for (int k = 0; k < sz; k++)
                    {
                        const Key key = bkeys[size_t(b) * kCap + k];
                        uint32_t sl = h.slot(key, d, T);
                        slots[size_t(g) * kCap + k] = sl;
                        if (occ[sl]) ok = false;
                        for (int m = 0; m < k; m++)
                            if (slots[size_t(g) * kCap + m] == sl) ok = false;
                    }
                    valid[g] = ok;
                    if (ok)
                        for (int k = 0; k < sz; k++)
                        {
                            uint32_t sl = slots[size_t(g) * kCap + k];
                            if (g < claim[sl]) claim[sl] = g;
                        }
                }
                // resolve
                std::vector<uint32_t> next;
                next.reserve(pend.size());
                std::vector<uint32_t> placed;
                for (uint32_t g : pend)
                {
                    bool won = valid[g];
                    if (won)
                        for (int k = 0; k < sz; k++)
                            if (claim[slots[size_t(g) * kCap + k]] != g) won = false;
                    if (won)
                    {
                        placed.push_back(g);
                        h.D[list[g]] = att[g] * (g + 1u);
                    }
                    else if (++att[g] >= kMaxAttempts)
                        failures++;
                    else
                        next.push_back(g);
                }
                for (uint32_t g : pend)
                    if (valid[g])
                        for (int k = 0; k < sz; k++) claim[slots[size_t(g) * kCap + k]] = EMPTY;
                for (uint32_t g : placed)
                    for (int k = 0; k < sz; k++) occ[slots[size_t(g) * kCap + k]] = 1;
                pend.swap(next);
            }
        }
        if (failures == 0) break;
        T = T + step;
    }
    // PerfectHashFunction.cu:486-502 — injectivity check on the final function
    // (after a last failure the table size has grown past the one used).
    if (h.tableSize != T) h.tableSize = T;
    std::vector<uint8_t> seen(h.tableSize, 0);
    bool dup = false;
    for (Key k : keys)
    {
        uint32_t v = h.hash(k);
        if (seen[v]) dup = true;
        seen[v] = 1;
    }
    return dup ? 1 : 0;
}

// rt/BVH.cpp:24-38 — table[hash(key_i)] = i, EMPTY elsewhere.
template<class Key = uint32_t>
inline std::vector<uint32_t> hashTable(const PHF& h, std::span<const Key> keys)
{
    std::vector<uint32_t> t(h.tableSize, EMPTY);
    for (size_t i = 0; i < keys.size(); i++) t[h.hash(keys[i])] = (uint32_t)i;
    return t;
}

// ---------------------------------------------------------------------------
// Hashed stackless traversal (rt/BVH.h:147-350)
// ---------------------------------------------------------------------------

struct Ray
{
    V3<float> origin, direction;
    float max;
};
struct Hit
{
    V3<float> coordinates {0, 0, 0};
    uint32_t index = EMPTY;
    float t = std::numeric_limits<float>::infinity();
};

// rt/BVH.h:147-176
inline bool nodeIntersect(float& tMin, float& tMax, const Box<float>& b, V3<float> o, V3<float> Dinv)
{
    tMin = -std::numeric_limits<float>::infinity();
    tMax = std::numeric_limits<float>::infinity();
    for (int i = 0; i < 3; i++)
    {
        int s = std::signbit(Dinv[i]) ? 1 : 0;
        float t[2];
        t[0] = (b.mn[i] - o[i]) * Dinv[i];
        t[1] = (b.mx[i] - o[i]) * Dinv[i];
        if (t[0 + s] > tMin) tMin = t[0 + s];
        if (t[1 - s] < tMax) tMax = t[1 - s];
    }
    return tMin < tMax && tMax > 0.0001f;
}

struct BVHView
{
    const Node* nodes;
    const uint32_t* indices;
    const uint32_t* table;
    const PHF* hash;
};

// rt/BVH.h:178-266 — closest-hit traversal; `fn(hit, index)` tests a leaf element.
template<class F>
bool hashIntersect(const BVHView& bv, Hit& hit, const Ray& ray, F&& fn, Work* w = nullptr)
{
    const V3<float> Dinv = ray.direction.inverse();
    const Node* node = bv.nodes;
    uint32_t key = ROOT_KEY, postponed = 0;
    bool result = false;
    while (true)
    {
        if (!node->isLeaf())
        {
            if (w) w->nodes++;
            float L0, L1, R0, R1;
            bool iL = nodeIntersect(L0, L1, node->leftBox, ray.origin, Dinv) && L0 < hit.t;
            bool iR = nodeIntersect(R0, R1, node->rightBox, ray.origin, Dinv) && R0 < hit.t;
            if (iL || iR) postponed <<= 1;
            if (iL)
            {
                if (iR)
                {
                    postponed |= 1;
                    if (w) w->maxPostponed = std::max<uint64_t>(w->maxPostponed, std::popcount(postponed));
                    if (L0 < R0) { key = key << 1; node = bv.nodes + node->left; }
                    else { key = (key << 1) ^ 1u; node = bv.nodes + node->right; }
                }
                else { key = key << 1; node = bv.nodes + node->left; }
                continue;
            }
            else if (iR)
            {
                key = (key << 1) ^ 1u;
                node = bv.nodes + node->right;
                continue;
            }
        }
        else
        {
            if (w) w->leaves++;
            uint32_t i = node->right;
            while (bv.indices[i] != EMPTY)
            {
                bool b = fn(hit, bv.indices[i++]);
                result |= b;
            }
        }
        if (!postponed) break;
        int levels = std::countr_zero(postponed);
        postponed >>= levels;
        key = (key >> levels) ^ 1u;
        postponed ^= 1;
        if (w) w->hashProbes++;
        node = bv.nodes + bv.table[bv.hash->hash(key)];
    }
    return result;
}

// rt/BVH.h:268-350 — any-hit traversal against ray.max.
template<class F>
bool hashIntersectAny(const BVHView& bv, const Ray& ray, F&& fn, Work* w = nullptr)
{
    const V3<float> Dinv = ray.direction.inverse();
    const Node* node = bv.nodes;
    uint32_t key = ROOT_KEY, postponed = 0;
    while (true)
    {
        if (!node->isLeaf())
        {
            if (w) w->nodes++;
            float L0, L1, R0, R1;
            bool iL = nodeIntersect(L0, L1, node->leftBox, ray.origin, Dinv) && L0 < ray.max;
            bool iR = nodeIntersect(R0, R1, node->rightBox, ray.origin, Dinv) && R0 < ray.max;
            if (iL || iR) postponed <<= 1;
            if (iL)
            {
                if (iR)
                {
                    postponed |= 1;
                    if (w) w->maxPostponed = std::max<uint64_t>(w->maxPostponed, std::popcount(postponed));
                    if (L0 < R0) { key = key << 1; node = bv.nodes + node->left; }
                    else { key = (key << 1) ^ 1u; node = bv.nodes + node->right; }
                }
                else { key = key << 1; node = bv.nodes + node->left; }
                continue;
            }
            else if (iR)
            {
                key = (key << 1) ^ 1u;
                node = bv.nodes + node->right;
                continue;
            }
        }
        else
        {
            if (w) w->leaves++;
            uint32_t i = node->right;
            while (bv.indices[i] != EMPTY)
                if (fn(bv.indices[i++])) return true;
        }
        if (!postponed) break;
        int levels = std::countr_zero(postponed);
        postponed >>= levels;
        key = (key >> levels) ^ 1u;
        postponed ^= 1;
        if (w) w->hashProbes++;
        node = bv.nodes + bv.table[bv.hash->hash(key)];
    }
    return false;
}

// rt/BVH.cpp:126-189 — explicit-stack closest-hit traversal (cross-check).
template<class F>
bool stackIntersect(const BVHView& bv, Hit& hit, const Ray& ray, F&& fn)
{
    const V3<float> Dinv = ray.direction.inverse();
    bool result = false;
    std::vector<const Node*> S;
    S.push_back(bv.nodes);
    while (!S.empty())
    {
        const Node* node = S.back();
        S.pop_back();
        if (node->isLeaf())
        {
            uint32_t i = node->right;
            while (bv.indices[i] != EMPTY) result |= fn(hit, bv.indices[i++]);
            continue;
        }
        float L0, L1, R0, R1;
        bool iL = nodeIntersect(L0, L1, node->leftBox, ray.origin, Dinv) && L0 < hit.t;
        bool iR = nodeIntersect(R0, R1, node->rightBox, ray.origin, Dinv) && R0 < hit.t;
        if (iL)
        {
            if (iR)
            {
                if (L0 < R0) { S.push_back(bv.nodes + node->right); S.push_back(bv.nodes + node->left); }
                else { S.push_back(bv.nodes + node->left); S.push_back(bv.nodes + node->right); }
            }
            else S.push_back(bv.nodes + node->left);
        }
        else if (iR) S.push_back(bv.nodes + node->right);
    }
    return result;
}

} // namespace orc
