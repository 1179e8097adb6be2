// SPDX-License-Identifier: MIT
//
// Model readers on the host: the `.be` Bézier patch format read by
// Surface::load_be (Surface.cpp:149-213); and the BVH element data of a
// triangle mesh (Mesh::buildBVH, rt/Primitive.cu:90-108).
//
// Layout of a `.be` text: the patch count; per patch 16 one-based indices of
// its control points (u fastest, Spline.h:82); the point count; x, y, z per
// point.  Every number is followed by exactly one separator character (a
// comma or a newline), and whitespace before a number is skipped, which is
// how `input >> value; input.ignore();` reads it.
//
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>

#include "cgs_internal.hpp"

namespace
{

// std::istream's `>>` for the two token kinds the format uses.
struct Cursor
{
    const char* p;
    const char* end;

    void skipSpace()
    {
        while (p < end && std::isspace((unsigned char)*p)) ++p;
    }
    void ignoreOne()
    {
        if (p < end) ++p;
    }
    // Decimal unsigned integer; fails on no digits or a value above 2^32-1.
    bool readU32(uint32_t& v)
    {
        skipSpace();
        if (p < end && *p == '+') ++p;
        uint64_t x = 0;
        const char* s = p;
        while (p < end && *p >= '0' && *p <= '9')
        {
            x = 10 * x + uint64_t(*p - '0');
            if (x > 0xffffffffull) return false;
            ++p;
        }
        v = uint32_t(x);
        return p > s;
    }
    // [+-]digits[.digits][(e|E)[+-]digits], converted with strtof (correctly
    // rounded, as the stream's num_get does).  No inf/nan/hex forms.
    bool readFloat(float& v)
    {
        skipSpace();
        char buf[96];
        size_t n = 0;
        auto take = [&]() { if (n + 1 < sizeof buf) buf[n++] = *p; ++p; };
        if (p < end && (*p == '+' || *p == '-')) take();
        size_t digits = 0;
        while (p < end && std::isdigit((unsigned char)*p)) take(), ++digits;
        if (p < end && *p == '.')
        {
            take();
            while (p < end && std::isdigit((unsigned char)*p)) take(), ++digits;
        }
        if (!digits) return false;
        if (p < end && (*p == 'e' || *p == 'E'))
        {
            const char* save = p;
            size_t saveN = n;
            take();
            if (p < end && (*p == '+' || *p == '-')) take();
            size_t ed = 0;
            while (p < end && std::isdigit((unsigned char)*p)) take(), ++ed;
            if (!ed) p = save, n = saveN;
        }
        if (n + 1 >= sizeof buf) return false;
        buf[n] = 0;
        v = std::strtof(buf, nullptr);
        return true;
    }
};

} // namespace

cgs_status cgs_load_be(const char* text, size_t length, float** vertices, uint32_t* npoints,
                       uint32_t** indices, uint32_t* npatches)
{
    if (!vertices || !npoints || !indices || !npatches || (length && !text))
        return cgs::setError(CGS_ERR_INVALID, "cgs_load_be: null argument");
    *vertices = nullptr;
    *indices = nullptr;
    *npoints = *npatches = 0;
    Cursor in {text, text + length};

    uint32_t nPatches;
    if (!in.readU32(nPatches)) return cgs::setError(CGS_ERR_INVALID, "Could not read the no. patches");
    // Each index takes at least two characters (a digit and its separator):
    // a count the text cannot hold is a parse failure, not an allocation.
    if (uint64_t(nPatches) * 32 > length) return cgs::setError(CGS_ERR_INVALID, "Parsing patch failed");
    const size_t nIdx = size_t(16) * nPatches;
    uint32_t* idx = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<size_t>(nIdx, 1));
    if (!idx) return cgs::setError(CGS_ERR_OUT_OF_MEMORY, "cgs_load_be: out of host memory");
    for (size_t k = 0; k < nIdx; ++k)
    {
        uint32_t value;
        if (!in.readU32(value))
        {
            std::free(idx);
            return cgs::setError(CGS_ERR_INVALID, "Parsing patch failed");
        }
        idx[k] = value - 1; // zero-based; 0 wraps and fails the range check
        in.ignoreOne();
    }

    uint32_t nPoints;
    if (!in.readU32(nPoints))
    {
        std::free(idx);
        return cgs::setError(CGS_ERR_INVALID, "Could not read the no. vertices");
    }
    // three numbers per point with a separator after all but the last one
    if (uint64_t(nPoints) * 6 > uint64_t(length) + 1)
    {
        std::free(idx);
        return cgs::setError(CGS_ERR_INVALID, "Parsing point failed");
    }
    float* pts = (float*)std::malloc(sizeof(float) * 4 * std::max<size_t>(nPoints, 1));
    if (!pts)
    {
        std::free(idx);
        return cgs::setError(CGS_ERR_OUT_OF_MEMORY, "cgs_load_be: out of host memory");
    }
    for (uint32_t i = 0; i < nPoints; ++i)
    {
        for (int c = 0; c < 3; ++c)
        {
            if (!in.readFloat(pts[4 * i + c]))
            {
                std::free(idx);
                std::free(pts);
                return cgs::setError(CGS_ERR_INVALID, "Parsing point failed");
            }
            in.ignoreOne();
        }
        pts[4 * i + 3] = 1.0f;
    }

    // The reference accepts index == point count (its test is
    // `points.size() < i`), which reads one point past the end when the
    // patch is used; here every index must name a point.
    for (size_t k = 0; k < nIdx; ++k)
    {
        if (idx[k] >= nPoints)
        {
            std::free(idx);
            std::free(pts);
            return cgs::setError(CGS_ERR_INVALID, "Invalid/corrupted file. Indices out of range");
        }
    }
    *vertices = pts;
    *indices = idx;
    *npoints = nPoints;
    *npatches = nPatches;
    return CGS_OK;
}

// Mesh::buildBVH (rt/Primitive.cu:90-108): per triangle its index, the box of
// its three vertices and the centroid (1/3)·((p0 + p1) + p2), in FP32.  The
// reference builds the mesh BVH from these with 16 triangles per leaf.
cgs_status cgs_mesh_elements(const float* vertices, uint32_t nvertices, const uint32_t* indices, uint32_t ntriangles,
                             cgs_bvh_element* elements)
{
    if (ntriangles && (!vertices || !indices || !elements))
        return cgs::setError(CGS_ERR_INVALID, "cgs_mesh_elements: null argument");
    for (size_t k = 0; k < 3 * (size_t)ntriangles; ++k)
        if (indices[k] >= nvertices) return cgs::setError(CGS_ERR_INVALID, "cgs_mesh_elements: vertex index out of range");
    const float third = 1.0f / 3.0f;
    for (uint32_t i = 0; i < ntriangles; ++i)
    {
        const float* p[3] = {vertices + 3 * (size_t)indices[3 * (size_t)i], vertices + 3 * (size_t)indices[3 * (size_t)i + 1],
                             vertices + 3 * (size_t)indices[3 * (size_t)i + 2]};
        cgs_bvh_element& e = elements[i];
        for (int c = 0; c < 3; ++c)
        {
            e.bmin[c] = 3.402823466e+38f;    // the empty box, then inflate by each vertex
            e.bmax[c] = -3.402823466e+38f;
            for (int v = 0; v < 3; ++v)
            {
                if (p[v][c] < e.bmin[c]) e.bmin[c] = p[v][c];
                if (p[v][c] > e.bmax[c]) e.bmax[c] = p[v][c];
            }
            e.centroid[c] = third * ((p[0][c] + p[1][c]) + p[2][c]);
        }
        e.index = i;
    }
    return CGS_OK;
}
