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

// ---------------------------------------------------------------------------
// Surface element files (`.se`, Surface::load_se, Surface.cpp:214-348) and
// the patch table they fill (PatchTable::set, Surface.cpp:36-148), with the
// boxes cg::PatchBVH gives their patches (PatchBVH.cpp:40-80).
// ---------------------------------------------------------------------------
namespace
{

// std::istream's `>>` of a uint16_t: digits, then the value must fit.
bool readU16(Cursor& in, uint16_t& v)
{
    uint32_t x;
    if (!in.readU32(x) || x > 0xffffu) return false;
    v = (uint16_t)x;
    return true;
}

// Signed decimal integer (`>> int`).
bool readI32(Cursor& in, int32_t& v)
{
    in.skipSpace();
    bool neg = false;
    if (in.p < in.end && (*in.p == '+' || *in.p == '-')) neg = *in.p++ == '-';
    uint32_t x;
    if (!in.readU32(x) || x > 0x80000000u || (!neg && x == 0x80000000u)) return false;
    v = neg ? (int32_t)(0u - x) : (int32_t)x;
    return true;
}

// Surface::load_se's discard(): '#' comments to the end of the line and
// whitespace.
void discard(Cursor& in)
{
    while (in.p < in.end)
    {
        if (*in.p == '#')
        {
            while (in.p < in.end && *in.p != '\n') ++in.p;
            if (in.p < in.end) ++in.p;
        }
        else if (std::isspace((unsigned char)*in.p))
            ++in.p;
        else
            return;
    }
}

struct SeElement
{
    uint32_t type;                 // CGS_PATCH_BSPLINE or CGS_PATCH_GREGORY
    std::vector<uint32_t> indices;
    std::vector<float> matrix;     // rows × nodes, column-major (row + col·rows)
};

uint32_t originalSize(uint32_t type) { return type == CGS_PATCH_GREGORY ? 20u : 16u; }

cgs_status seError(const char* msg) { return cgs::setError(CGS_ERR_INVALID, msg); }

} // namespace

extern "C" cgs_status cgs_load_se(const char* text, size_t length, cgs_patch_table** out)
{
    if (!out || (length && !text)) return cgs::setError(CGS_ERR_INVALID, "cgs_load_se: null argument");
    *out = nullptr;
    Cursor in {text, text + length};
    uint32_t nodeCount;
    discard(in);
    if (!in.readU32(nodeCount)) return seError("could not read no. points / nodes");
    // each node takes at least ten characters (five numbers and separators)
    if ((uint64_t)nodeCount * 10 > length) return seError("unexpected point id");
    std::vector<float> points(4 * (size_t)nodeCount);
    for (uint32_t i = 0; i < nodeCount; i++)
    {
        uint16_t id;
        discard(in);
        float* v = &points[4 * (size_t)i];
        if (!readU16(in, id) || !in.readFloat(v[0]) || !in.readFloat(v[1]) || !in.readFloat(v[2]) ||
            !in.readFloat(v[3]) || id != i)
            return seError("unexpected point id");
    }
    uint32_t elementCount;
    discard(in);
    if (!in.readU32(elementCount)) return seError("could not read no. elements");
    if ((uint64_t)elementCount * 10 > length) return seError("read failed during element header");
    std::vector<SeElement> elements(elementCount);
    for (uint32_t i = 0; i < elementCount; i++)
    {
        uint16_t id, typeRead, nodes, degree, boundary;
        discard(in);
        if (!readU16(in, id) || !readU16(in, typeRead) || !readU16(in, nodes) || !readU16(in, degree) ||
            !readU16(in, boundary))
            return seError("read failed during element header");
        if (degree != 3) return seError("degree value unsupported");
        SeElement& e = elements[i];
        uint32_t lines;
        if (typeRead == 1)
        {
            lines = boundary ? 16 : 0;
            e.type = CGS_PATCH_BSPLINE;
        }
        else if (typeRead == 2)
        {
            lines = 20;
            e.type = CGS_PATCH_GREGORY;
        }
        else
            return seError("unknown patch type");
        e.indices.resize(nodes);
        discard(in);
        for (uint32_t j = 0; j < nodes; j++)
            if (!in.readU32(e.indices[j]) || e.indices[j] >= nodeCount)
                return seError("Invalid/corrupted file. Indices out of range");
        discard(in);
        for (uint32_t j = 0; j < nodes; j++)
        {
            int32_t unknown;
            if (!readI32(in, unknown)) return seError("read failed during element data");
        }
        e.matrix.resize((size_t)lines * nodes);
        for (uint32_t j = 0; j < lines; j++)
        {
            discard(in);
            for (uint32_t k = 0; k < nodes; k++)
                if (!in.readFloat(e.matrix[j + (size_t)k * lines])) return seError("read failed during element data");
        }
    }

    // PatchTable::set: groups of one type, node count and matrix presence,
    // in order of first appearance; each group's elements in file order
    struct Group
    {
        uint32_t type, nodes;
        bool matrix;
        std::vector<uint32_t> members;
    };
    std::vector<Group> groups;
    for (uint32_t i = 0; i < elementCount; i++)
    {
        const SeElement& e = elements[i];
        const bool m = !e.matrix.empty();
        Group* g = nullptr;
        for (Group& q : groups)
            if (q.type == e.type && q.nodes == e.indices.size() && q.matrix == m)
            {
                g = &q;
                break;
            }
        if (!g)
        {
            groups.push_back({e.type, (uint32_t)e.indices.size(), m, {}});
            g = &groups.back();
        }
        if (m && e.matrix.size() != (size_t)originalSize(e.type) * e.indices.size())
            return seError("invalid element matrix size");
        g->members.push_back(i);
    }
    size_t nIdx = 0, nMat = 0;
    for (const SeElement& e : elements)
    {
        nIdx += e.indices.size();
        nMat += e.matrix.size();
    }
    auto* t = (cgs_patch_table*)std::calloc(1, sizeof(cgs_patch_table));
    if (!t) return cgs::setError(CGS_ERR_OUT_OF_MEMORY, "cgs_load_se: out of host memory");
    t->points = (float*)std::malloc(std::max<size_t>(1, points.size()) * 4);
    t->indices = (uint32_t*)std::malloc(std::max<size_t>(1, nIdx) * 4);
    t->matrices = (float*)std::malloc(std::max<size_t>(1, nMat) * 4);
    t->patches = (cgs_patch_ref*)std::malloc(std::max<size_t>(1, elementCount) * sizeof(cgs_patch_ref));
    if (!t->points || !t->indices || !t->matrices || !t->patches)
    {
        cgs_patch_table_free(t);
        return cgs::setError(CGS_ERR_OUT_OF_MEMORY, "cgs_load_se: out of host memory");
    }
    std::memcpy(t->points, points.data(), points.size() * 4);
    t->npoints = nodeCount;
    uint32_t off = 0, moff = 0, k = 0;
    for (const Group& g : groups)
        for (uint32_t i : g.members)
        {
            const SeElement& e = elements[i];
            std::memcpy(t->indices + off, e.indices.data(), e.indices.size() * 4);
            if (g.matrix) std::memcpy(t->matrices + moff, e.matrix.data(), e.matrix.size() * 4);
            t->patches[k++] = {off, g.matrix ? (int32_t)moff : -1, (uint32_t)e.indices.size(), e.type};
            off += (uint32_t)e.indices.size();
            moff += (uint32_t)e.matrix.size();
        }
    t->nindices = off;
    t->nmatrices = moff;
    t->npatches = elementCount;
    *out = t;
    return CGS_OK;
}

extern "C" void cgs_patch_table_free(cgs_patch_table* t)
{
    if (!t) return;
    std::free(t->points);
    std::free(t->indices);
    std::free(t->matrices);
    std::free(t->patches);
    std::free(t);
}

// PatchBVH::init (PatchBVH.cpp:40-80): a patch with a basis matrix is boxed
// by its n = 16 (B-spline) or 20 (Gregory) transformed points, row r being
// Σ_col M[r + col·n]·p[col] summed from col 0 in FP32 and projected by w;
// one without is boxed by its first n points.  The centroid is the box's
// centre.  A patch whose leaf test would read past the index table (16
// indices from its offset; the reference reads undefined memory there) gets
// the empty box, so no ray reaches it.
extern "C" cgs_status cgs_patch_table_elements(const cgs_patch_table* t, cgs_bvh_element* elements)
{
    if (!t || (t->npatches && !elements)) return cgs::setError(CGS_ERR_INVALID, "cgs_patch_table_elements: null argument");
    for (uint32_t i = 0; i < t->npatches; i++)
    {
        const cgs_patch_ref& p = t->patches[i];
        const uint32_t n = originalSize(p.type);
        const bool m = p.matrix_offset >= 0;
        if ((uint64_t)p.offset + (m ? p.size : n) > t->nindices ||
            (m && (uint64_t)p.matrix_offset + (uint64_t)n * p.size > t->nmatrices))
            return cgs::setError(CGS_ERR_INVALID, "cgs_patch_table_elements: patch " + std::to_string(i) +
                                                      " reads past the table");
        float mn[3] = {3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f};
        float mx[3] = {-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f};
        auto inflate = [&](const float* q) {
            for (int c = 0; c < 3; c++)
            {
                if (q[c] < mn[c]) mn[c] = q[c];
                if (q[c] > mx[c]) mx[c] = q[c];
            }
        };
        for (uint32_t r = 0; r < n; r++)
        {
            if (m)
            {
                const float* M = t->matrices + p.matrix_offset;
                float s[4];
                for (uint32_t col = 0; col < p.size; col++)
                {
                    const float* q = t->points + 4 * (size_t)t->indices[p.offset + col];
                    const float w = M[r + (size_t)col * n];
                    for (int c = 0; c < 4; c++) s[c] = col == 0 ? w * q[c] : s[c] + w * q[c];
                }
                const float pr[3] = {s[0] / s[3], s[1] / s[3], s[2] / s[3]};
                inflate(pr);
            }
            else
                inflate(t->points + 4 * (size_t)t->indices[p.offset + r]);
        }
        if ((uint64_t)p.offset + 16 > t->nindices)
            for (int c = 0; c < 3; c++)
            {
                mn[c] = 3.402823466e+38f;
                mx[c] = -3.402823466e+38f;
            }
        cgs_bvh_element& e = elements[i];
        for (int c = 0; c < 3; c++)
        {
            e.bmin[c] = mn[c];
            e.bmax[c] = mx[c];
            e.centroid[c] = 0.5f * (mn[c] + mx[c]);
        }
        e.index = i;
    }
    return CGS_OK;
}

// The 16 control points PatchBVH::intersectLeaf clips for each patch
// (PatchBVH.cpp:84-121): the 16 indices from the patch's offset, whatever
// its node count and type (so a patch of fewer nodes reads on into the
// next patch's).  A patch whose 16 would run past the index table (undefined
// in the reference) gets 16 zeros; its element box is empty, so no ray
// tests it.
extern "C" cgs_status cgs_patch_table_patches(const cgs_patch_table* t, uint32_t* indices16)
{
    if (!t || (t->npatches && !indices16)) return cgs::setError(CGS_ERR_INVALID, "cgs_patch_table_patches: null argument");
    for (uint32_t i = 0; i < t->npatches; i++)
    {
        const uint32_t off = t->patches[i].offset;
        if ((uint64_t)off + 16 > t->nindices)
            std::memset(indices16 + 16 * (size_t)i, 0, 64);
        else
            std::memcpy(indices16 + 16 * (size_t)i, t->indices + off, 64);
    }
    return CGS_OK;
}
