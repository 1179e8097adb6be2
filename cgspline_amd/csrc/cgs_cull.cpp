// SPDX-License-Identifier: MIT
//
// Cull trees: a surface-area-heuristic BVH over the leaf boxes of a scene's
// top-level BVH (or of a large surface's patch BVH), which the ray queries
// use to find the objects (patches) a ray can reach without walking the
// reference's median-split tree (whose split axis
// is always y, rt/BVH.cpp:46-52, so its boxes are thin y-slabs spanning the
// whole scene and a ray through the scene crosses most of them).  The tree
// only chooses which objects to test: every object test still uses the
// reference's box (the child box of its leaf, as the reference tests it) and
// the reference's per-object traversal, and the kernels fall back to the
// reference's traversal where the result could depend on the visit order
// (cgs_trace.hip, sceneClosest).
//
// Layout (device, 16-B aligned):
//   nodes : 2 float4 per node: (bmin.xyz, first element as bits),
//           (bmax.xyz, element count as bits; 0 for an inner node)
//   links : 8 uint2 per node, one per ray direction octant (bit i set: the
//           direction's component i is negative): (next node when the ray
//           enters this node, next node when it does not); EMPTY ends
//   elems : 2 float4 per element: (box min, object id as bits),
//           (box max, 0)
// An inner node's box is the exact union of its elements' boxes, so the
// reference's slab test passes on every ancestor of an element whose box it
// passes (IEEE subtraction, multiplication, min and max are monotone).
//
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "cgs_internal.hpp"

namespace cgs
{

namespace
{

struct Item
{
    float mn[3], mx[3], c[3];
    uint32_t id;
};

struct Node
{
    float mn[3], mx[3];
    uint32_t left = CGS_EMPTY, right = CGS_EMPTY;   // inner: children
    uint32_t first = 0, count = 0;                  // leaf: element range
    int axis = 0;
};

float area(const float* mn, const float* mx)
{
    const float a = mx[0] - mn[0], b = mx[1] - mn[1], c = mx[2] - mn[2];
    return 2.0f * (a * b + b * c + c * a);
}

struct Builder
{
    std::vector<Item>& items;
    std::vector<Node>& nodes;
    static constexpr int kBins = 16;
    static constexpr uint32_t kLeaf = 2;

    uint32_t build(uint32_t b, uint32_t e)
    {
        const uint32_t id = (uint32_t)nodes.size();
        nodes.emplace_back();
        Node n;
        float cmn[3], cmx[3];
        for (int a = 0; a < 3; a++)
        {
            n.mn[a] = cmn[a] = INFINITY;
            n.mx[a] = cmx[a] = -INFINITY;
        }
        for (uint32_t i = b; i < e; i++)
            for (int a = 0; a < 3; a++)
            {
                n.mn[a] = std::fmin(n.mn[a], items[i].mn[a]);
                n.mx[a] = std::fmax(n.mx[a], items[i].mx[a]);
                cmn[a] = std::fmin(cmn[a], items[i].c[a]);
                cmx[a] = std::fmax(cmx[a], items[i].c[a]);
            }
        const uint32_t count = e - b;
        int bestAxis = -1;
        int bestSplit = 0;
        float bestCost = (float)count * area(n.mn, n.mx);
        if (count > kLeaf)
        {
            for (int a = 0; a < 3; a++)
            {
                const float ext = cmx[a] - cmn[a];
                if (!(ext > 0.0f)) continue;
                float bmn[kBins][3], bmx[kBins][3];
                uint32_t cnt[kBins] = {};
                for (int k = 0; k < kBins; k++)
                    for (int q = 0; q < 3; q++)
                    {
                        bmn[k][q] = INFINITY;
                        bmx[k][q] = -INFINITY;
                    }
                for (uint32_t i = b; i < e; i++)
                {
                    int k = (int)((items[i].c[a] - cmn[a]) / ext * kBins);
                    k = k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
                    cnt[k]++;
                    for (int q = 0; q < 3; q++)
                    {
                        bmn[k][q] = std::fmin(bmn[k][q], items[i].mn[q]);
                        bmx[k][q] = std::fmax(bmx[k][q], items[i].mx[q]);
                    }
                }
                // sweep: areas of the bins right of each split
                float rA[kBins];
                uint32_t rN[kBins];
                float smn[3] = {INFINITY, INFINITY, INFINITY}, smx[3] = {-INFINITY, -INFINITY, -INFINITY};
                uint32_t sn = 0;
                for (int k = kBins - 1; k > 0; k--)
                {
                    for (int q = 0; q < 3; q++)
                    {
                        smn[q] = std::fmin(smn[q], bmn[k][q]);
                        smx[q] = std::fmax(smx[q], bmx[k][q]);
                    }
                    sn += cnt[k];
                    rA[k] = sn ? area(smn, smx) : 0.0f;
                    rN[k] = sn;
                }
                float lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
                uint32_t ln = 0;
                for (int k = 0; k < kBins - 1; k++)
                {
                    for (int q = 0; q < 3; q++)
                    {
                        lmn[q] = std::fmin(lmn[q], bmn[k][q]);
                        lmx[q] = std::fmax(lmx[q], bmx[k][q]);
                    }
                    ln += cnt[k];
                    if (!ln || !rN[k + 1]) continue;
                    const float cost = (float)ln * area(lmn, lmx) + (float)rN[k + 1] * rA[k + 1] +
                                       0.125f * area(n.mn, n.mx);
                    if (cost < bestCost)
                    {
                        bestCost = cost;
                        bestAxis = a;
                        bestSplit = k;
                    }
                }
            }
        }
        uint32_t mid = b;
        if (bestAxis >= 0)
        {
            const int a = bestAxis;
            const float ext = cmx[a] - cmn[a];
            auto it = std::partition(items.begin() + b, items.begin() + e, [&](const Item& x) {
                int k = (int)((x.c[a] - cmn[a]) / ext * kBins);
                k = k < 0 ? 0 : (k >= kBins ? kBins - 1 : k);
                return k <= bestSplit;
            });
            mid = (uint32_t)(it - items.begin());
        }
        else if (count > 4 * kLeaf)
        {
            // no useful split (coincident centroids): halve by order so the
            // leaves stay small
            bestAxis = 0;
            mid = b + count / 2;
        }
        if (bestAxis < 0 || mid == b || mid == e)
        {
            n.first = b;
            n.count = count;
            nodes[id] = n;
            return id;
        }
        n.axis = bestAxis;
        nodes[id] = n;
        const uint32_t l = build(b, mid);
        const uint32_t r = build(mid, e);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

// Per-octant threaded links: entering an inner node leads to its nearer
// child along its split axis (the child of lower coordinates when the
// direction is positive on that axis), leaving a subtree to the next one in
// that order.
void link(const std::vector<Node>& nodes, uint32_t n, int oct, uint32_t miss, std::vector<uint32_t>& out)
{
    const Node& x = nodes[n];
    uint32_t* l = &out[2 * (8 * (size_t)n + oct)];
    l[1] = miss;
    if (x.count)
    {
        l[0] = miss;
        return;
    }
    const bool neg = (oct >> x.axis) & 1;
    const uint32_t first = neg ? x.right : x.left, second = neg ? x.left : x.right;
    l[0] = first;
    link(nodes, first, oct, second, out);
    link(nodes, second, oct, miss, out);
}

} // namespace

bool buildCullTree(const cgs_bvh_node* refNodes, size_t nnodes, const uint32_t* indices, CullTreeHost& out)
{
    out = CullTreeHost {};
    if (nnodes < 3 || refNodes[0].left == CGS_EMPTY) return false;
    std::vector<Item> items;
    for (size_t i = 0; i < nnodes; i++)
    {
        const cgs_bvh_node& n = refNodes[i];
        if (n.left == CGS_EMPTY) continue;
        for (int side = 0; side < 2; side++)
        {
            const uint32_t c = side ? n.right : n.left;
            if (c >= nnodes || refNodes[c].left != CGS_EMPTY) continue;
            const float* mn = side ? n.rmin : n.lmin;
            const float* mx = side ? n.rmax : n.lmax;
            bool ok = true;
            for (int a = 0; a < 3; a++)
                ok = ok && std::isfinite(mn[a]) && std::isfinite(mx[a]) && mn[a] <= mx[a];
            if (!ok) continue;   // an empty box (an object without geometry): no ray enters it
            for (uint32_t s = refNodes[c].right; indices[s] != CGS_EMPTY; s++)
            {
                Item it;
                for (int a = 0; a < 3; a++)
                {
                    it.mn[a] = mn[a];
                    it.mx[a] = mx[a];
                    it.c[a] = 0.5f * (mn[a] + mx[a]);
                }
                it.id = indices[s];
                items.push_back(it);
            }
        }
    }
    if (items.empty()) return false;
    std::vector<Node> nodes;
    nodes.reserve(2 * items.size());
    Builder bld {items, nodes};
    bld.build(0, (uint32_t)items.size());
    std::vector<uint32_t> links(16 * nodes.size());
    for (int oct = 0; oct < 8; oct++) link(nodes, 0, oct, CGS_EMPTY, links);
    out.nodes.resize(8 * nodes.size());
    for (size_t i = 0; i < nodes.size(); i++)
    {
        float* q = &out.nodes[8 * i];
        const Node& n = nodes[i];
        std::memcpy(q, n.mn, 12);
        std::memcpy(q + 3, &n.first, 4);
        std::memcpy(q + 4, n.mx, 12);
        std::memcpy(q + 7, &n.count, 4);
    }
    out.links = std::move(links);
    out.elems.resize(8 * items.size());
    for (size_t i = 0; i < items.size(); i++)
    {
        float* q = &out.elems[8 * i];
        std::memcpy(q, items[i].mn, 12);
        std::memcpy(q + 3, &items[i].id, 4);
        std::memcpy(q + 4, items[i].mx, 12);
        q[7] = 0.0f;
    }
    out.nodeCount = (uint32_t)nodes.size();
    return true;
}

uint32_t appendCullTree(CullTreeHost& all, const CullTreeHost& t)
{
    const uint32_t nodeBase = all.nodeCount, elemBase = (uint32_t)(all.elems.size() / 8);
    const size_t n0 = all.nodes.size(), l0 = all.links.size();
    all.nodes.insert(all.nodes.end(), t.nodes.begin(), t.nodes.end());
    all.links.insert(all.links.end(), t.links.begin(), t.links.end());
    all.elems.insert(all.elems.end(), t.elems.begin(), t.elems.end());
    for (size_t i = 0; i < t.nodeCount; i++)
    {
        uint32_t first;
        std::memcpy(&first, &all.nodes[n0 + 8 * i + 3], 4);
        first += elemBase;
        std::memcpy(&all.nodes[n0 + 8 * i + 3], &first, 4);
    }
    for (size_t i = l0; i < all.links.size(); i++)
        if (all.links[i] != CGS_EMPTY) all.links[i] += nodeBase;
    all.nodeCount += t.nodeCount;
    return nodeBase;
}

} // namespace cgs
