// SPDX-License-Identifier: MIT
//
// Device-resident scene layout shared by the kernels and the host library.
//
// HBM layout (all arrays 256-B aligned allocations):
//   patches  : float4[16] per patch, control points pre-gathered patch-major
//              (256 B per patch; one patch test reads one contiguous block
//              instead of 16 index loads + 16 gathers)
//   nodes    : 64-B BVH nodes of every surface BVH, then the top-level BVH;
//              child offsets, leaf `first` slots and hash-table entries are
//              absolute (rebased at upload)
//   indices  : leaf element lists ([id, EMPTY] per leaf), ids local to the BVH
//   displ    : perfect-hash displacement tables, concatenated
//   tables   : perfect-hash node tables, concatenated
//   surfaces : one DevSurface per surface;  objects : one DevObject per object
//   cull*    : the cull tree over the top-level BVH's leaf boxes (optional)
//
#pragma once

#include <stdint.h>

#include "cgs_device.hpp"

namespace cgs
{

struct DevBVH
{
    uint32_t root;        // absolute node offset of the root
    uint32_t displBase;   // offset into displ
    uint32_t tableBase;   // offset into tables
    uint32_t A, B, mask;
    uint32_t T;           // table size
    uint32_t patchBase;   // first patch of the surface (surfaces only)
    uint64_t Tm;          // fastmod multiplier for T
    uint32_t finiteBoxes; // 1: every child box of an inner node is finite
    uint32_t wide;        // 1: hashed with the wide family (exact arithmetic)
    uint32_t cullRoot;    // root of its cull tree in the scene's cull arrays, or EMPTY
    uint32_t pad[3];
};
static_assert(sizeof(DevBVH) == 64, "DevBVH layout");

struct DevObject
{
    float4 w2l[4];        // world→local, column-major
    float4 mat0;          // diffuse rgb, metalness
    float4 mat1;          // specular rgb, roughness
    uint32_t surface;
    uint32_t pad[3];
};
static_assert(sizeof(DevObject) == 112, "DevObject layout");

struct DevLight
{
    float4 a;             // color rgb, strength
    float4 b;             // position xyz, range
    float4 c;             // direction xyz, angle
};

// Top-level nodes with heap key below kTopKeyCap are listed by key
// (topByKey[key] = absolute node offset, EMPTY where the tree has no such
// key) so kernels can stage the upper levels of the top-level BVH in LDS.
constexpr uint32_t kTopKeyCap = 4096;

struct SceneDev
{
    const float4* __restrict__ patches;
    const float4* __restrict__ nodes;
    const uint32_t* __restrict__ indices;
    const uint32_t* __restrict__ displ;
    const uint32_t* __restrict__ tables;
    const DevBVH* __restrict__ surfaces;
    const DevObject* __restrict__ objects;
    const DevLight* __restrict__ lights;
    const uint32_t* __restrict__ topByKey;   // kTopKeyCap entries
    // cull trees (cgs_cull.cpp): over the top-level leaf boxes where the
    // scene allows one (top.cullRoot), and over the leaf boxes of each large
    // surface (its cullRoot); null when there is none
    const float4* __restrict__ cullNodes;
    const uint2* __restrict__ cullLinks;
    const float4* __restrict__ cullElems;
    uint32_t nlight;
    uint32_t nobj;
    uint32_t nodeCount;
    uint32_t big;         // 1: some BVH is hashed with the wide family or some surface has a cull tree
    DevBVH top;
    float background[3];
    float ambient[3];
};

// One job of a batched launch: a resident scene and a ray range.  The jobs'
// rays are numbered consecutively (rayBase = rays of the jobs before it).
struct BatchJob
{
    SceneDev sc;
    const cgs_ray* rays;
    cgs_hit* hits;          // closest-hit jobs
    int32_t* objects;       // optional object ids (closest)
    uint32_t* occluded;     // any-hit jobs
    uint32_t n, rayBase, any, pad;
};

// One shading job: n world rays against one scene, each traced and shaded as
// the render traces and shades a camera ray (RayTracer.cu:194-544).
struct ShadeJob
{
    SceneDev sc;
    const cgs_ray* rays;
    float* color;           // 3 floats per ray (unclamped radiance)
    uint32_t* pixels;       // packed sRGB per ray, optional
    cgs_hit* hits;          // the first segment's closest hit, optional
    int32_t* objects;       // its object id, optional
    uint32_t n, rayBase;
    float eps;
    int32_t depth;
};

CGS_DEV BVHRef bvhRef(const SceneDev& sc, const DevBVH& d)
{
    BVHRef b;
    b.nodes = sc.nodes;
    b.indices = sc.indices;
    b.table = sc.tables + d.tableBase;
    b.displ = sc.displ + d.displBase;
    b.root = d.root;
    b.A = d.A;
    b.B = d.B;
    b.mask = d.mask;
    b.T = {d.Tm, d.T};
    b.nodeCount = sc.nodeCount;
    b.finiteBoxes = d.finiteBoxes;
    b.wide = d.wide;
    return b;
}

} // namespace cgs
