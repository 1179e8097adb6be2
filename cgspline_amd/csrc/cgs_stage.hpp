// SPDX-License-Identifier: MIT
//
// Data of the staged tracer (structure of arrays in HBM) and its launchers.
//
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cgs.h"

namespace cgs
{

struct SceneDev;

// Resumable per-ray traversal state, one slot per ray (16-B fields so a
// wave's accesses to one field are one contiguous 1-KB run).
struct RayBatch
{
    float4* od0;    // origin xyz, tmax
    float4* od1;    // direction xyz, best t (world)
    float4* hit;    // u, v, patch id bits, object id bits (EMPTY = none)
    uint4* top;     // top-level node, key, postponed bits, leaf position
    uint4* bot;     // object-level node, key, postponed bits, leaf position
    float4* lr0;    // local origin xyz, local max
    float4* lr1;    // local direction xyz, 1/|M d|
    float4* lh;     // local best t, u, v, patch id bits
    uint32_t* flags;
    uint32_t* obj;  // object being traversed
    uint32_t* pixel;// render: pixel (closest rays) or pixel·32 + light (shadow rays)
};

struct ClipQueue
{
    uint2* req;            // (ray slot, patch id within the object's surface)
    unsigned int* count;
};

// Per-pixel shading state of the render passes.
struct PixelState
{
    float4* acc;      // accumulated colour
    float4* att;      // attenuation of the current bounce
    float4* attNext;  // attenuation of the next bounce
    float4* base;     // ambient term of the current hit
    float4* term;     // per (pixel, light) lit term
    uint32_t* state;  // bit 31: hit this bounce; bit l: light l pending
    uint32_t nlight;
};

struct CamDevS
{
    float4 c[4];
    float px, py, pz;
    float tlx, tly, hdx, hdy;
};

hipError_t stageGrid(int which, uint32_t& blocksPerCU);
hipError_t launchTraverse(const SceneDev& sc, const RayBatch& rb, const uint32_t* active, const unsigned int* nActive,
                          const ClipQueue& cq, unsigned int* work, uint32_t grid, hipStream_t s);
hipError_t launchFinish(const SceneDev& sc, const RayBatch& rb, const uint32_t* active, const unsigned int* nActive,
                        unsigned int* work, float2* stack, uint32_t lanes, uint32_t grid, hipStream_t s);
hipError_t launchClip(const SceneDev& sc, const RayBatch& rb, const ClipQueue& cq, float2* stack, uint32_t lanes,
                      uint32_t* nextActive, unsigned int* nNext, unsigned int* work, uint32_t grid, hipStream_t s);
hipError_t launchLoadRays(const SceneDev& sc, const RayBatch& rb, const cgs_ray* rays, uint32_t n, int any,
                          uint32_t* active, hipStream_t s);
hipError_t launchStoreClosest(const RayBatch& rb, uint32_t n, cgs_hit* hits, int32_t* objects,
                              unsigned long long* overflow, hipStream_t s);
hipError_t launchStoreAny(const RayBatch& rb, uint32_t n, uint32_t* occluded, unsigned long long* overflow, hipStream_t s);
hipError_t launchGenPrimary(const SceneDev& sc, const RayBatch& rb, const CamDevS& cam, uint32_t w, uint32_t h,
                            const PixelState& px, uint32_t* active, unsigned int* nActive, unsigned long long* counts,
                            hipStream_t s);
hipError_t launchShade(const SceneDev& sc, const RayBatch& rb, uint32_t n, const PixelState& px, int depth, float eps,
                       const RayBatch& next, uint32_t* nextActive, unsigned int* nNext, const RayBatch& shadow,
                       uint32_t* shadowActive, unsigned int* nShadow, unsigned long long* counts, hipStream_t s);
hipError_t launchScatterShadow(const RayBatch& shadow, uint32_t n, uint32_t* occ, unsigned long long* counts,
                               hipStream_t s);
hipError_t launchAccumulate(const RayBatch& rb, uint32_t n, const PixelState& px, const uint32_t* occ, hipStream_t s);
hipError_t launchAdvanceAtt(const RayBatch& next, uint32_t n, const PixelState& px, hipStream_t s);
hipError_t launchPack(const PixelState& px, uint32_t npix, uint32_t width, uint32_t stride, uint32_t* frame, float* color,
                      hipStream_t s);
hipError_t launchZero(uint32_t* p, size_t n, hipStream_t s);

} // namespace cgs
