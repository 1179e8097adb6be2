// SPDX-License-Identifier: MIT
//
// Declarations shared between the host library and the kernel translation
// units (not part of the public ABI).
//
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <bit>
#include <string>
#include <vector>

#include "../../include/cgs.h"

namespace cgs
{

// Records `msg` as the calling thread's cgs_last_error() and returns `s`.
cgs_status setError(cgs_status s, const std::string& msg);

struct SceneDev;
struct BatchJob;
struct ShadeJob;

// One perfect-hash build (one BVH's key set).
struct PhfJob
{
    uint32_t keyOff, n, M, mask;
    uint32_t scratchOff;   // per-job scratch words: phfScratchWords(M, tcap, key words)
    uint32_t tcap;         // capacity for the growing table: n + 16·step
    uint32_t displOff;     // output displacement (M words)
    uint32_t tableOff;     // output table (tcap words)
    uint32_t step;         // table growth per failed function
    uint32_t abOff;        // 16 (A,B) pairs drawn by the host
    uint32_t nodeBase;     // value added to every table entry
    uint32_t pad;
};

struct PhfResult
{
    uint32_t A, B, T, attempts, status;
};

// Scratch words of one build: 9·M words of bucket state, 4·M keys of
// keyWords words each, 2·tcap words of slot state.
constexpr uint32_t phfScratchWords(uint32_t M, uint32_t tcap, uint32_t keyWords = 1)
{
    return (9u + 4u * keyWords) * M + 2u * tcap;
}

// Displacement-table size of n keys: bit_floor(max(2, (n-1) >> 1))
// (PerfectHashFunction.cu:234-248).
constexpr uint32_t phfDisplSize(uint64_t n)
{
    int s = int(n - 1) >> 1;
    if (s < 2) s = 2;
    return std::bit_floor((unsigned)s);
}
// Table growth after a failed function: max(4, 0.05·n) (PerfectHashFunction.cu:251,481).
constexpr uint32_t phfGrowthStep(uint64_t n) { return std::max<uint32_t>(4u, (uint32_t)(0.05f * (float)n)); }

// Device scratch of cgs_phf_build_static, a file-scope array in cgs_phf.hip
// sized for kStaticPhfMaxKeys keys, in words: the job record, the 16 (A,B)
// pairs, the result, the build's working arrays and the growing table.
constexpr uint32_t kStaticPhfMaxKeys = 1u << 19;
constexpr uint32_t kStaticPhfM = phfDisplSize(kStaticPhfMaxKeys);
constexpr uint32_t kStaticPhfTcap = kStaticPhfMaxKeys + 16u * phfGrowthStep(kStaticPhfMaxKeys);
constexpr uint32_t kStaticPhfJobWord = 0, kStaticPhfAbWord = 16, kStaticPhfResultWord = 48, kStaticPhfScratchWord = 64;
constexpr uint32_t kStaticPhfScratchWords = phfScratchWords(kStaticPhfM, kStaticPhfTcap);
constexpr uint32_t kStaticPhfTableWord = (kStaticPhfScratchWord + kStaticPhfScratchWords + 63u) & ~63u;
constexpr uint32_t kStaticPhfWords = kStaticPhfTableWord + kStaticPhfTcap;
static_assert(sizeof(PhfJob) <= 4 * (kStaticPhfAbWord - kStaticPhfJobWord), "static scratch: job record");
static_assert(sizeof(PhfResult) <= 4 * (kStaticPhfScratchWord - kStaticPhfResultWord), "static scratch: result record");

// Records the scratch array's address on `device` (the current device) once
// per process; cgs_open calls it.  phfStaticScratch returns that address, or
// null before the first cgs_open on the device.
hipError_t phfStaticInit(int device);
uint32_t* phfStaticScratch(int device);

// Perfect-hash builds of 32-bit keys, with the reference's family or (wide)
// the same function in exact arithmetic; and of 64-bit keys (wide family).
hipError_t launchPhfBuild(const PhfJob* jobs, uint32_t njobs, const uint32_t* keys, const uint32_t* ab, uint32_t* scratch,
                          uint32_t* displ, uint32_t* tables, PhfResult* results, hipStream_t s, bool wide = false);
hipError_t launchPhfBuildWide64(const PhfJob* jobs, uint32_t njobs, const uint64_t* keys, const uint32_t* ab,
                                uint32_t* scratch, uint32_t* displ, uint32_t* tables, PhfResult* results, hipStream_t s);
hipError_t launchPhfHash(const cgs_phf& h, const uint32_t* D, const uint32_t* keys, uint32_t n, uint32_t* out, hipStream_t s);
hipError_t launchPhfHashWide64(const cgs_phf& h, const uint32_t* D, const uint64_t* keys, uint32_t n, uint32_t* out,
                               hipStream_t s);
hipError_t launchPhfFill(const cgs_phf& h, const uint32_t* D, const uint32_t* keys, uint32_t n, uint32_t base,
                         uint32_t* table, hipStream_t s);
hipError_t launchPhfFillWide64(const cgs_phf& h, const uint32_t* D, const uint64_t* keys, uint32_t n, uint32_t* table,
                               hipStream_t s);
hipError_t launchFill(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

// A cull tree over the leaf boxes of a BVH (cgs_cull.cpp): nodes 8 floats,
// links 16 words and elements 8 floats each.  buildCullTree returns false
// (and builds nothing) when the BVH's root is a leaf or no leaf box is finite.
struct CullTreeHost
{
    std::vector<float> nodes;
    std::vector<uint32_t> links;
    std::vector<float> elems;
    uint32_t nodeCount = 0;
};
bool buildCullTree(const cgs_bvh_node* refNodes, size_t nnodes, const uint32_t* indices, CullTreeHost& out);
// Appends tree t to the combined arrays `all` (node links and element starts
// rebased); returns t's root node in `all`.
uint32_t appendCullTree(CullTreeHost& all, const CullTreeHost& t);
// Surfaces of at least this many patches get a cull tree of their own.
constexpr uint32_t kSurfaceCullMinPatches = 4096;

// Persistent kernels: the grid is the resident workgroups; every lane owns a
// slice of the region-stack buffer and waves pull work from a counter.
enum class PersistKernel
{
    TraceClosest,
    TraceAny,
    ClipBatch,
    Render,
    Batch,
    Pick,
    Shade,
};
constexpr int kPersistKernels = 7;

struct PersistArgs
{
    float2* stack;          // lanes × CGS_CLIP_STACK × 20 float2
    uint32_t lanes;         // grid × 256
    unsigned int* counter;  // zeroed before the launch
    uint32_t grid;
};

hipError_t persistentGrid(PersistKernel k, uint32_t& blocksPerCU);

// Leaf method of a ray query (cgs_trace_options): the default is Bézier
// clipping at the reference's tolerance, unpolished and uncapped.
struct TraceMode
{
    uint32_t method = CGS_METHOD_CLIPPING;
    uint32_t polish = 0;
    float tol = 0x1p-12f;
    float threshold = 0x1p-4f;
    uint32_t maxIters = 0;
};

hipError_t launchTraceClosest(const SceneDev& sc, const PersistArgs& pa, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                              int32_t* objects, unsigned long long* overflow, hipStream_t s, const TraceMode& m);
hipError_t launchTraceAny(const SceneDev& sc, const PersistArgs& pa, const cgs_ray* rays, uint32_t n, uint32_t* occluded,
                          unsigned long long* overflow, hipStream_t s, const TraceMode& m);
// Batched queries with leaf method `m` (closest-hit jobs polished when it
// asks); the counter must be zero at launch (cleared with a fill kernel, so
// the pair can be captured in a graph).
// big: some job's scene has a wide-hashed BVH or a surface cull tree.
hipError_t launchTraceBatch(const BatchJob* jobs, uint32_t njobs, uint32_t total, uint32_t chunk, float2* stack,
                            uint32_t lanes, unsigned int* counter, uint32_t grid, unsigned long long* overflow, hipStream_t s,
                            const TraceMode& m, bool big);
hipError_t launchClipBatch(const PersistArgs& pa, const cgs_ray* rays, const float* patches, uint32_t n, float tol,
                           cgs_hit* hits, hipStream_t s);
hipError_t launchSubdivBatch(const cgs_ray* rays, const float* patches, uint32_t n, float threshold, cgs_hit* hits,
                             hipStream_t s);
hipError_t launchPatchEval(uint32_t kind, const float* patches, const float* uv, uint32_t n, float* out, hipStream_t s);
hipError_t launchSubpatch(const float* patches, const float* ranges, uint32_t n, float* out, hipStream_t s);
// PatchBVH queries against surface 0 of `sc`: closest hits into `hits`, or
// (occluded non-null) any hits into `occluded`.
hipError_t launchPick(const SceneDev& sc, const PersistArgs& pa, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                      uint32_t* occluded, hipStream_t s);
// Shaded ray queries over `jobs` (consecutive ray numbers, `total` rays); with
// `one` non-null every job is on that scene.  The counter must be zero at
// launch.  counts (nullable): cgs_ray_counts, accumulated.
hipError_t launchShade(const SceneDev* one, const ShadeJob* jobs, uint32_t njobs, uint32_t total, uint32_t chunk,
                       float2* stack, uint32_t lanes, unsigned int* counter, uint32_t grid, unsigned long long* counts,
                       hipStream_t s);
// kRender reads its camera, options and output pointers from `argsDev`, a
// device block of kRenderArgsBytes that launchRender fills on stream `s`.
constexpr size_t kRenderArgsBytes = 256;
// `order` (rank -> tile) sets the order in which the waves take the tiles
// (null: row by row); `tileTicks`, when set, receives each tile's time.
// samples > 0: the sampled (Monte Carlo) render with that many camera rays
// per pixel and draws from `seed`.
hipError_t launchRender(const SceneDev& sc, const PersistArgs& pa, const float* toWorld16, const float* pos3, float tlx,
                        float tly, float hdx, float hdy, float eps, int depth, uint32_t w, uint32_t h, uint32_t* frame,
                        float* color, unsigned long long* counts, uint32_t stride, const uint32_t* order,
                        uint32_t* tileTicks, void* argsDev, hipStream_t s, uint32_t samples = 0, uint32_t seed = 0);
// The sampled render's kernel (its own translation unit, cgs_render_sampled.hip:
// in the plain render kernels' code object it raised their register spills
// ninefold); argsDev as filled by launchRender.
hipError_t launchRenderSampled(const SceneDev& sc, const PersistArgs& pa, const void* argsDev, hipStream_t s);
// The render's work-event totals (a CGS_WORK_COUNTS=1 build; zeros
// otherwise), kWorkEvents of them (cgs_device.hpp); cleared when `reset`.
hipError_t workCounts(unsigned long long* out, bool reset);
// order[0..n) = the tiles 0..n-1, longest `ticks` first.
hipError_t launchTileOrder(const uint32_t* ticks, uint32_t n, uint32_t* order, hipStream_t s);

} // namespace cgs
