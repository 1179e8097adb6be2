// SPDX-License-Identifier: MIT
//
// Host side of libcgs: the C ABI (include/cgs.h), the BVH build
// (rt/BVH.cpp) and the scene upload (MainWindow::convertScene).
//
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <bit>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <random>
#include <span>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "cgs_internal.hpp"
#include "cgs_limits.hpp"
#include "cgs_scene_dev.hpp"

using namespace cgs;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_lastError;

static cgs_status fail(cgs_status s, const std::string& msg)
{
    g_lastError = msg;
    return s;
}
cgs_status cgs::setError(cgs_status s, const std::string& msg) { return fail(s, msg); }

#define HIP_TRY(expr)                                                                        \
    do                                                                                       \
    {                                                                                        \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return fail(CGS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)

struct DeviceBuffer
{
    void* p = nullptr;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        return hipMalloc(&p, std::max<size_t>(bytes, 256));
    }
    template<class T> T* as() const { return (T*)p; }
};

// Tile scheduling state of one scene's renders through one context: each
// render records its tiles' times, and the next render of a frame of the same
// size takes the tiles longest first by them (an earlier view's order is still
// a good guess for a nearby one).  Results do not depend on the order.  Kept
// per context, so two contexts may render one scene at the same time.
struct TileState
{
    uint64_t scene = 0;            // cgs_scene::serial
    DeviceBuffer ticks, order;
    uint32_t cap = 0, w = 0, h = 0;
    bool ready = false;
    uint64_t used = 0;             // last use (eviction order)
};

struct cgs_ctx
{
    int device = 0;
    int numCUs = 0;
    hipEvent_t renderDone = nullptr;
    bool renderLaunched = false;
    // Persistent launches of this context share its region stacks, work
    // counters and render argument block.  Each one waits (on the GPU) for
    // the previous one, whatever stream either was issued on, so calls on
    // different streams never overlap on those buffers.
    hipEvent_t persistDone = nullptr;
    bool persistLaunched = false;
    // batched queries: launch shape, job table, completion of the last launch
    uint32_t batchChunk = 64, batchBlocksPerCU = 0;
    TraceMode batchMode;                     // leaf method of batched launches
    cgs_trace_options batchOptions {};       // ... as the caller set it (zero: the default)
    DeviceBuffer batchTable;
    size_t batchTableCap = 0;
    hipEvent_t batchDone = nullptr;
    bool batchLaunched = false;
    unsigned long long* dOverflow = nullptr; // device counter of overflowed clip stacks
    unsigned int* dCounters = nullptr;       // one work counter per persistent kernel
    void* dRenderArgs = nullptr;             // kRender's argument block (one render in flight)
    ShadeJob* dShadeJob = nullptr;           // cgs_shade's job (persistent launches run one at a time)
    float2* dStack = nullptr;                // clipping region stacks of the resident lanes
    size_t stackLanes = 0;
    uint32_t grid[kPersistKernels] = {};
    std::vector<std::unique_ptr<TileState>> tiles;
    uint64_t tileClock = 0;
    unsigned int* hCount = nullptr;          // pinned host word for count read-backs
    std::mutex mu;
};

// Grid, stack slice and zeroed work counter for one persistent launch over
// `items` work items.  The caller holds ctx->mu, launches the kernel on `s`
// and then calls persistEnd.  The counter grows by 64 per wave fetch past
// the last item, so items + 64 × (resident waves) must stay below 2^32.
static cgs_status persistArgs(cgs_ctx* ctx, PersistKernel k, hipStream_t s, uint64_t items, PersistArgs& pa)
{
    const int ki = (int)k;
    if (ctx->grid[ki] == 0)
    {
        uint32_t bpc = 1;
        HIP_TRY(persistentGrid(k, bpc));
        ctx->grid[ki] = bpc * (uint32_t)ctx->numCUs;
    }
    if (!counterFits(items, ctx->grid[ki], 64))
        return fail(CGS_ERR_INVALID, "work of " + std::to_string(items) + " items exceeds the 32-bit work counter");
    const size_t lanes = (size_t)ctx->grid[ki] * 256;
    if (lanes > ctx->stackLanes)
    {
        if (ctx->dStack)
        {
            HIP_TRY(hipDeviceSynchronize());
            HIP_TRY(hipFree(ctx->dStack));
            ctx->dStack = nullptr;
            ctx->stackLanes = 0;
        }
        HIP_TRY(hipMalloc(&ctx->dStack, lanes * CGS_CLIP_STACK * kStackWords * sizeof(float2)));
        ctx->stackLanes = lanes;
    }
    pa.stack = ctx->dStack;
    pa.lanes = (uint32_t)lanes;
    pa.counter = ctx->dCounters + ki;
    pa.grid = ctx->grid[ki];
    if (ctx->persistLaunched) HIP_TRY(hipStreamWaitEvent(s, ctx->persistDone, 0));
    HIP_TRY(launchFill(pa.counter, 0u, 1, s));
    return CGS_OK;
}

// Marks the end of a persistent launch on `s` (the next one waits for it).
static cgs_status persistEnd(cgs_ctx* ctx, hipStream_t s)
{
    HIP_TRY(hipEventRecord(ctx->persistDone, s));
    ctx->persistLaunched = true;
    return CGS_OK;
}

// The tile state of `serial` in this context (created on first use; the
// least recently used of more than 16 is dropped).
static TileState& tileState(cgs_ctx* ctx, uint64_t serial)
{
    for (auto& t : ctx->tiles)
        if (t->scene == serial)
        {
            t->used = ++ctx->tileClock;
            return *t;
        }
    if (ctx->tiles.size() >= 16)
    {
        auto lru = std::min_element(ctx->tiles.begin(), ctx->tiles.end(),
                                    [](const auto& a, const auto& b) { return a->used < b->used; });
        ctx->tiles.erase(lru);
    }
    ctx->tiles.push_back(std::make_unique<TileState>());
    TileState& t = *ctx->tiles.back();
    t.scene = serial;
    t.used = ++ctx->tileClock;
    return t;
}

// ---------------------------------------------------------------------------
// BVH build (rt/BVH.cpp:9-108)
// ---------------------------------------------------------------------------
namespace
{

constexpr float kFMax = std::numeric_limits<float>::max();

struct Box3
{
    float mn[3] {kFMax, kFMax, kFMax};
    float mx[3] {-kFMax, -kFMax, -kFMax};
    void inflate(const float* p)
    {
        for (int i = 0; i < 3; i++)
        {
            if (p[i] < mn[i]) mn[i] = p[i];
            if (p[i] > mx[i]) mx[i] = p[i];
        }
    }
    void inflate(const Box3& b)
    {
        inflate(b.mn);
        inflate(b.mx);
    }
};

struct BuiltBVH
{
    std::vector<cgs_bvh_node> nodes;
    std::vector<uint32_t> keys;
    std::vector<uint32_t> indices;
};

// Node and index-slot counts of a median-split subtree of c elements at the
// given depth: they depend only on (c, depth), so a subtree's position in the
// pre-order arrays is known before its siblings are built, and subtrees can
// be built concurrently with the sequential build's exact output.
struct Sizes
{
    uint32_t epn;
    std::map<std::pair<uint64_t, int>, std::pair<uint64_t, uint64_t>> memo;
    // Fills the memo for every (count, depth) reachable from the root.
    std::pair<uint64_t, uint64_t> fill(uint64_t c, int depth)
    {
        if (c <= epn || depth >= 31) return {1, c + 1};
        auto it = memo.find({c, depth});
        if (it != memo.end()) return it->second;
        auto l = fill(c / 2, depth + 1);
        auto r = fill(c - c / 2, depth + 1);
        std::pair<uint64_t, uint64_t> v {1 + l.first + r.first, l.second + r.second};
        memo[{c, depth}] = v;
        return v;
    }
    // Read-only lookup, safe from concurrent subtree builds after fill().
    std::pair<uint64_t, uint64_t> operator()(uint64_t c, int depth) const
    {
        if (c <= epn || depth >= 31) return {1, c + 1};
        return memo.at({c, depth});
    }
};

struct Builder
{
    BuiltBVH& out;
    uint32_t epn;
    Sizes& sizes;
    std::atomic<int>& spawned;

    // BVH.cpp:71-78: a child's box is its left box, inflated by its right
    // box when the child is an inner node.
    static void nodeBoxOf(const cgs_bvh_node& n, Box3& b)
    {
        std::memcpy(b.mn, n.lmin, 12);
        std::memcpy(b.mx, n.lmax, 12);
        if (n.left != CGS_EMPTY)
        {
            b.inflate(n.rmin);
            b.inflate(n.rmax);
        }
    }

    // Restates BVH::split / BVH::wrap with the node written at `ni`, its
    // index slots at `ii`.
    void build(cgs_bvh_element* el, uint64_t count, uint32_t key, int depth, uint64_t ni, uint64_t ii)
    {
        cgs_bvh_node& node = out.nodes[ni];
        out.keys[ni] = key;
        if (count <= epn || (key & (1u << 31)))
        {
            Box3 box;
            for (uint64_t i = 0; i < count; i++)
            {
                out.indices[ii + i] = el[i].index;
                box.inflate(el[i].bmin);
                box.inflate(el[i].bmax);
            }
            out.indices[ii + count] = CGS_EMPTY;
            Box3 empty;
            std::memcpy(node.lmin, box.mn, 12);
            std::memcpy(node.lmax, box.mx, 12);
            std::memcpy(node.rmin, empty.mn, 12);
            std::memcpy(node.rmax, empty.mx, 12);
            node.left = CGS_EMPTY;
            node.right = (uint32_t)ii;
            node.pad[0] = node.pad[1] = 0;
            return;
        }
        // the split axis comes from a never-inflated box: every component of
        // its size is equal, so the comparison chain always selects Y
        Box3 nobox;
        float v[3] = {nobox.mx[0] - nobox.mn[0], nobox.mx[1] - nobox.mn[1], nobox.mx[2] - nobox.mn[2]};
        const int dim = v[2] > v[1] ? (v[0] > v[2] ? 0 : 2) : (v[0] > v[1] ? 0 : 1);
        const uint64_t middle = count / 2;
        std::nth_element(el, el + middle, el + count,
                         [dim](const cgs_bvh_element& a, const cgs_bvh_element& b) {
                             return a.centroid[dim] < b.centroid[dim];
                         });
        const auto ls = sizes(middle, depth + 1);
        const uint64_t L = ni + 1, R = L + ls.first;
        const uint64_t IL = ii, IR = ii + ls.second;
        const uint32_t kl = key << 1, kr = (key << 1) ^ 1u;
        if (count > (1u << 15) && spawned.fetch_add(1) < 64)
        {
            auto f = std::async(std::launch::async, [&] { build(el, middle, kl, depth + 1, L, IL); });
            build(el + middle, count - middle, kr, depth + 1, R, IR);
            f.get();
        }
        else
        {
            build(el, middle, kl, depth + 1, L, IL);
            build(el + middle, count - middle, kr, depth + 1, R, IR);
        }
        Box3 lb, rb;
        nodeBoxOf(out.nodes[L], lb);
        nodeBoxOf(out.nodes[R], rb);
        std::memcpy(node.lmin, lb.mn, 12);
        std::memcpy(node.lmax, lb.mx, 12);
        std::memcpy(node.rmin, rb.mn, 12);
        std::memcpy(node.rmax, rb.mx, 12);
        node.left = (uint32_t)L;
        node.right = (uint32_t)R;
        node.pad[0] = node.pad[1] = 0;
    }
};

cgs_status buildBVH(std::vector<cgs_bvh_element>& el, uint32_t epn, BuiltBVH& out)
{
    if (el.empty()) return fail(CGS_ERR_EMPTY_SPLIT, "BVH build over an empty element range");
    if (epn == 0) epn = 1;
    Sizes sizes {epn, {}};
    auto total = sizes.fill(el.size(), 0);
    out.nodes.assign(total.first, cgs_bvh_node {});
    out.keys.assign(total.first, 0);
    out.indices.assign(total.second, 0);
    std::atomic<int> spawned {0};
    Builder b {out, epn, sizes, spawned};
    b.build(el.data(), el.size(), 1u, 0, 0, 0);
    return CGS_OK;
}

// BezierSurface::buildBVH element data (rt/Primitive.cu:199-220).
void patchElements(const float* verts, const uint32_t* idx, uint32_t npatch, cgs_bvh_element* out)
{
    for (uint32_t i = 0; i < npatch; i++)
    {
        Box3 b;
        float c[3] = {0.0f, 0.0f, 0.0f};
        const float inv = 1.0f / 16.0f;
        for (int j = 0; j < 16; j++)
        {
            const float* w = verts + 4 * (size_t)idx[16 * (size_t)i + j];
            float p[3] = {w[0] / w[3], w[1] / w[3], w[2] / w[3]};
            b.inflate(p);
            for (int k = 0; k < 3; k++) c[k] = c[k] + inv * p[k];
        }
        std::memcpy(out[i].bmin, b.mn, 12);
        std::memcpy(out[i].bmax, b.mx, 12);
        std::memcpy(out[i].centroid, c, 12);
        out[i].index = i;
    }
}

// PerfectHashFunction's (A,B) draws (rt/PerfectHashFunction.h:45-46,
// .cu:386-400): mt19937(seed), uniform_int_distribution<uint32_t>(0, n),
// A = 0 replaced by 31, redrawn while gcd(n, A) > 1.  All 16 functions'
// pairs, in order, into ab[0..32).
void drawAB(uint32_t n, uint32_t seed, uint32_t* ab)
{
    std::mt19937 mt(seed);
    std::uniform_int_distribution<uint32_t> d(0, n);
    for (int fn = 0; fn < 16; fn++)
    {
        uint32_t A, B;
        do
        {
            A = d(mt);
            if (A == 0) A = 31;
            B = d(mt);
        } while (std::gcd((size_t)n, (size_t)A) > 1);
        ab[2 * fn] = A;
        ab[2 * fn + 1] = B;
    }
}

std::vector<uint32_t> drawAB(uint32_t n, uint32_t seed)
{
    std::vector<uint32_t> ab(32);
    drawAB(n, seed, ab.data());
    return ab;
}

uint32_t displSize(uint64_t n) { return phfDisplSize(n); }

uint32_t growthStep(uint64_t n) { return phfGrowthStep(n); }

template<class F> void parallelFor(size_t n, F&& f)
{
    unsigned th = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 2 || th == 1)
    {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<size_t> next {0};
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < th; t++)
        ts.emplace_back([&] {
            size_t i;
            while ((i = next.fetch_add(1)) < n) f(i);
        });
    for (auto& t : ts) t.join();
}


// Batched PHF build: jobs over keys already on the device.
struct PhfBatch
{
    std::vector<PhfJob> jobs;
    std::vector<uint32_t> ab;
    uint64_t scratchWords = 0, displWords = 0, tableWords = 0;
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> abCache;

    // Adds a job; returns its index.
    uint32_t add(uint32_t keyOff, uint32_t n, uint32_t seed, uint32_t nodeBase)
    {
        PhfJob j {};
        j.keyOff = keyOff;
        j.n = n;
        j.M = displSize(n);
        j.mask = j.M - 1;
        j.step = growthStep(n);
        j.tcap = n + 16 * j.step;
        j.scratchOff = (uint32_t)scratchWords;
        scratchWords += phfScratchWords(j.M, j.tcap);
        j.displOff = (uint32_t)displWords;
        displWords += j.M;
        j.tableOff = (uint32_t)tableWords;
        tableWords += j.tcap;
        auto key = std::make_pair(n, seed);
        auto it = abCache.find(key);
        if (it == abCache.end())
        {
            uint32_t off = (uint32_t)(ab.size() / 2);
            auto v = drawAB(n, seed);
            ab.insert(ab.end(), v.begin(), v.end());
            it = abCache.emplace(key, off).first;
        }
        j.abOff = it->second;
        j.nodeBase = nodeBase;
        jobs.push_back(j);
        return (uint32_t)jobs.size() - 1;
    }
};

cgs_status runPhfBatch(PhfBatch& b, const uint32_t* dKeys, uint32_t* dDispl, uint32_t* dTables,
                       std::vector<PhfResult>& results, hipStream_t s, bool wide = false)
{
    DeviceBuffer jobs, ab, scratch, res;
    HIP_TRY(jobs.alloc(b.jobs.size() * sizeof(PhfJob)));
    HIP_TRY(ab.alloc(b.ab.size() * 4));
    HIP_TRY(scratch.alloc(b.scratchWords * 4));
    HIP_TRY(res.alloc(b.jobs.size() * sizeof(PhfResult)));
    HIP_TRY(hipMemcpyAsync(jobs.p, b.jobs.data(), b.jobs.size() * sizeof(PhfJob), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ab.p, b.ab.data(), b.ab.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(launchPhfBuild(jobs.as<PhfJob>(), (uint32_t)b.jobs.size(), dKeys, ab.as<uint32_t>(), scratch.as<uint32_t>(),
                           dDispl, dTables, res.as<PhfResult>(), s, wide));
    results.resize(b.jobs.size());
    HIP_TRY(hipMemcpyAsync(results.data(), res.p, b.jobs.size() * sizeof(PhfResult), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return CGS_OK;
}

uint64_t fastmodM(uint32_t T) { return 0xFFFFFFFFFFFFFFFFull / T + 1; }

} // namespace

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------
static std::atomic<uint64_t> g_sceneSerial {0};

// A resident scene.  Its device arrays are read-only after cgs_scene_create,
// so any context on the same device may trace or render it.
struct cgs_scene
{
    int device = 0;
    uint64_t serial = 0;           // unique per scene of this process
    uint32_t nsurf = 0, nobj = 0, nlight = 0, nodeCount = 0;
    std::vector<BuiltBVH> bvh;     // per surface, then the top level (local offsets)
    std::vector<uint32_t> nodeBase, indexBase;
    std::vector<DevBVH> dev;       // per surface, then the top level
    std::vector<cgs_phf> phf;
    DeviceBuffer patches, nodes, indices, keys, displ, tables, surfaces, objects, lights, topByKey;
    DeviceBuffer cullNodes, cullLinks, cullElems;
    bool cull = false;             // the top-level cull tree is built
    float background[3], ambient[3];
    SceneDev view() const
    {
        SceneDev s {};
        s.patches = patches.as<float4>();
        s.nodes = nodes.as<float4>();
        s.indices = indices.as<uint32_t>();
        s.displ = displ.as<uint32_t>();
        s.tables = tables.as<uint32_t>();
        s.surfaces = surfaces.as<DevBVH>();
        s.objects = objects.as<DevObject>();
        s.lights = lights.as<DevLight>();
        s.topByKey = topByKey.as<uint32_t>();
        s.cullNodes = cullNodes.as<float4>();
        s.cullLinks = cullLinks.as<uint2>();
        s.cullElems = cullElems.as<float4>();
        s.nlight = nlight;
        s.nobj = nobj;
        s.nodeCount = nodeCount;
        s.big = 0;
        for (const DevBVH& b : dev) s.big |= b.wide | (b.cullRoot != CGS_EMPTY && &b != &dev.back() ? 1u : 0u);
        s.top = dev.back();
        for (int i = 0; i < 3; i++)
        {
            s.background[i] = background[i];
            s.ambient[i] = ambient[i];
        }
        return s;
    }
};


extern "C" {

int cgs_abi_version(void) { return CGS_ABI_VERSION; }

const char* cgs_status_string(int s)
{
    switch (s)
    {
    case CGS_OK: return "ok";
    case CGS_ERR_INVALID: return "invalid argument";
    case CGS_ERR_HIP: return "HIP runtime error";
    case CGS_ERR_HASH_FAILED: return "hash table build failed";
    case CGS_ERR_HASH_DUPLICATE: return "perfect hash: duplicate found";
    case CGS_ERR_BUCKET_OVERFLOW: return "perfect hash: bucket overflow";
    case CGS_ERR_BUSY: return "ray tracer not ready: previous render still running";
    case CGS_ERR_EMPTY_SPLIT: return "invalid";
    case CGS_ERR_STACK_OVERFLOW: return "clipping region stack overflow";
    case CGS_ERR_NO_DEVICE: return "no HIP device";
    case CGS_ERR_OUT_OF_MEMORY: return "out of memory";
    default: return "unknown status";
    }
}

const char* cgs_last_error(void) { return g_lastError.c_str(); }

cgs_status cgs_open(int device, cgs_ctx** out)
{
    if (!out) return fail(CGS_ERR_INVALID, "cgs_open: null out");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CGS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(CGS_ERR_INVALID, "cgs_open: device out of range");
    HIP_TRY(hipSetDevice(device));
    auto* c = new cgs_ctx;
    c->device = device;
    hipError_t e = hipEventCreateWithFlags(&c->renderDone, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->batchDone, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->persistDone, hipEventDisableTiming);
    // the no-allocation hash build's scratch array: its address on this device
    if (e == hipSuccess) e = phfStaticInit(device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->numCUs, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipMalloc(&c->dOverflow, 256);
    if (e == hipSuccess) e = hipMemset(c->dOverflow, 0, 256);
    if (e == hipSuccess) e = hipMalloc(&c->dCounters, 256);
    if (e == hipSuccess) e = hipMemset(c->dCounters, 0, 256);
    if (e == hipSuccess) e = hipMalloc(&c->dRenderArgs, kRenderArgsBytes);
    if (e == hipSuccess) e = hipMalloc(&c->dShadeJob, sizeof(ShadeJob));
    if (e != hipSuccess)
    {
        delete c;
        return fail(CGS_ERR_HIP, std::string("cgs_open: ") + hipGetErrorString(e));
    }
    *out = c;
    return CGS_OK;
}

cgs_status cgs_close(cgs_ctx* ctx)
{
    if (!ctx) return CGS_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->renderDone) (void)hipEventSynchronize(ctx->renderDone), (void)hipEventDestroy(ctx->renderDone);
    if (ctx->batchDone) (void)hipEventSynchronize(ctx->batchDone), (void)hipEventDestroy(ctx->batchDone);
    if (ctx->persistDone) (void)hipEventSynchronize(ctx->persistDone), (void)hipEventDestroy(ctx->persistDone);
    (void)hipDeviceSynchronize();
    if (ctx->dOverflow) (void)hipFree(ctx->dOverflow);
    if (ctx->dCounters) (void)hipFree(ctx->dCounters);
    if (ctx->dRenderArgs) (void)hipFree(ctx->dRenderArgs);
    if (ctx->dShadeJob) (void)hipFree(ctx->dShadeJob);
    if (ctx->dStack) (void)hipFree(ctx->dStack);
    if (ctx->hCount) (void)hipHostFree(ctx->hCount);
    delete ctx;
    return CGS_OK;
}

cgs_status cgs_bezier_clip(cgs_ctx* ctx, const cgs_ray* rays, const float* patches, uint32_t n, float tol, cgs_hit* hits,
                           void* stream)
{
    if (!ctx || (n && (!rays || !patches || !hits))) return fail(CGS_ERR_INVALID, "cgs_bezier_clip: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    if (n == 0) return CGS_OK;
    PersistArgs pa;
    cgs_status st = persistArgs(ctx, PersistKernel::ClipBatch, (hipStream_t)stream, n, pa);
    if (st != CGS_OK) return st;
    HIP_TRY(launchClipBatch(pa, rays, patches, n, tol, hits, (hipStream_t)stream));
    return persistEnd(ctx, (hipStream_t)stream);
}

cgs_status cgs_subdivision(cgs_ctx* ctx, const cgs_ray* rays, const float* patches, uint32_t n, float threshold,
                           cgs_hit* hits, void* stream)
{
    if (!ctx || (n && (!rays || !patches || !hits))) return fail(CGS_ERR_INVALID, "cgs_subdivision: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchSubdivBatch(rays, patches, n, threshold, hits, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_patch_eval(cgs_ctx* ctx, uint32_t kind, const float* patches, const float* uv, uint32_t n, float* out,
                          void* stream)
{
    if (!ctx || (n && (!patches || !uv || !out))) return fail(CGS_ERR_INVALID, "cgs_patch_eval: null argument");
    if (kind > CGS_EVAL_NORMAL_MAT) return fail(CGS_ERR_INVALID, "cgs_patch_eval: unknown kind");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchPatchEval(kind, patches, uv, n, out, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_patch_subpatch(cgs_ctx* ctx, const float* patches, const float* ranges, uint32_t n, float* out,
                              void* stream)
{
    if (!ctx || (n && (!patches || !ranges || !out))) return fail(CGS_ERR_INVALID, "cgs_patch_subpatch: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchSubpatch(patches, ranges, n, out, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_bvh_build(const cgs_bvh_element* elements, uint32_t n, uint32_t epn, cgs_bvh_node** nodes,
                         uint32_t** keys, uint32_t* node_count, uint32_t** indices, uint32_t* index_count)
{
    if (!nodes || !keys || !node_count || !indices || !index_count || (n && !elements))
        return fail(CGS_ERR_INVALID, "cgs_bvh_build: null argument");
    std::vector<cgs_bvh_element> el(elements, elements + n);
    BuiltBVH b;
    cgs_status st = buildBVH(el, epn, b);
    if (st != CGS_OK) return st;
    *nodes = (cgs_bvh_node*)std::malloc(b.nodes.size() * sizeof(cgs_bvh_node));
    *keys = (uint32_t*)std::malloc(b.keys.size() * 4);
    *indices = (uint32_t*)std::malloc(b.indices.size() * 4);
    if (!*nodes || !*keys || !*indices) return fail(CGS_ERR_OUT_OF_MEMORY, "cgs_bvh_build: malloc");
    std::memcpy(*nodes, b.nodes.data(), b.nodes.size() * sizeof(cgs_bvh_node));
    std::memcpy(*keys, b.keys.data(), b.keys.size() * 4);
    std::memcpy(*indices, b.indices.data(), b.indices.size() * 4);
    *node_count = (uint32_t)b.nodes.size();
    *index_count = (uint32_t)b.indices.size();
    return CGS_OK;
}

cgs_status cgs_patch_elements(const float* vertices, const uint32_t* indices, uint32_t npatch, cgs_bvh_element* elements)
{
    if (npatch && (!vertices || !indices || !elements)) return fail(CGS_ERR_INVALID, "cgs_patch_elements: null argument");
    patchElements(vertices, indices, npatch, elements);
    return CGS_OK;
}

void cgs_free_host(void* p) { std::free(p); }

uint32_t cgs_phf_displacement_size(uint32_t n) { return displSize(n); }

cgs_status cgs_phf_build(cgs_ctx* ctx, const uint32_t* keys, uint32_t n, uint32_t seed, uint32_t* displacement,
                         cgs_phf* out, void* stream)
{
    if (!ctx || !out || !displacement || (n && !keys)) return fail(CGS_ERR_INVALID, "cgs_phf_build: null argument");
    if (n == 0) return fail(CGS_ERR_INVALID, "cgs_phf_build: empty key set");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    PhfBatch b;
    b.add(0, n, seed, 0);
    DeviceBuffer tables;
    HIP_TRY(tables.alloc(b.tableWords * 4));
    std::vector<PhfResult> r;
    cgs_status st = runPhfBatch(b, keys, displacement, tables.as<uint32_t>(), r, s);
    if (st != CGS_OK) return st;
    out->A = r[0].A;
    out->B = r[0].B;
    out->mask = b.jobs[0].mask;
    out->table_size = r[0].T;
    out->attempts = r[0].attempts;
    out->displacement_size = b.jobs[0].M;
    if (r[0].status != CGS_OK)
        return fail((cgs_status)r[0].status, r[0].status == CGS_ERR_HASH_DUPLICATE ? "PerfectHashFunction: DUPLICATE FOUND!"
                                                                                  : "PerfectHashFunction: bucket overflow");
    return CGS_OK;
}

cgs_status cgs_phf_build_static(cgs_ctx* ctx, const uint32_t* keys, uint32_t n, uint32_t seed, uint32_t* displacement,
                                cgs_phf* out, void* stream)
{
    if (!ctx || !out || !displacement || (n && !keys)) return fail(CGS_ERR_INVALID, "cgs_phf_build_static: null argument");
    if (n == 0) return fail(CGS_ERR_INVALID, "cgs_phf_build_static: empty key set");
    if (n > kStaticPhfMaxKeys)
        return fail(CGS_ERR_INVALID, "cgs_phf_build_static: " + std::to_string(n) + " keys exceed its static scratch (at most " +
                                         std::to_string(kStaticPhfMaxKeys) + " keys); use cgs_phf_build");
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t* base = phfStaticScratch(ctx->device);
    if (!base) return fail(CGS_ERR_HIP, "cgs_phf_build_static: no scratch address for this device");
    hipStream_t s = (hipStream_t)stream;
    PhfJob j {};
    j.n = n;
    j.M = displSize(n);
    j.mask = j.M - 1;
    j.step = growthStep(n);
    j.tcap = n + 16 * j.step;
    if (phfScratchWords(j.M, j.tcap) > kStaticPhfScratchWords || j.tcap > kStaticPhfTcap)
        return fail(CGS_ERR_INVALID, "cgs_phf_build_static: scratch layout too small");
    uint32_t ab[32];
    drawAB(n, seed, ab);
    PhfResult r {};
    HIP_TRY(hipMemcpyAsync(base + kStaticPhfJobWord, &j, sizeof j, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(base + kStaticPhfAbWord, ab, sizeof ab, hipMemcpyHostToDevice, s));
    HIP_TRY(launchPhfBuild((const PhfJob*)(base + kStaticPhfJobWord), 1, keys, base + kStaticPhfAbWord,
                           base + kStaticPhfScratchWord, displacement, base + kStaticPhfTableWord,
                           (PhfResult*)(base + kStaticPhfResultWord), s));
    HIP_TRY(hipMemcpyAsync(&r, base + kStaticPhfResultWord, sizeof r, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out->A = r.A;
    out->B = r.B;
    out->mask = j.mask;
    out->table_size = r.T;
    out->attempts = r.attempts;
    out->displacement_size = j.M;
    if (r.status != CGS_OK)
        return fail((cgs_status)r.status, r.status == CGS_ERR_HASH_DUPLICATE ? "PerfectHashFunction: DUPLICATE FOUND!"
                                                                            : "PerfectHashFunction: bucket overflow");
    return CGS_OK;
}

cgs_status cgs_phf_build_wide(cgs_ctx* ctx, const uint64_t* keys, uint32_t n, uint32_t seed, uint32_t* displacement,
                              cgs_phf* out, void* stream)
{
    if (!ctx || !out || !displacement || (n && !keys)) return fail(CGS_ERR_INVALID, "cgs_phf_build_wide: null argument");
    if (n == 0) return fail(CGS_ERR_INVALID, "cgs_phf_build_wide: empty key set");
    if (n >= (1u << 30)) return fail(CGS_ERR_INVALID, "cgs_phf_build_wide: more than 2^30 keys");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    PhfJob j {};
    j.n = n;
    j.M = displSize(n);
    j.mask = j.M - 1;
    j.step = growthStep(n);
    j.tcap = n + 16 * j.step;
    const std::vector<uint32_t> ab = drawAB(n, seed);
    DeviceBuffer job, dab, scratch, tables, res;
    HIP_TRY(job.alloc(sizeof j));
    HIP_TRY(dab.alloc(ab.size() * 4));
    HIP_TRY(scratch.alloc((size_t)phfScratchWords(j.M, j.tcap, 2) * 4));
    HIP_TRY(tables.alloc((size_t)j.tcap * 4));
    HIP_TRY(res.alloc(sizeof(PhfResult)));
    HIP_TRY(hipMemcpyAsync(job.p, &j, sizeof j, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dab.p, ab.data(), ab.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(launchPhfBuildWide64(job.as<PhfJob>(), 1, keys, dab.as<uint32_t>(), scratch.as<uint32_t>(), displacement,
                                 tables.as<uint32_t>(), res.as<PhfResult>(), s));
    PhfResult r {};
    HIP_TRY(hipMemcpyAsync(&r, res.p, sizeof r, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out->A = r.A;
    out->B = r.B;
    out->mask = j.mask;
    out->table_size = r.T;
    out->attempts = r.attempts;
    out->displacement_size = j.M;
    if (r.status != CGS_OK)
        return fail((cgs_status)r.status, r.status == CGS_ERR_HASH_DUPLICATE ? "PerfectHashFunction: DUPLICATE FOUND!"
                                                                            : "PerfectHashFunction: bucket overflow");
    return CGS_OK;
}

cgs_status cgs_phf_fill_table_wide(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement, const uint64_t* keys,
                                   uint32_t n, uint32_t* table, void* stream)
{
    if (!ctx || !h || !displacement || !table || (n && !keys) || h->table_size == 0)
        return fail(CGS_ERR_INVALID, "cgs_phf_fill_table_wide: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchPhfFillWide64(*h, displacement, keys, n, table, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_phf_hash_wide(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement, const uint64_t* keys,
                             uint32_t n, uint32_t* out, void* stream)
{
    if (!ctx || !h || !displacement || !out || (n && !keys) || h->table_size == 0)
        return fail(CGS_ERR_INVALID, "cgs_phf_hash_wide: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchPhfHashWide64(*h, displacement, keys, n, out, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_phf_fill_table(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement, const uint32_t* keys,
                              uint32_t n, uint32_t* table, void* stream)
{
    if (!ctx || !h || !displacement || !table || (n && !keys) || h->table_size == 0)
        return fail(CGS_ERR_INVALID, "cgs_phf_fill_table: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchPhfFill(*h, displacement, keys, n, 0, table, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_phf_hash(cgs_ctx* ctx, const cgs_phf* h, const uint32_t* displacement, const uint32_t* keys, uint32_t n,
                        uint32_t* out, void* stream)
{
    if (!ctx || !h || !displacement || !out || (n && !keys) || h->table_size == 0)
        return fail(CGS_ERR_INVALID, "cgs_phf_hash: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(launchPhfHash(*h, displacement, keys, n, out, (hipStream_t)stream));
    return CGS_OK;
}

// Whether the queries may enumerate a scene's objects through a cull tree
// (cgs_cull.cpp).  The enumeration skips an object whose top-level box the
// ray enters well beyond the nearest hit found so far, which is sound when a
// hit on an object lies inside that object's box: every control point of
// weight 1 (the box is then the control points' box, and a Bézier patch lies
// in its control hull) and affine object transforms whose world→local matrix
// undoes local→world on the object's box corners to within 1e-5 of the box's
// size plus distance from the origin, in world units.
// CGS_CULL=0 in the environment turns the tree off (a diagnostic: the
// results are the same).
static bool cullEnabled()
{
    const char* e = std::getenv("CGS_CULL");
    return !(e && e[0] == '0');
}

static bool cullAllowed(const cgs_scene_desc* d)
{
    if (!cullEnabled()) return false;
    if (d->nobj < 2) return false;
    for (uint64_t v = 0; v < d->surf_vbase[d->nsurf]; v++)
        if (d->vertices[4 * v + 3] != 1.0f) return false;
    for (uint32_t i = 0; i < d->nobj; i++)
    {
        const uint32_t k = d->obj_surface[i];
        if (k == CGS_EMPTY) continue;
        const float* L = d->obj_local2world + 16 * (size_t)i;
        const float* W = d->obj_world2local + 16 * (size_t)i;
        for (const float* M : {L, W})
        {
            if (M[3] != 0.0f || M[7] != 0.0f || M[11] != 0.0f || M[15] != 1.0f) return false;
            for (int q = 0; q < 16; q++)
                if (!std::isfinite(M[q])) return false;
        }
        // the object's local control-point box, mapped out and back
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const float* V = d->vertices + 4 * (size_t)d->surf_vbase[k];
        for (uint64_t q = 16ull * d->surf_pbase[k]; q < 16ull * d->surf_pbase[k + 1]; q++)
            for (int a = 0; a < 3; a++)
            {
                lo[a] = std::min(lo[a], (double)V[4 * (size_t)d->indices[q] + a]);
                hi[a] = std::max(hi[a], (double)V[4 * (size_t)d->indices[q] + a]);
            }
        // the round trip's error, mapped back to world units, against the
        // box's size and distance from the origin
        double err = 0.0, mag = 0.0, wlo[3] = {INFINITY, INFINITY, INFINITY}, whi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; c++)
        {
            const double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
            double cw[3], e[3];
            for (int r = 0; r < 3; r++)
            {
                cw[r] = (double)L[r] * p[0] + (double)L[4 + r] * p[1] + (double)L[8 + r] * p[2] + (double)L[12 + r];
                wlo[r] = std::min(wlo[r], cw[r]);
                whi[r] = std::max(whi[r], cw[r]);
                mag = std::max(mag, std::fabs(cw[r]));
            }
            for (int r = 0; r < 3; r++)
                e[r] = (double)W[r] * cw[0] + (double)W[4 + r] * cw[1] + (double)W[8 + r] * cw[2] + (double)W[12 + r] - p[r];
            for (int r = 0; r < 3; r++)
                err = std::max(err, std::fabs((double)L[r] * e[0] + (double)L[4 + r] * e[1] + (double)L[8 + r] * e[2]));
        }
        double ext = 0.0;
        for (int r = 0; r < 3; r++) ext = std::max(ext, whi[r] - wlo[r]);
        if (!(err <= 1e-5 * (mag + ext))) return false;
    }
    return true;
}

// cgs_scene_create, with surface 0's BVH built over the given element boxes
// when `surface0` is not null (a PatchBVH over a patch table, whose boxes
// come from basis-transformed points: that surface gets no cull tree, whose
// exactness needs every hit inside its patch's box).
static cgs_status sceneCreate(cgs_ctx* ctx, const cgs_scene_desc* d, cgs_scene** out, const cgs_bvh_element* surface0)
{
    if (!ctx || !d || !out) return fail(CGS_ERR_INVALID, "cgs_scene_create: null argument");
    if (d->nsurf == 0 || d->nobj == 0 || !d->vertices || !d->indices || !d->surf_vbase || !d->surf_pbase ||
        !d->obj_surface || !d->obj_local2world || !d->obj_world2local || !d->obj_material || (d->nlight && !d->lights))
        return fail(CGS_ERR_INVALID, "cgs_scene_create: incomplete description");
    for (uint32_t i = 0; i < d->nobj; i++)
        if (d->obj_surface[i] >= d->nsurf && d->obj_surface[i] != CGS_EMPTY)
            return fail(CGS_ERR_INVALID, "cgs_scene_create: object surface out of range");
    for (uint32_t k = 0; k < d->nsurf; k++)
        if (d->surf_pbase[k + 1] <= d->surf_pbase[k] || d->surf_vbase[k + 1] < d->surf_vbase[k])
            return fail(CGS_ERR_EMPTY_SPLIT, "cgs_scene_create: a surface has no patch");
    HIP_TRY(hipSetDevice(ctx->device));
    auto sc = std::make_unique<cgs_scene>();
    sc->device = ctx->device;
    sc->serial = ++g_sceneSerial;
    sc->nsurf = d->nsurf;
    sc->nobj = d->nobj;
    sc->nlight = d->nlight;
    std::memcpy(sc->background, d->background, 12);
    std::memcpy(sc->ambient, d->ambient, 12);
    const uint32_t S = d->nsurf;
    const uint64_t npatch = d->surf_pbase[S];
    // validate indices
    for (uint32_t k = 0; k < S; k++)
    {
        const uint32_t nv = d->surf_vbase[k + 1] - d->surf_vbase[k];
        for (uint64_t q = 16ull * d->surf_pbase[k]; q < 16ull * d->surf_pbase[k + 1]; q++)
            if (d->indices[q] >= nv) return fail(CGS_ERR_INVALID, "cgs_scene_create: patch index out of range");
    }

    // 1. per-surface BVHs (BezierSurface::buildBVH), in parallel over surfaces
    sc->bvh.resize(S + 1);
    std::vector<cgs_status> st(S, CGS_OK);
    parallelFor(S, [&](size_t k) {
        const uint32_t p0 = d->surf_pbase[k], p1 = d->surf_pbase[k + 1];
        std::vector<cgs_bvh_element> el(p1 - p0);
        if (k == 0 && surface0)
            std::memcpy(el.data(), surface0, el.size() * sizeof(cgs_bvh_element));
        else
            patchElements(d->vertices + 4 * (size_t)d->surf_vbase[k], d->indices + 16 * (size_t)p0, p1 - p0, el.data());
        st[k] = buildBVH(el, 1, sc->bvh[k]);
    });
    for (auto s : st)
        if (s != CGS_OK) return s;

    // 2. top-level BVH over object world bounds (Scene::buildBVH, rt/Scene.cu:39-53).
    //    An object without a surface (CGS_EMPTY: the reference's Mesh and
    //    Sphere primitives, whose bounds() are empty, Primitive.cu:15-23,47-56)
    //    keeps the empty box, centred at 0: it shapes the BVH like the
    //    reference's and no ray enters it.
    {
        std::vector<cgs_bvh_element> el(d->nobj);
        for (uint32_t i = 0; i < d->nobj; i++)
        {
            const uint32_t k = d->obj_surface[i];
            const float* M = d->obj_local2world + 16 * (size_t)i;
            Box3 b;
            const float* V = k == CGS_EMPTY ? nullptr : d->vertices + 4 * (size_t)d->surf_vbase[k];
            const uint64_t q0 = k == CGS_EMPTY ? 0 : 16ull * d->surf_pbase[k];
            const uint64_t q1 = k == CGS_EMPTY ? 0 : 16ull * d->surf_pbase[k + 1];
            for (uint64_t q = q0; q < q1; q++)
            {
                const float* v = V + 4 * (size_t)d->indices[q];
                float p[3];
                for (int r = 0; r < 3; r++)
                    p[r] = ((v[0] * M[r] + v[1] * M[4 + r]) + v[2] * M[8 + r]) + v[3] * M[12 + r];
                b.inflate(p);
            }
            std::memcpy(el[i].bmin, b.mn, 12);
            std::memcpy(el[i].bmax, b.mx, 12);
            for (int r = 0; r < 3; r++) el[i].centroid[r] = 0.5f * (b.mn[r] + b.mx[r]);
            el[i].index = i;
        }
        cgs_status s = buildBVH(el, 1, sc->bvh[S]);
        if (s != CGS_OK) return s;
    }

    // 2b. cull trees: over the top-level leaf boxes where the scene allows
    // one, and over the leaf boxes of each surface of at least
    // kSurfaceCullMinPatches patches whose control points all have weight 1
    // (a hit then lies in its patch's box); all in one set of arrays
    CullTreeHost cullTree, one;
    std::vector<uint32_t> cullRoot(S + 1, CGS_EMPTY);
    if (cullAllowed(d) && buildCullTree(sc->bvh[S].nodes.data(), sc->bvh[S].nodes.size(), sc->bvh[S].indices.data(), one))
        cullRoot[S] = appendCullTree(cullTree, one);
    sc->cull = cullRoot[S] != CGS_EMPTY;
    if (cullEnabled())
        for (uint32_t k = 0; k < S; k++)
        {
            if (d->surf_pbase[k + 1] - d->surf_pbase[k] < kSurfaceCullMinPatches || (k == 0 && surface0)) continue;
            bool unit = true;
            for (uint64_t v = d->surf_vbase[k]; v < d->surf_vbase[k + 1] && unit; v++) unit = d->vertices[4 * v + 3] == 1.0f;
            if (unit && buildCullTree(sc->bvh[k].nodes.data(), sc->bvh[k].nodes.size(), sc->bvh[k].indices.data(), one))
                cullRoot[k] = appendCullTree(cullTree, one);
        }

    // 3. layout: absolute node offsets, leaf slots and patch bases.  The
    // top-level BVH (index S) comes first in the node array: the kernels
    // address its nodes with 32-bit byte offsets from the array's start.
    if (sc->bvh[S].nodes.size() > (1u << 26))
        return fail(CGS_ERR_INVALID, "cgs_scene_create: top-level BVH beyond 2^26 nodes");
    uint64_t nodeTotal = 0, indexTotal = 0, keyTotal = 0;
    sc->nodeBase.resize(S + 1);
    sc->indexBase.resize(S + 1);
    for (uint32_t j = 0; j <= S; j++)
    {
        const uint32_t k = j == 0 ? S : j - 1;
        sc->nodeBase[k] = (uint32_t)nodeTotal;
        sc->indexBase[k] = (uint32_t)indexTotal;
        nodeTotal += sc->bvh[k].nodes.size();
        indexTotal += sc->bvh[k].indices.size();
    }
    keyTotal = nodeTotal;
    sc->nodeCount = (uint32_t)nodeTotal;
    if (nodeTotal >= 0xFFFFFFF0ull || indexTotal >= 0xFFFFFFF0ull)
        return fail(CGS_ERR_INVALID, "cgs_scene_create: scene too large for 32-bit offsets");
    std::vector<cgs_bvh_node> nodes(nodeTotal);
    std::vector<uint32_t> indices(indexTotal), keys(keyTotal);
    parallelFor(S + 1, [&](size_t k) {
        const BuiltBVH& b = sc->bvh[k];
        const uint32_t nb = sc->nodeBase[k], ib = sc->indexBase[k];
        for (size_t i = 0; i < b.nodes.size(); i++)
        {
            cgs_bvh_node n = b.nodes[i];
            if (n.left != CGS_EMPTY)
            {
                n.left += nb;
                n.right += nb;
            }
            else
                n.right += ib;
            nodes[nb + i] = n;
            keys[nb + i] = b.keys[i];
        }
        std::copy(b.indices.begin(), b.indices.end(), indices.begin() + ib);
    });

    // 4. gathered patches (256 B each), objects, lights
    std::vector<float> patches(npatch * 64);
    parallelFor(S, [&](size_t k) {
        const float* V = d->vertices + 4 * (size_t)d->surf_vbase[k];
        for (uint64_t p = d->surf_pbase[k]; p < d->surf_pbase[k + 1]; p++)
            for (int j = 0; j < 16; j++)
                std::memcpy(&patches[64 * p + 4 * j], V + 4 * (size_t)d->indices[16 * p + j], 16);
    });
    std::vector<DevObject> objs(d->nobj);
    for (uint32_t i = 0; i < d->nobj; i++)
    {
        DevObject& o = objs[i];
        std::memcpy(o.w2l, d->obj_world2local + 16 * (size_t)i, 64);
        const float* m = d->obj_material + 8 * (size_t)i;
        o.mat0 = make_float4(m[0], m[1], m[2], m[6]);
        o.mat1 = make_float4(m[3], m[4], m[5], m[7]);
        o.surface = d->obj_surface[i];
        o.pad[0] = o.pad[1] = o.pad[2] = 0;
    }
    // upper levels of the top-level BVH by heap key (LDS staging in the kernels)
    std::vector<uint32_t> topByKey(kTopKeyCap, CGS_EMPTY);
    {
        const BuiltBVH& tb = sc->bvh[S];
        for (size_t i = 0; i < tb.keys.size(); i++)
            if (tb.keys[i] < kTopKeyCap) topByKey[tb.keys[i]] = sc->nodeBase[S] + (uint32_t)i;
    }
    hipStream_t s = nullptr;
    HIP_TRY(sc->topByKey.alloc(kTopKeyCap * 4));
    HIP_TRY(hipMemcpyAsync(sc->topByKey.p, topByKey.data(), kTopKeyCap * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(sc->patches.alloc(patches.size() * 4));
    HIP_TRY(sc->nodes.alloc(nodes.size() * sizeof(cgs_bvh_node)));
    HIP_TRY(sc->indices.alloc(indices.size() * 4));
    HIP_TRY(sc->keys.alloc(keys.size() * 4));
    HIP_TRY(sc->objects.alloc(objs.size() * sizeof(DevObject)));
    HIP_TRY(sc->lights.alloc(std::max<size_t>(1, d->nlight) * sizeof(DevLight)));
    HIP_TRY(hipMemcpyAsync(sc->patches.p, patches.data(), patches.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc->nodes.p, nodes.data(), nodes.size() * sizeof(cgs_bvh_node), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc->indices.p, indices.data(), indices.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc->keys.p, keys.data(), keys.size() * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc->objects.p, objs.data(), objs.size() * sizeof(DevObject), hipMemcpyHostToDevice, s));
    if (d->nlight)
        HIP_TRY(hipMemcpyAsync(sc->lights.p, d->lights, d->nlight * sizeof(DevLight), hipMemcpyHostToDevice, s));
    sc->nlight = d->nlight;
    if (cullTree.nodeCount)
    {
        HIP_TRY(sc->cullNodes.alloc(cullTree.nodes.size() * 4));
        HIP_TRY(sc->cullLinks.alloc(cullTree.links.size() * 4));
        HIP_TRY(sc->cullElems.alloc(cullTree.elems.size() * 4));
        HIP_TRY(hipMemcpyAsync(sc->cullNodes.p, cullTree.nodes.data(), cullTree.nodes.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(sc->cullLinks.p, cullTree.links.data(), cullTree.links.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(sc->cullElems.p, cullTree.elems.data(), cullTree.elems.size() * 4, hipMemcpyHostToDevice, s));
    }

    // 5. perfect hashes for every BVH in one launch (BVH::buildHashTable)
    PhfBatch batch;
    for (uint32_t k = 0; k <= S; k++)
        batch.add(sc->nodeBase[k], (uint32_t)sc->bvh[k].keys.size(), d->hash_seed, sc->nodeBase[k]);
    HIP_TRY(sc->displ.alloc(batch.displWords * 4));
    HIP_TRY(sc->tables.alloc(batch.tableWords * 4));
    std::vector<PhfResult> res;
    cgs_status ps = runPhfBatch(batch, sc->keys.as<uint32_t>(), sc->displ.as<uint32_t>(), sc->tables.as<uint32_t>(), res, s);
    if (ps != CGS_OK) return ps;
    // A BVH whose reference-family hash has a duplicate (every attempt fails
    // from 2^21 heap keys on) is hashed again with the wide family, into the
    // same displacement and table ranges; a bucket overflow fails both.
    std::vector<uint8_t> wide(S + 1, 0);
    {
        PhfBatch again;
        std::vector<uint32_t> which;
        for (uint32_t k = 0; k <= S; k++)
            if (res[k].status == CGS_ERR_HASH_DUPLICATE)
            {
                PhfJob j = batch.jobs[k];
                j.scratchOff = (uint32_t)again.scratchWords;
                again.scratchWords += phfScratchWords(j.M, j.tcap);
                j.abOff = (uint32_t)(again.ab.size() / 2);
                again.ab.insert(again.ab.end(), batch.ab.begin() + 2 * batch.jobs[k].abOff,
                                batch.ab.begin() + 2 * (batch.jobs[k].abOff + 16));
                again.jobs.push_back(j);
                which.push_back(k);
            }
        if (!which.empty())
        {
            std::vector<PhfResult> res2;
            ps = runPhfBatch(again, sc->keys.as<uint32_t>(), sc->displ.as<uint32_t>(), sc->tables.as<uint32_t>(), res2, s,
                             true);
            if (ps != CGS_OK) return ps;
            for (size_t q = 0; q < which.size(); q++)
            {
                res[which[q]] = res2[q];
                wide[which[q]] = 1;
            }
        }
    }
    sc->dev.resize(S + 1);
    sc->phf.resize(S + 1);
    for (uint32_t k = 0; k <= S; k++)
    {
        const PhfJob& j = batch.jobs[k];
        const PhfResult& r = res[k];
        if (r.status != CGS_OK)
            return fail(CGS_ERR_HASH_FAILED, "hash table build failed (BVH " + std::to_string(k) + ", " +
                                                 cgs_status_string((int)r.status) + ")");
        DevBVH& b = sc->dev[k];
        b = DevBVH {};
        b.root = sc->nodeBase[k];
        b.displBase = j.displOff;
        b.tableBase = j.tableOff;
        b.A = r.A;
        b.B = r.B;
        b.mask = j.mask;
        b.T = r.T;
        b.patchBase = k < S ? d->surf_pbase[k] : 0;
        b.Tm = fastmodM(r.T);
        b.wide = wide[k];
        b.cullRoot = cullRoot[k];
        b.finiteBoxes = 1;
        for (const cgs_bvh_node& n : sc->bvh[k].nodes)
            if (n.left != CGS_EMPTY)
                for (int q = 0; q < 3; q++)
                    if (!std::isfinite(n.lmin[q]) || !std::isfinite(n.lmax[q]) ||
                        !std::isfinite(n.rmin[q]) || !std::isfinite(n.rmax[q]))
                        b.finiteBoxes = 0;
        sc->phf[k] = cgs_phf {r.A, r.B, j.mask, r.T, r.attempts, j.M};
    }
    HIP_TRY(sc->surfaces.alloc(S * sizeof(DevBVH)));
    HIP_TRY(hipMemcpyAsync(sc->surfaces.p, sc->dev.data(), S * sizeof(DevBVH), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    *out = sc.release();
    return CGS_OK;
}

cgs_status cgs_scene_create(cgs_ctx* ctx, const cgs_scene_desc* d, cgs_scene** out)
{
    return sceneCreate(ctx, d, out, nullptr);
}

cgs_status cgs_scene_destroy(cgs_scene* scene)
{
    if (!scene) return CGS_OK;
    (void)hipSetDevice(scene->device);
    (void)hipDeviceSynchronize();
    delete scene;
    return CGS_OK;
}

cgs_status cgs_scene_bvh_info(cgs_scene* sc, int32_t surface, uint32_t* node_count, uint32_t* index_count, cgs_phf* hash)
{
    if (!sc || surface < -1 || surface >= (int32_t)sc->nsurf) return fail(CGS_ERR_INVALID, "cgs_scene_bvh_info: bad surface");
    const uint32_t k = surface < 0 ? sc->nsurf : (uint32_t)surface;
    if (node_count) *node_count = (uint32_t)sc->bvh[k].nodes.size();
    if (index_count) *index_count = (uint32_t)sc->bvh[k].indices.size();
    if (hash) *hash = sc->phf[k];
    return CGS_OK;
}

int cgs_scene_culled(cgs_scene* sc) { return sc ? (sc->cull ? 1 : 0) : -1; }

int cgs_scene_bvh_culled(cgs_scene* sc, int32_t surface)
{
    if (!sc || surface < -1 || surface >= (int32_t)sc->nsurf) return -1;
    return sc->dev[surface < 0 ? sc->nsurf : (uint32_t)surface].cullRoot != CGS_EMPTY ? 1 : 0;
}

int cgs_scene_bvh_wide(cgs_scene* sc, int32_t surface)
{
    if (!sc || surface < -1 || surface >= (int32_t)sc->nsurf) return -1;
    return sc->dev[surface < 0 ? sc->nsurf : (uint32_t)surface].wide ? 1 : 0;
}

} // extern "C"

// A PatchBVH is a scene of its one surface, as one object in its own frame:
// the surface's BVH, hash and gathered patches are what the pick kernel reads.
struct cgs_patchbvh
{
    cgs_scene* scene = nullptr;
    int device = 0;
};

extern "C" {

static cgs_status patchbvhCreate(cgs_ctx* ctx, const float* vertices, uint32_t nv, const uint32_t* indices, uint32_t np,
                                 const cgs_bvh_element* elements, cgs_patchbvh** out)
{
    if (!ctx || !out || !vertices || !indices || nv == 0 || np == 0)
        return fail(CGS_ERR_INVALID, "cgs_patchbvh_create: empty or null surface");
    static const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const uint32_t vbase[2] = {0, nv}, pbase[2] = {0, np}, surf = 0;
    const float mat[8] = {};
    cgs_scene_desc d {};
    d.vertices = vertices;
    d.indices = indices;
    d.surf_vbase = vbase;
    d.surf_pbase = pbase;
    d.nsurf = 1;
    d.obj_surface = &surf;
    d.obj_local2world = I;
    d.obj_world2local = I;
    d.obj_material = mat;
    d.nobj = 1;
    d.hash_seed = 42;
    auto pb = std::make_unique<cgs_patchbvh>();
    cgs_status st = sceneCreate(ctx, &d, &pb->scene, elements);
    if (st != CGS_OK) return st;
    pb->device = ctx->device;
    *out = pb.release();
    return CGS_OK;
}

cgs_status cgs_patchbvh_create(cgs_ctx* ctx, const float* vertices, uint32_t nv, const uint32_t* indices, uint32_t np,
                               cgs_patchbvh** out)
{
    return patchbvhCreate(ctx, vertices, nv, indices, np, nullptr, out);
}

cgs_status cgs_patchbvh_create_elements(cgs_ctx* ctx, const float* vertices, uint32_t nv, const uint32_t* indices,
                                        uint32_t np, const cgs_bvh_element* elements, cgs_patchbvh** out)
{
    if (!elements) return fail(CGS_ERR_INVALID, "cgs_patchbvh_create_elements: null elements");
    // the leaf lists hold the element indices, which address the patches
    for (uint32_t i = 0; i < np; i++)
        if (elements[i].index >= np) return fail(CGS_ERR_INVALID, "cgs_patchbvh_create_elements: element index out of range");
    return patchbvhCreate(ctx, vertices, nv, indices, np, elements, out);
}

cgs_status cgs_patchbvh_destroy(cgs_patchbvh* pb)
{
    if (!pb) return CGS_OK;
    cgs_status st = cgs_scene_destroy(pb->scene);
    delete pb;
    return st;
}

static cgs_status patchbvhQuery(cgs_ctx* ctx, cgs_patchbvh* pb, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                                uint32_t* occluded, void* stream, const char* what)
{
    if (!ctx || !pb || (n && (!rays || (!hits && !occluded)))) return fail(CGS_ERR_INVALID, std::string(what) + ": null argument");
    if (pb->device != ctx->device) return fail(CGS_ERR_INVALID, std::string(what) + ": PatchBVH on another device");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    if (n == 0) return CGS_OK;
    PersistArgs pa;
    cgs_status st = persistArgs(ctx, PersistKernel::Pick, (hipStream_t)stream, n, pa);
    if (st != CGS_OK) return st;
    HIP_TRY(launchPick(pb->scene->view(), pa, rays, n, hits, occluded, (hipStream_t)stream));
    return persistEnd(ctx, (hipStream_t)stream);
}

cgs_status cgs_patchbvh_intersect(cgs_ctx* ctx, cgs_patchbvh* pb, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                                  void* stream)
{
    if (!hits && n) return fail(CGS_ERR_INVALID, "cgs_patchbvh_intersect: null hits");
    return patchbvhQuery(ctx, pb, rays, n, hits, nullptr, stream, "cgs_patchbvh_intersect");
}

cgs_status cgs_patchbvh_intersect_any(cgs_ctx* ctx, cgs_patchbvh* pb, const cgs_ray* rays, uint32_t n,
                                      uint32_t* occluded, void* stream)
{
    if (!occluded && n) return fail(CGS_ERR_INVALID, "cgs_patchbvh_intersect_any: null output");
    return patchbvhQuery(ctx, pb, rays, n, nullptr, occluded, stream, "cgs_patchbvh_intersect_any");
}

cgs_status cgs_scene_bvh_copy(cgs_scene* sc, int32_t surface, cgs_bvh_node* nodes, uint32_t* keys, uint32_t* indices,
                              uint32_t* displacement, uint32_t* table)
{
    if (!sc || surface < -1 || surface >= (int32_t)sc->nsurf) return fail(CGS_ERR_INVALID, "cgs_scene_bvh_copy: bad surface");
    const uint32_t k = surface < 0 ? sc->nsurf : (uint32_t)surface;
    const BuiltBVH& b = sc->bvh[k];
    if (nodes) std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(cgs_bvh_node));
    if (keys) std::memcpy(keys, b.keys.data(), b.keys.size() * 4);
    if (indices) std::memcpy(indices, b.indices.data(), b.indices.size() * 4);
    HIP_TRY(hipSetDevice(sc->device));
    if (displacement)
        HIP_TRY(hipMemcpy(displacement, sc->displ.as<uint32_t>() + sc->dev[k].displBase, sc->phf[k].displacement_size * 4,
                          hipMemcpyDeviceToHost));
    if (table)
    {
        HIP_TRY(hipMemcpy(table, sc->tables.as<uint32_t>() + sc->dev[k].tableBase, sc->phf[k].table_size * 4,
                          hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < sc->phf[k].table_size; i++)
            if (table[i] != CGS_EMPTY) table[i] -= sc->nodeBase[k];
    }
    return CGS_OK;
}

// The trace mode of a query's options (null: the default).
static cgs_status traceMode(const cgs_trace_options* o, TraceMode& m, const char* what)
{
    m = TraceMode {};
    if (!o) return CGS_OK;
    if (o->method != CGS_METHOD_CLIPPING && o->method != CGS_METHOD_SUBDIVISION)
        return fail(CGS_ERR_INVALID, std::string(what) + ": unknown method");
    if (o->subdiv_depth > 16) return fail(CGS_ERR_INVALID, std::string(what) + ": subdivision depth above 16");
    if (!(o->clip_tol >= 0.0f) || !std::isfinite(o->clip_tol))
        return fail(CGS_ERR_INVALID, std::string(what) + ": bad clipping tolerance");
    m.method = o->method;
    m.polish = o->polish ? 1u : 0u;
    if (o->clip_tol > 0.0f) m.tol = o->clip_tol;
    if (o->subdiv_depth) m.threshold = std::ldexp(1.0f, 1 - (int)o->subdiv_depth);
    m.maxIters = o->max_clip_iters;
    return CGS_OK;
}

cgs_status cgs_trace_closest_ex(cgs_ctx* ctx, cgs_scene* sc, const cgs_ray* rays, uint32_t n, cgs_hit* hits,
                                int32_t* objects, const cgs_trace_options* options, void* stream)
{
    if (!ctx || !sc || (n && (!rays || !hits))) return fail(CGS_ERR_INVALID, "cgs_trace_closest: null argument");
    if (sc->device != ctx->device) return fail(CGS_ERR_INVALID, "cgs_trace_closest: scene on another device");
    TraceMode m;
    cgs_status st = traceMode(options, m, "cgs_trace_closest");
    if (st != CGS_OK) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    if (n == 0) return CGS_OK;
    PersistArgs pa;
    if ((st = persistArgs(ctx, PersistKernel::TraceClosest, (hipStream_t)stream, n, pa)) != CGS_OK) return st;
    HIP_TRY(launchTraceClosest(sc->view(), pa, rays, n, hits, objects, ctx->dOverflow, (hipStream_t)stream, m));
    return persistEnd(ctx, (hipStream_t)stream);
}

cgs_status cgs_trace_any_ex(cgs_ctx* ctx, cgs_scene* sc, const cgs_ray* rays, uint32_t n, uint32_t* occluded,
                            const cgs_trace_options* options, void* stream)
{
    if (!ctx || !sc || (n && (!rays || !occluded))) return fail(CGS_ERR_INVALID, "cgs_trace_any: null argument");
    if (sc->device != ctx->device) return fail(CGS_ERR_INVALID, "cgs_trace_any: scene on another device");
    TraceMode m;
    cgs_status st = traceMode(options, m, "cgs_trace_any");
    if (st != CGS_OK) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    if (n == 0) return CGS_OK;
    PersistArgs pa;
    if ((st = persistArgs(ctx, PersistKernel::TraceAny, (hipStream_t)stream, n, pa)) != CGS_OK) return st;
    HIP_TRY(launchTraceAny(sc->view(), pa, rays, n, occluded, ctx->dOverflow, (hipStream_t)stream, m));
    return persistEnd(ctx, (hipStream_t)stream);
}

cgs_status cgs_trace_closest(cgs_ctx* ctx, cgs_scene* sc, const cgs_ray* rays, uint32_t n, cgs_hit* hits, int32_t* objects,
                             void* stream)
{
    return cgs_trace_closest_ex(ctx, sc, rays, n, hits, objects, nullptr, stream);
}

cgs_status cgs_trace_any(cgs_ctx* ctx, cgs_scene* sc, const cgs_ray* rays, uint32_t n, uint32_t* occluded, void* stream)
{
    return cgs_trace_any_ex(ctx, sc, rays, n, occluded, nullptr, stream);
}

namespace
{
// cgs_render and cgs_render_sampled: samples = 0 is the plain render.
cgs_status renderFrame(const char* name, cgs_ctx* ctx, cgs_scene* sc, const cgs_camera* cam, const cgs_render_options* op,
                       uint32_t samples, uint32_t seed, uint32_t w, uint32_t h, uint32_t* frame, float* color,
                       cgs_ray_counts* counts, void* stream)
{
    const std::string fn(name);
    if (!ctx || !sc || !cam || !op || !frame) return fail(CGS_ERR_INVALID, fn + ": null argument");
    if (sc->device != ctx->device) return fail(CGS_ERR_INVALID, fn + ": scene on another device");
    if ((uint64_t)w * h >= (1ull << 32)) return fail(CGS_ERR_INVALID, fn + ": frame too large");
    const uint32_t stride = op->frame_stride ? op->frame_stride : w;
    if (stride < w) return fail(CGS_ERR_INVALID, fn + ": frame stride smaller than the width");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    if (ctx->renderLaunched)
    {
        hipError_t q = hipEventQuery(ctx->renderDone);
        if (q == hipErrorNotReady)
            return fail(CGS_ERR_BUSY, "CUDA ray tracer not ready: wait the previous work to finish or create a new instance");
        HIP_TRY(q);
    }
    // camera setup (RayTracer.cu:141-163), on the host
    const float halfHeight = std::tan(0.5f * cam->fovy);
    float hdy = halfHeight / (float)h;
    const float hdx = hdy;
    const float tlx = -hdx * (float)(w - 1);
    float tly = halfHeight - hdy;
    if (cam->flip_y)
    {
        tly = -tly;
        hdy = -hdy;
    }
    hipStream_t s = (hipStream_t)stream;
    if (w == 0 || h == 0) return CGS_OK;
    PersistArgs pa;
    cgs_status pst = persistArgs(ctx, PersistKernel::Render, s, renderItems(w, h), pa);
    if (pst != CGS_OK) return pst;
    const uint32_t nTiles = ((w + 7) / 8) * ((h + 7) / 8);
    TileState& ts = tileState(ctx, sc->serial);
    if (nTiles > ts.cap)
    {
        HIP_TRY(ts.ticks.alloc((size_t)nTiles * 4));
        HIP_TRY(ts.order.alloc((size_t)nTiles * 4));
        ts.cap = nTiles;
        ts.ready = false;
    }
    const bool ordered = ts.ready && ts.w == w && ts.h == h;
    HIP_TRY(launchRender(sc->view(), pa, cam->to_world, cam->position, tlx, tly, hdx, hdy, op->eps, op->recursion_depth,
                         w, h, frame, color, (unsigned long long*)counts, stride,
                         ordered ? ts.order.as<uint32_t>() : nullptr, ts.ticks.as<uint32_t>(), ctx->dRenderArgs, s,
                         samples, seed));
    HIP_TRY(launchTileOrder(ts.ticks.as<uint32_t>(), nTiles, ts.order.as<uint32_t>(), s));
    ts.ready = true;
    ts.w = w;
    ts.h = h;
    HIP_TRY(hipEventRecord(ctx->renderDone, s));
    ctx->renderLaunched = true;
    return persistEnd(ctx, s);
}
} // namespace

cgs_status cgs_render(cgs_ctx* ctx, cgs_scene* sc, const cgs_camera* cam, const cgs_render_options* op, uint32_t w,
                      uint32_t h, uint32_t* frame, float* color, cgs_ray_counts* counts, void* stream)
{
    return renderFrame("cgs_render", ctx, sc, cam, op, 0, 0, w, h, frame, color, counts, stream);
}

cgs_status cgs_render_sampled(cgs_ctx* ctx, cgs_scene* sc, const cgs_camera* cam, const cgs_render_options* op,
                              uint32_t samples, uint32_t seed, uint32_t w, uint32_t h, uint32_t* frame, float* color,
                              cgs_ray_counts* counts, void* stream)
{
    if (samples == 0 || samples > CGS_MAX_SAMPLES)
        return fail(CGS_ERR_INVALID, "cgs_render_sampled: samples must lie in 1.." + std::to_string(CGS_MAX_SAMPLES));
    return renderFrame("cgs_render_sampled", ctx, sc, cam, op, samples, seed, w, h, frame, color, counts, stream);
}

cgs_status cgs_shade(cgs_ctx* ctx, cgs_scene* sc, const cgs_ray* rays, uint32_t n, const cgs_render_options* op,
                     float* color, uint32_t* pixels, cgs_hit* hits, int32_t* objects, cgs_ray_counts* counts,
                     void* stream)
{
    if (!ctx || !sc || !op || (n && (!rays || !color))) return fail(CGS_ERR_INVALID, "cgs_shade: null argument");
    if (sc->device != ctx->device) return fail(CGS_ERR_INVALID, "cgs_shade: scene on another device");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    if (n == 0) return CGS_OK;
    hipStream_t s = (hipStream_t)stream;
    PersistArgs pa;
    cgs_status st = persistArgs(ctx, PersistKernel::Shade, s, n, pa);
    if (st != CGS_OK) return st;
    const SceneDev view = sc->view();
    ShadeJob job {};
    job.sc = view;
    job.rays = rays;
    job.color = color;
    job.pixels = pixels;
    job.hits = hits;
    job.objects = objects;
    job.n = n;
    job.rayBase = 0;
    job.eps = op->eps;
    job.depth = op->recursion_depth;
    // `job` is pageable (the stack): HIP performs such a copy synchronously,
    // so it has been read when the call returns; the copy itself runs in
    // stream order after the previous persistent launch.
    HIP_TRY(hipMemcpyAsync(ctx->dShadeJob, &job, sizeof job, hipMemcpyHostToDevice, s));
    HIP_TRY(launchShade(&view, ctx->dShadeJob, 1, n, 64, pa.stack, pa.lanes, pa.counter, pa.grid,
                        (unsigned long long*)counts, s));
    return persistEnd(ctx, s);
}

cgs_status cgs_work_counts(cgs_ctx* ctx, uint64_t* out, uint32_t n, int reset)
{
    if (!ctx || (n && !out)) return fail(CGS_ERR_INVALID, "cgs_work_counts: null argument");
    if (!CGS_WORK_COUNTS) return fail(CGS_ERR_INVALID, "cgs_work_counts: this build does not count work events");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    unsigned long long all[kWorkEvents];
    HIP_TRY(workCounts(all, reset != 0));
    for (uint32_t k = 0; k < n && k < (uint32_t)kWorkEvents; k++) out[k] = all[k];
    return CGS_OK;
}

int cgs_render_running(cgs_ctx* ctx)
{
    if (!ctx || !ctx->renderLaunched) return 0;
    return hipEventQuery(ctx->renderDone) == hipErrorNotReady ? 1 : 0;
}

} // extern "C"

// ---------------------------------------------------------------------------
// batched queries
// ---------------------------------------------------------------------------
// The job table of a batched launch of `grid` workgroups taking `chunk` rays
// per wave fetch.
static cgs_status buildBatchJobs(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, uint32_t chunk, uint32_t grid,
                                 std::vector<BatchJob>& out, uint32_t& total)
{
    if (njobs && !jobs) return fail(CGS_ERR_INVALID, "batch: null jobs");
    uint64_t sum = 0;
    out.clear();
    for (uint32_t j = 0; j < njobs; j++)
    {
        const cgs_batch_job& q = jobs[j];
        if (!q.scene || q.scene->device != ctx->device)
            return fail(CGS_ERR_INVALID, "batch: job " + std::to_string(j) + " has no scene on this context's device");
        if (q.n == 0) continue;
        if (!q.rays || (q.any ? !q.occluded : !q.hits))
            return fail(CGS_ERR_INVALID, "batch: job " + std::to_string(j) + " lacks rays or outputs");
        BatchJob b {};
        b.sc = q.scene->view();
        b.rays = q.rays;
        b.hits = q.hits;
        b.objects = q.objects;
        b.occluded = q.occluded;
        b.n = q.n;
        b.rayBase = (uint32_t)sum;
        b.any = q.any ? 1u : 0u;
        out.push_back(b);
        sum += q.n;
        if (!counterFits(sum, grid, chunk))
            return fail(CGS_ERR_INVALID, "batch: " + std::to_string(sum) + " rays exceed the 32-bit work counter");
    }
    total = (uint32_t)sum;
    return CGS_OK;
}

static cgs_status batchGrid(cgs_ctx* ctx, uint32_t blocksPerCU, uint32_t& grid)
{
    const int ki = (int)PersistKernel::Batch;
    if (ctx->grid[ki] == 0)
    {
        uint32_t bpc = 1;
        HIP_TRY(persistentGrid(PersistKernel::Batch, bpc));
        ctx->grid[ki] = bpc * (uint32_t)ctx->numCUs;
    }
    const uint32_t maxBpc = ctx->grid[ki] / (uint32_t)ctx->numCUs;
    const uint32_t bpc = blocksPerCU ? std::min(blocksPerCU, maxBpc) : maxBpc;
    grid = bpc * (uint32_t)ctx->numCUs;
    return CGS_OK;
}

static cgs_status validChunk(uint32_t chunk)
{
    if (chunk == 0 || chunk % 64 || chunk > 4096) return fail(CGS_ERR_INVALID, "batch: chunk must be a multiple of 64 in [64, 4096]");
    return CGS_OK;
}

// Whether some job's scene needs the kernels' wide-hash and surface-cull paths.
static bool anyBig(const std::vector<BatchJob>& table)
{
    for (const BatchJob& b : table)
        if (b.sc.big) return true;
    return false;
}

// One batched launch through the context's own buffers (waits for the
// previous one to finish before reusing them).
static cgs_status traceBatch(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, uint32_t chunk, uint32_t bpc,
                             hipStream_t s, const TraceMode& mode)
{
    uint32_t grid = 0;
    cgs_status st = batchGrid(ctx, bpc, grid);
    if (st != CGS_OK) return st;
    std::vector<BatchJob> table;
    uint32_t total = 0;
    if ((st = buildBatchJobs(ctx, jobs, njobs, chunk, grid, table, total)) != CGS_OK) return st;
    if (total == 0) return CGS_OK;
    if (ctx->batchLaunched) HIP_TRY(hipEventSynchronize(ctx->batchDone));
    if (ctx->batchTableCap < table.size())
    {
        HIP_TRY(ctx->batchTable.alloc(table.size() * sizeof(BatchJob)));
        ctx->batchTableCap = table.size();
    }
    // region stacks: the context's, sized for the largest persistent grid
    PersistArgs pa;
    if ((st = persistArgs(ctx, PersistKernel::Batch, s, total, pa)) != CGS_OK) return st;
    HIP_TRY(hipMemcpyAsync(ctx->batchTable.p, table.data(), table.size() * sizeof(BatchJob), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s)); // the host table is released on return
    HIP_TRY(launchTraceBatch(ctx->batchTable.as<BatchJob>(), (uint32_t)table.size(), total, chunk, pa.stack,
                             std::min<uint32_t>(pa.lanes, (uint32_t)ctx->stackLanes), pa.counter, grid,
                             ctx->dOverflow, s, mode, anyBig(table)));
    HIP_TRY(hipEventRecord(ctx->batchDone, s));
    ctx->batchLaunched = true;
    return persistEnd(ctx, s);
}

struct cgs_batch
{
    cgs_ctx* ctx = nullptr;
    DeviceBuffer table, counter, stack, overflow;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~cgs_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

extern "C" {

cgs_status cgs_trace_batch(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, void* stream)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_trace_batch: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    return traceBatch(ctx, jobs, njobs, ctx->batchChunk, ctx->batchBlocksPerCU, (hipStream_t)stream, ctx->batchMode);
}

cgs_status cgs_batch_create(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, cgs_batch** out)
{
    if (!ctx || !out) return fail(CGS_ERR_INVALID, "cgs_batch_create: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    uint32_t grid = 0;
    cgs_status st = batchGrid(ctx, ctx->batchBlocksPerCU, grid);
    if (st != CGS_OK) return st;
    std::vector<BatchJob> table;
    uint32_t total = 0;
    if ((st = buildBatchJobs(ctx, jobs, njobs, ctx->batchChunk, grid, table, total)) != CGS_OK) return st;
    auto b = std::make_unique<cgs_batch>();
    b->ctx = ctx;
    const uint32_t lanes = grid * 256;
    HIP_TRY(b->table.alloc(std::max<size_t>(1, table.size()) * sizeof(BatchJob)));
    HIP_TRY(b->counter.alloc(256));
    HIP_TRY(b->overflow.alloc(256));
    HIP_TRY(b->stack.alloc((size_t)lanes * CGS_CLIP_STACK * kStackWords * sizeof(float2)));
    if (!table.empty())
        HIP_TRY(hipMemcpy(b->table.p, table.data(), table.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(b->overflow.p, 0, 256));
    // capture: counter cleared by a fill kernel (no memset node), then the launch
    hipStream_t cs = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = launchFill(b->counter.as<uint32_t>(), 0u, 1, cs);
    if (e == hipSuccess)
        e = launchTraceBatch(b->table.as<BatchJob>(), (uint32_t)table.size(), total, ctx->batchChunk,
                             b->stack.as<float2>(), lanes, b->counter.as<unsigned int>(), grid,
                             b->overflow.as<unsigned long long>(), cs, ctx->batchMode, anyBig(table));
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess) e = e2;
    b->graph = g;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return fail(CGS_ERR_HIP, std::string("cgs_batch_create: graph capture: ") + hipGetErrorString(e));
    *out = b.release();
    return CGS_OK;
}

cgs_status cgs_batch_launch(cgs_batch* batch, void* stream)
{
    if (!batch) return fail(CGS_ERR_INVALID, "cgs_batch_launch: null batch");
    HIP_TRY(hipSetDevice(batch->ctx->device));
    HIP_TRY(hipGraphLaunch(batch->exec, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_batch_destroy(cgs_batch* batch)
{
    if (!batch) return CGS_OK;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipDeviceSynchronize();
    delete batch;
    return CGS_OK;
}

cgs_status cgs_batch_set_config(cgs_ctx* ctx, const cgs_batch_config* c)
{
    if (!ctx || !c) return fail(CGS_ERR_INVALID, "cgs_batch_set_config: null argument");
    cgs_status st = validChunk(c->chunk);
    if (st != CGS_OK) return st;
    std::lock_guard lk(ctx->mu);
    ctx->batchChunk = c->chunk;
    ctx->batchBlocksPerCU = c->blocks_per_cu;
    return CGS_OK;
}

cgs_status cgs_batch_get_config(cgs_ctx* ctx, cgs_batch_config* c)
{
    if (!ctx || !c) return fail(CGS_ERR_INVALID, "cgs_batch_get_config: null argument");
    std::lock_guard lk(ctx->mu);
    c->chunk = ctx->batchChunk;
    c->blocks_per_cu = ctx->batchBlocksPerCU;
    return CGS_OK;
}

cgs_status cgs_batch_autotune(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs, cgs_batch_config* best)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_batch_autotune: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    uint32_t grid = 0;
    cgs_status st = batchGrid(ctx, 0, grid);
    if (st != CGS_OK) return st;
    const uint32_t maxBpc = grid / (uint32_t)ctx->numCUs;
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    float bestMs = std::numeric_limits<float>::infinity();
    cgs_batch_config pick {ctx->batchChunk, ctx->batchBlocksPerCU};
    for (uint32_t chunk : {64u, 128u, 256u})
        for (uint32_t bpc = maxBpc; bpc >= 1; bpc--)
        {
            float ms = 0.0f;
            for (int rep = 0; rep < 2 && st == CGS_OK; rep++) // first run warms up, second is timed
            {
                hipError_t e = hipEventRecord(t0, s);
                if (e != hipSuccess) { st = fail(CGS_ERR_HIP, hipGetErrorString(e)); break; }
                st = traceBatch(ctx, jobs, njobs, chunk, bpc, s, ctx->batchMode);
                if (st != CGS_OK) break;
                e = hipEventRecord(t1, s);
                if (e == hipSuccess) e = hipEventSynchronize(t1);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
                if (e != hipSuccess) st = fail(CGS_ERR_HIP, hipGetErrorString(e));
            }
            if (st != CGS_OK) break;
            if (ms < bestMs)
            {
                bestMs = ms;
                pick = {chunk, bpc == maxBpc ? 0u : bpc};
            }
            if (bpc == 1) break;
        }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipStreamDestroy(s);
    if (st != CGS_OK) return st;
    ctx->batchChunk = pick.chunk;
    ctx->batchBlocksPerCU = pick.blocks_per_cu;
    if (best) *best = pick;
    return CGS_OK;
}

cgs_status cgs_batch_set_trace_options(cgs_ctx* ctx, const cgs_trace_options* options)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_batch_set_trace_options: null context");
    TraceMode m;
    cgs_status st = traceMode(options, m, "cgs_batch_set_trace_options");
    if (st != CGS_OK) return st;
    std::lock_guard lk(ctx->mu);
    ctx->batchMode = m;
    ctx->batchOptions = options ? *options : cgs_trace_options {};
    return CGS_OK;
}

cgs_status cgs_batch_get_trace_options(cgs_ctx* ctx, cgs_trace_options* options)
{
    if (!ctx || !options) return fail(CGS_ERR_INVALID, "cgs_batch_get_trace_options: null argument");
    std::lock_guard lk(ctx->mu);
    *options = ctx->batchOptions;
    return CGS_OK;
}

} // extern "C"

// The outputs of a batch's jobs, read back to the host.
namespace
{
struct JobResults
{
    std::vector<cgs_hit> hits;
    std::vector<int32_t> objects;
    std::vector<uint32_t> occluded;
};

cgs_status readJobResults(const cgs_batch_job* jobs, uint32_t njobs, std::vector<JobResults>& out)
{
    out.resize(njobs);
    for (uint32_t j = 0; j < njobs; j++)
    {
        const cgs_batch_job& q = jobs[j];
        JobResults& r = out[j];
        if (q.n == 0) continue;
        if (q.any)
        {
            r.occluded.resize(q.n);
            HIP_TRY(hipMemcpy(r.occluded.data(), q.occluded, (size_t)q.n * 4, hipMemcpyDeviceToHost));
        }
        else
        {
            r.hits.resize(q.n);
            HIP_TRY(hipMemcpy(r.hits.data(), q.hits, (size_t)q.n * sizeof(cgs_hit), hipMemcpyDeviceToHost));
            if (q.objects)
            {
                r.objects.resize(q.n);
                HIP_TRY(hipMemcpy(r.objects.data(), q.objects, (size_t)q.n * 4, hipMemcpyDeviceToHost));
            }
        }
    }
    return CGS_OK;
}

// A candidate's results against the reference's: the largest relative t and
// (u,v) differences over rays both hit on the same patch, and the fraction
// of rays outside the limits: hit or miss, object or occlusion differ, the
// patch differs with t beyond max_rel_t (a hit on the neighbouring patch of a
// seam at the same distance agrees; its (u,v) is on another patch and is not
// compared), or t or (u,v) beyond the limits on the same patch.
void compareResults(const cgs_batch_job* jobs, uint32_t njobs, const std::vector<JobResults>& ref,
                    const std::vector<JobResults>& got, float maxRelT, float maxUV, cgs_autotune_entry& e)
{
    double relT = 0.0, uv = 0.0;
    uint64_t disagree = 0, total = 0;
    for (uint32_t j = 0; j < njobs; j++)
    {
        const cgs_batch_job& q = jobs[j];
        total += q.n;
        if (q.any)
        {
            for (uint32_t k = 0; k < q.n; k++) disagree += ref[j].occluded[k] != got[j].occluded[k];
            continue;
        }
        for (uint32_t k = 0; k < q.n; k++)
        {
            const cgs_hit& a = ref[j].hits[k];
            const cgs_hit& b = got[j].hits[k];
            const bool ha = a.patch != EMPTY && a.patch != CGS_PATCH_OVERFLOW;
            const bool hb = b.patch != EMPTY && b.patch != CGS_PATCH_OVERFLOW;
            if (!ha && !hb) continue;
            if (ha != hb || (!ref[j].objects.empty() && ref[j].objects[k] != got[j].objects[k]))
            {
                disagree++;
                continue;
            }
            const double dt = std::fabs((double)b.t - (double)a.t) / std::max(std::fabs((double)a.t), 1e-30);
            if (a.patch != b.patch)
            {
                if (dt > maxRelT) disagree++;
                continue;
            }
            const double duv = std::max(std::fabs((double)b.u - a.u), std::fabs((double)b.v - a.v));
            relT = std::max(relT, dt);
            uv = std::max(uv, duv);
            if (dt > maxRelT || duv > maxUV) disagree++;
        }
    }
    e.max_rel_t = (float)relT;
    e.max_uv = (float)uv;
    e.disagree = total ? (float)((double)disagree / (double)total) : 0.0f;
}
} // namespace

extern "C" {

cgs_status cgs_batch_autotune_accuracy(cgs_ctx* ctx, const cgs_batch_job* jobs, uint32_t njobs,
                                       const cgs_accuracy_limits* limits, cgs_trace_options* best,
                                       cgs_autotune_entry* report, uint32_t report_cap, uint32_t* report_count)
{
    if (!ctx || !limits || (njobs && !jobs) || (report_cap && !report))
        return fail(CGS_ERR_INVALID, "cgs_batch_autotune_accuracy: null argument");
    if (!(limits->max_rel_t >= 0.0f) || !(limits->max_uv >= 0.0f) || !(limits->max_disagree >= 0.0f))
        return fail(CGS_ERR_INVALID, "cgs_batch_autotune_accuracy: negative or NaN limit");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard lk(ctx->mu);
    // candidates: the reference's clipping, clipping with a pass cap, and
    // subdivision at several depths, each also with the FP64 polish if allowed
    std::vector<cgs_trace_options> cands;
    for (uint32_t polish : {0u, 1u})
    {
        if (polish && !limits->allow_polish) break;
        cands.push_back({CGS_METHOD_CLIPPING, 0.0f, 0, 0, polish, {}});
        for (uint32_t it : {8u, 12u, 16u, 24u, 32u}) cands.push_back({CGS_METHOD_CLIPPING, 0.0f, 0, it, polish, {}});
        for (uint32_t d : {4u, 5u, 6u, 8u, 10u, 12u}) cands.push_back({CGS_METHOD_SUBDIVISION, 0.0f, d, 0, polish, {}});
    }
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    cgs_status st = CGS_OK;
    // the reference results: the reference's clipping, polished when allowed
    std::vector<JobResults> ref, got;
    {
        TraceMode rm;
        rm.polish = limits->allow_polish ? 1u : 0u;
        st = traceBatch(ctx, jobs, njobs, ctx->batchChunk, ctx->batchBlocksPerCU, s, rm);
        if (st == CGS_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(CGS_ERR_HIP, "cgs_batch_autotune_accuracy");
        if (st == CGS_OK) st = readJobResults(jobs, njobs, ref);
    }
    std::vector<cgs_autotune_entry> entries;
    int pick = -1;
    for (size_t c = 0; c < cands.size() && st == CGS_OK; c++)
    {
        TraceMode m;
        if ((st = traceMode(&cands[c], m, "cgs_batch_autotune_accuracy")) != CGS_OK) break;
        cgs_autotune_entry e {};
        e.options = cands[c];
        // one run for the results, then the faster of two timed runs
        st = traceBatch(ctx, jobs, njobs, ctx->batchChunk, ctx->batchBlocksPerCU, s, m);
        if (st == CGS_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(CGS_ERR_HIP, "cgs_batch_autotune_accuracy");
        if (st == CGS_OK) st = readJobResults(jobs, njobs, got);
        if (st != CGS_OK) break;
        compareResults(jobs, njobs, ref, got, limits->max_rel_t, limits->max_uv, e);
        float bestMs = std::numeric_limits<float>::infinity();
        for (int rep = 0; rep < 2 && st == CGS_OK; rep++)
        {
            hipError_t he = hipEventRecord(t0, s);
            if (he == hipSuccess) st = traceBatch(ctx, jobs, njobs, ctx->batchChunk, ctx->batchBlocksPerCU, s, m);
            if (st != CGS_OK) break;
            if (he == hipSuccess) he = hipEventRecord(t1, s);
            if (he == hipSuccess) he = hipEventSynchronize(t1);
            float ms = 0.0f;
            if (he == hipSuccess) he = hipEventElapsedTime(&ms, t0, t1);
            if (he != hipSuccess) st = fail(CGS_ERR_HIP, hipGetErrorString(he));
            bestMs = std::min(bestMs, ms);
        }
        e.ms = bestMs;
        e.admissible = e.disagree <= limits->max_disagree ? 1u : 0u;
        entries.push_back(e);
        if (e.admissible && (pick < 0 || e.ms < entries[pick].ms)) pick = (int)entries.size() - 1;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipStreamDestroy(s);
    if (st != CGS_OK) return st;
    if (report_count) *report_count = (uint32_t)entries.size();
    for (size_t k = 0; k < entries.size() && k < report_cap; k++) report[k] = entries[k];
    // the reference candidate agrees with the reference exactly, so one is always admissible
    if (pick < 0) return fail(CGS_ERR_INVALID, "cgs_batch_autotune_accuracy: no candidate within the limits");
    TraceMode m;
    (void)traceMode(&entries[pick].options, m, "cgs_batch_autotune_accuracy");
    ctx->batchMode = m;
    ctx->batchOptions = entries[pick].options;
    if (best) *best = entries[pick].options;
    return CGS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// serving: continuous batching over resident scenes
// ---------------------------------------------------------------------------
namespace
{

using Clock = std::chrono::steady_clock;

enum class RequestKind
{
    Closest,
    Any,
    Shade,
};

struct ServerRequest
{
    uint64_t id = 0;
    uint32_t scene = 0, n = 0;
    RequestKind kind = RequestKind::Closest;
    const cgs_ray* rays = nullptr;
    cgs_hit* hits = nullptr;
    int32_t* objects = nullptr;
    uint32_t* occluded = nullptr;
    float* color = nullptr;        // shaded requests
    uint32_t* pixels = nullptr;
    float eps = 0.0f;
    int32_t depth = 0;
    size_t bytes = 0;              // device bytes while in flight
    Clock::time_point submitted;
    uint64_t rayOff = 0;           // position in the slot's launch (among requests of its kernel)
    uint64_t outOff = 0;           // byte offset of its results in the slot's output buffer
};

// Result bytes per ray: a 16-B hit + a 4-B object id (or occlusion flag);
// shaded requests add 12 B of colour and a 4-B packed pixel.
inline size_t resultBytes(RequestKind k) { return sizeof(cgs_hit) + 4 + (k == RequestKind::Shade ? 16 : 0); }
// Device bytes of a request: rays in, results out.
inline size_t requestBytes(uint32_t n, RequestKind k) { return (size_t)n * (sizeof(cgs_ray) + resultBytes(k)); }

struct PinnedBuffer
{
    void* p = nullptr;
    size_t cap = 0;
    ~PinnedBuffer()
    {
        if (p) (void)hipHostFree(p);
    }
    hipError_t reserve(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
};

// Host copies of one launch, cut into 4-MiB pieces and spread over worker
// threads when there is more than a few MiB to move.
struct CopyList
{
    struct Part
    {
        void* d;
        const void* s;
        size_t n;
    };
    std::vector<Part> parts;
    size_t bytes = 0;
    void add(void* d, const void* s, size_t n)
    {
        constexpr size_t kPiece = (size_t)4 << 20;
        for (size_t o = 0; o < n; o += kPiece)
            parts.push_back({(char*)d + o, (const char*)s + o, std::min(kPiece, n - o)});
        bytes += n;
    }
    void run()
    {
        if (bytes < ((size_t)8 << 20))
            for (const Part& p : parts) std::memcpy(p.d, p.s, p.n);
        else
            parallelFor(parts.size(), [&](size_t i) { std::memcpy(parts[i].d, parts[i].s, parts[i].n); });
        parts.clear();
        bytes = 0;
    }
};

struct ServerSlot
{
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DeviceBuffer rays, out, table, shadeTable, counter, stack, overflow;
    size_t rayCap = 0, outCap = 0, tableCap = 0, shadeTableCap = 0;
    PinnedBuffer hRays, hOut, hTable, hShadeTable;
    std::vector<ServerRequest> reqs;
    size_t bytes = 0;
    uint64_t total = 0;
    bool busy = false;
};

struct Done
{
    uint64_t latencyNs;
    int ok;
};

} // namespace

struct cgs_server
{
    cgs_ctx* ctx = nullptr;
    std::vector<cgs_scene*> scenes;
    std::vector<SceneDev> views;
    size_t budget = 0;
    uint32_t maxBatchRays = 0;
    uint32_t grid = 0, lanes = 0, chunk = 64;
    uint32_t shadeGrid = 0;        // kShade's resident workgroups (at most grid: the stacks hold grid × 256 lanes)
    std::vector<ServerSlot> slots;
    std::mutex mu;
    std::condition_variable wake, finished;
    std::deque<ServerRequest> pending;
    // Requests not yet done (queued or in a launch), and done requests not
    // yet collected by cgs_server_wait / cgs_server_poll, which erase them.
    std::unordered_set<uint64_t> open;
    std::unordered_map<uint64_t, Done> done;
    uint64_t nextId = 1;
    size_t inflight = 0;
    bool stop = false;
    std::thread worker;

    // Copies a finished slot's results to the requests' host buffers.
    void complete(ServerSlot& sl, int ok)
    {
        const uint8_t* base = (const uint8_t*)sl.hOut.p;
        if (ok)
        {
            CopyList copies;
            for (const ServerRequest& r : sl.reqs)
            {
                // per request: hits, then object ids (or flags), then colours and pixels
                const uint8_t* o = base + r.outOff;
                const size_t n = r.n;
                if (r.kind == RequestKind::Any)
                    copies.add(r.occluded, o + n * sizeof(cgs_hit), n * 4);
                else
                {
                    if (r.hits) copies.add(r.hits, o, n * sizeof(cgs_hit));
                    if (r.objects) copies.add(r.objects, o + n * sizeof(cgs_hit), n * 4);
                }
                if (r.kind == RequestKind::Shade)
                {
                    copies.add(r.color, o + n * (sizeof(cgs_hit) + 4), n * 12);
                    if (r.pixels) copies.add(r.pixels, o + n * (sizeof(cgs_hit) + 16), n * 4);
                }
            }
            copies.run();
        }
        const Clock::time_point now = Clock::now();
        for (const ServerRequest& r : sl.reqs)
        {
            std::lock_guard lk(mu);
            done[r.id] = {(uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(now - r.submitted).count(), ok};
            open.erase(r.id);
            inflight -= r.bytes;
        }
        sl.reqs.clear();
        sl.busy = false;
        finished.notify_all();
    }

    // One batched launch of the slot's requests: rays up, then kTraceBatch
    // over the closest and occlusion requests and kShade over the shaded
    // ones (each numbering its own requests' rays from 0, with its own
    // counter), results down.
    hipError_t launch(ServerSlot& sl)
    {
        uint64_t total = 0, outBytes = 0;
        size_t nShade = 0;
        for (ServerRequest& r : sl.reqs)
        {
            r.rayOff = total;
            r.outOff = outBytes;
            total += r.n;
            outBytes += (uint64_t)r.n * resultBytes(r.kind);
            nShade += r.kind == RequestKind::Shade;
        }
        const size_t nTrace = sl.reqs.size() - nShade;
        sl.total = total;
        hipError_t e = hipSuccess;
        if (sl.rayCap < total)
        {
            if ((e = sl.rays.alloc(total * sizeof(cgs_ray))) != hipSuccess) return e;
            sl.rayCap = total;
        }
        if (sl.outCap < outBytes)
        {
            if ((e = sl.out.alloc(outBytes)) != hipSuccess) return e;
            sl.outCap = outBytes;
        }
        if (sl.tableCap < nTrace)
        {
            if ((e = sl.table.alloc(nTrace * sizeof(BatchJob))) != hipSuccess) return e;
            sl.tableCap = nTrace;
        }
        if (sl.shadeTableCap < nShade)
        {
            if ((e = sl.shadeTable.alloc(nShade * sizeof(ShadeJob))) != hipSuccess) return e;
            sl.shadeTableCap = nShade;
        }
        if ((e = sl.hRays.reserve(total * sizeof(cgs_ray))) != hipSuccess) return e;
        if ((e = sl.hOut.reserve(outBytes)) != hipSuccess) return e;
        if ((e = sl.hTable.reserve(std::max<size_t>(1, nTrace) * sizeof(BatchJob))) != hipSuccess) return e;
        if ((e = sl.hShadeTable.reserve(std::max<size_t>(1, nShade) * sizeof(ShadeJob))) != hipSuccess) return e;
        cgs_ray* hr = (cgs_ray*)sl.hRays.p;
        BatchJob* jt = (BatchJob*)sl.hTable.p;
        ShadeJob* js = (ShadeJob*)sl.hShadeTable.p;
        uint8_t* dOut = sl.out.as<uint8_t>();
        cgs_ray* dRays = sl.rays.as<cgs_ray>();
        CopyList copies;
        uint32_t traceRays = 0, shadeRays = 0, kt = 0, ks = 0;
        for (const ServerRequest& r : sl.reqs)
        {
            copies.add(hr + r.rayOff, r.rays, (size_t)r.n * sizeof(cgs_ray));
            uint8_t* o = dOut + r.outOff;
            const size_t n = r.n;
            if (r.kind == RequestKind::Shade)
            {
                ShadeJob j {};
                j.sc = views[r.scene];
                j.rays = dRays + r.rayOff;
                j.hits = r.hits ? (cgs_hit*)o : nullptr;
                j.objects = r.objects ? (int32_t*)(o + n * sizeof(cgs_hit)) : nullptr;
                j.color = (float*)(o + n * (sizeof(cgs_hit) + 4));
                j.pixels = r.pixels ? (uint32_t*)(o + n * (sizeof(cgs_hit) + 16)) : nullptr;
                j.n = r.n;
                j.rayBase = shadeRays;
                j.eps = r.eps;
                j.depth = r.depth;
                js[ks++] = j;
                shadeRays += r.n;
            }
            else
            {
                const bool any = r.kind == RequestKind::Any;
                BatchJob b {};
                b.sc = views[r.scene];
                b.rays = dRays + r.rayOff;
                b.hits = any ? nullptr : (cgs_hit*)o;
                b.objects = any ? nullptr : (int32_t*)(o + n * sizeof(cgs_hit));
                b.occluded = any ? (uint32_t*)(o + n * sizeof(cgs_hit)) : nullptr;
                b.n = r.n;
                b.rayBase = traceRays;
                b.any = any ? 1u : 0u;
                jt[kt++] = b;
                traceRays += r.n;
            }
        }
        copies.run();
        if ((e = hipMemcpyAsync(sl.rays.p, hr, total * sizeof(cgs_ray), hipMemcpyHostToDevice, sl.stream)) != hipSuccess)
            return e;
        if ((e = launchFill(sl.counter.as<uint32_t>(), 0u, 2, sl.stream)) != hipSuccess) return e;
        if (nTrace)
        {
            if ((e = hipMemcpyAsync(sl.table.p, jt, nTrace * sizeof(BatchJob), hipMemcpyHostToDevice, sl.stream)) !=
                hipSuccess)
                return e;
            bool big = false;
            for (size_t k = 0; k < nTrace; k++) big |= jt[k].sc.big != 0;
            if ((e = launchTraceBatch(sl.table.as<BatchJob>(), (uint32_t)nTrace, traceRays, chunk, sl.stack.as<float2>(),
                                      lanes, sl.counter.as<unsigned int>(), grid, sl.overflow.as<unsigned long long>(),
                                      sl.stream, TraceMode {}, big)) != hipSuccess)
                return e;
        }
        if (nShade)
        {
            if ((e = hipMemcpyAsync(sl.shadeTable.p, js, nShade * sizeof(ShadeJob), hipMemcpyHostToDevice,
                                    sl.stream)) != hipSuccess)
                return e;
            if ((e = launchShade(nullptr, sl.shadeTable.as<ShadeJob>(), (uint32_t)nShade, shadeRays, chunk,
                                 sl.stack.as<float2>(), lanes, sl.counter.as<unsigned int>() + 1, shadeGrid, nullptr,
                                 sl.stream)) != hipSuccess)
                return e;
        }
        if ((e = hipMemcpyAsync(sl.hOut.p, sl.out.p, outBytes, hipMemcpyDeviceToHost, sl.stream)) != hipSuccess) return e;
        if ((e = hipEventRecord(sl.done, sl.stream)) != hipSuccess) return e;
        sl.busy = true;
        return hipSuccess;
    }

    void run()
    {
        (void)hipSetDevice(ctx->device);
        while (true)
        {
            bool anyBusy = false;
            for (ServerSlot& sl : slots)
            {
                if (!sl.busy) continue;
                const hipError_t q = hipEventQuery(sl.done);
                if (q == hipErrorNotReady)
                    anyBusy = true;
                else
                    complete(sl, q == hipSuccess ? 1 : 0);
            }
            for (ServerSlot& sl : slots)
            {
                if (sl.busy) continue;
                {
                    std::lock_guard lk(mu);
                    uint64_t rays = 0;
                    // admit in arrival order while the byte budget (and ray cap) allows
                    while (!pending.empty())
                    {
                        const ServerRequest& r = pending.front();
                        if (inflight + r.bytes > budget && !(sl.reqs.empty() && inflight == 0)) break;
                        if (maxBatchRays && !sl.reqs.empty() && rays + r.n > maxBatchRays) break;
                        if (!counterFits(rays + r.n, grid, chunk)) break;
                        inflight += r.bytes;
                        rays += r.n;
                        sl.reqs.push_back(r);
                        pending.pop_front();
                    }
                }
                if (sl.reqs.empty()) continue;
                if (launch(sl) != hipSuccess)
                {
                    (void)hipStreamSynchronize(sl.stream);
                    complete(sl, 0);
                }
                else
                    anyBusy = true;
            }
            std::unique_lock lk(mu);
            if (stop && pending.empty() && !anyBusy) break;
            if (anyBusy)
                wake.wait_for(lk, std::chrono::microseconds(50));
            else
                wake.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
};

extern "C" {

cgs_status cgs_server_create(cgs_ctx* ctx, cgs_scene* const* scenes, uint32_t nscenes, const cgs_server_options* opt,
                             cgs_server** out)
{
    if (!ctx || !out || (nscenes && !scenes)) return fail(CGS_ERR_INVALID, "cgs_server_create: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    auto sv = std::make_unique<cgs_server>();
    sv->ctx = ctx;
    for (uint32_t k = 0; k < nscenes; k++)
    {
        if (!scenes[k] || scenes[k]->device != ctx->device)
            return fail(CGS_ERR_INVALID, "cgs_server_create: scene on another device");
        sv->scenes.push_back(scenes[k]);
        sv->views.push_back(scenes[k]->view());
    }
    sv->budget = opt && opt->hbm_budget ? opt->hbm_budget : ((size_t)1 << 30);
    sv->maxBatchRays = opt ? opt->max_batch_rays : 0;
    const uint32_t nslots = opt && opt->slots ? opt->slots : 3;
    {
        std::lock_guard lk(ctx->mu);
        cgs_status st = batchGrid(ctx, ctx->batchBlocksPerCU, sv->grid);
        if (st != CGS_OK) return st;
        sv->chunk = ctx->batchChunk;
    }
    sv->lanes = sv->grid * 256;
    {
        uint32_t bpc = 1;
        HIP_TRY(persistentGrid(PersistKernel::Shade, bpc));
        sv->shadeGrid = std::min(sv->grid, bpc * (uint32_t)ctx->numCUs);
    }
    sv->slots.resize(nslots);
    for (ServerSlot& sl : sv->slots)
    {
        HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        HIP_TRY(sl.counter.alloc(256));
        HIP_TRY(sl.overflow.alloc(256));
        HIP_TRY(hipMemset(sl.overflow.p, 0, 256));
        HIP_TRY(sl.stack.alloc((size_t)sv->lanes * CGS_CLIP_STACK * kStackWords * sizeof(float2)));
    }
    cgs_server* raw = sv.get();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = sv.release();
    return CGS_OK;
}

// Queues a checked request (the submit entry points fill in its outputs).
static cgs_status serverQueue(cgs_server* sv, ServerRequest& r, uint64_t* id, const char* who)
{
    if (r.scene >= sv->scenes.size()) return fail(CGS_ERR_INVALID, std::string(who) + ": no such scene");
    r.bytes = requestBytes(r.n, r.kind);
    if (r.bytes > sv->budget) return fail(CGS_ERR_INVALID, std::string(who) + ": request exceeds the byte budget");
    // a request must fit one launch's 32-bit work counter on its own, or it
    // would never be admitted and would block every request behind it
    if (!counterFits(r.n, sv->grid, sv->chunk))
        return fail(CGS_ERR_INVALID, std::string(who) + ": request exceeds the 32-bit work counter of a launch");
    r.submitted = Clock::now();
    {
        std::lock_guard lk(sv->mu);
        if (sv->stop) return fail(CGS_ERR_INVALID, std::string(who) + ": server stopping");
        r.id = sv->nextId++;
        *id = r.id;
        if (r.n == 0)
            sv->done[r.id] = {0, 1};
        else
        {
            sv->open.insert(r.id);
            sv->pending.push_back(r);
        }
    }
    sv->wake.notify_one();
    return CGS_OK;
}

cgs_status cgs_server_submit(cgs_server* sv, uint32_t scene, const cgs_ray* rays, uint32_t n, int32_t any,
                             cgs_hit* hits, int32_t* objects, uint32_t* occluded, uint64_t* id)
{
    if (!sv || !id) return fail(CGS_ERR_INVALID, "cgs_server_submit: null argument");
    if (n && (!rays || (any ? !occluded : !hits))) return fail(CGS_ERR_INVALID, "cgs_server_submit: missing buffers");
    ServerRequest r;
    r.scene = scene;
    r.n = n;
    r.kind = any ? RequestKind::Any : RequestKind::Closest;
    r.rays = rays;
    r.hits = hits;
    r.objects = objects;
    r.occluded = occluded;
    return serverQueue(sv, r, id, "cgs_server_submit");
}

cgs_status cgs_server_submit_shaded(cgs_server* sv, uint32_t scene, const cgs_ray* rays, uint32_t n,
                                    const cgs_render_options* options, float* color, uint32_t* pixels, cgs_hit* hits,
                                    int32_t* objects, uint64_t* id)
{
    if (!sv || !id || !options) return fail(CGS_ERR_INVALID, "cgs_server_submit_shaded: null argument");
    if (n && (!rays || !color)) return fail(CGS_ERR_INVALID, "cgs_server_submit_shaded: missing buffers");
    ServerRequest r;
    r.scene = scene;
    r.n = n;
    r.kind = RequestKind::Shade;
    r.rays = rays;
    r.hits = hits;
    r.objects = objects;
    r.color = color;
    r.pixels = pixels;
    r.eps = options->eps;
    r.depth = options->recursion_depth;
    return serverQueue(sv, r, id, "cgs_server_submit_shaded");
}

int cgs_server_poll(cgs_server* sv, uint64_t id, uint64_t* latency_ns)
{
    if (!sv) return -1;
    std::lock_guard lk(sv->mu);
    auto it = sv->done.find(id);
    if (it == sv->done.end()) return sv->open.count(id) ? 0 : -1;
    const Done d = it->second;
    sv->done.erase(it);
    if (latency_ns) *latency_ns = d.latencyNs;
    return d.ok ? 1 : -1;
}

cgs_status cgs_server_wait(cgs_server* sv, uint64_t id, uint64_t* latency_ns)
{
    if (!sv) return fail(CGS_ERR_INVALID, "cgs_server_wait: null server");
    std::unique_lock lk(sv->mu);
    if (!sv->done.count(id) && !sv->open.count(id))
        return fail(CGS_ERR_INVALID, "cgs_server_wait: unknown or already collected request");
    sv->finished.wait(lk, [&] { return sv->done.count(id) != 0; });
    const Done d = sv->done[id];
    sv->done.erase(id);
    if (latency_ns) *latency_ns = d.latencyNs;
    return d.ok ? CGS_OK : fail(CGS_ERR_HIP, "cgs_server_wait: the request's launch failed");
}

cgs_status cgs_server_destroy(cgs_server* sv)
{
    if (!sv) return CGS_OK;
    {
        std::lock_guard lk(sv->mu);
        sv->stop = true;
    }
    sv->wake.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->ctx->device);
    for (ServerSlot& sl : sv->slots)
    {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream), (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    delete sv;
    return CGS_OK;
}

cgs_status cgs_malloc(cgs_ctx* ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return fail(CGS_ERR_INVALID, "cgs_malloc: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(dptr, std::max<size_t>(bytes, 1));
    if (e == hipErrorOutOfMemory) return fail(CGS_ERR_OUT_OF_MEMORY, "cgs_malloc: out of memory");
    HIP_TRY(e);
    return CGS_OK;
}

cgs_status cgs_free(cgs_ctx* ctx, void* dptr)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_free: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipFree(dptr));
    return CGS_OK;
}

cgs_status cgs_memcpy_h2d(cgs_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_memcpy_h2d: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_memcpy_d2h(cgs_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_memcpy_d2h: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_memset(cgs_ctx* ctx, void* dst, int value, size_t bytes, void* stream)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_memset: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_stream_create(cgs_ctx* ctx, void** stream)
{
    if (!ctx || !stream) return fail(CGS_ERR_INVALID, "cgs_stream_create: null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return CGS_OK;
}

cgs_status cgs_stream_destroy(cgs_ctx* ctx, void* stream)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_stream_destroy: null context");
    if (!stream) return CGS_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return CGS_OK;
}

cgs_status cgs_stream_sync(cgs_ctx* ctx, void* stream)
{
    if (!ctx) return fail(CGS_ERR_INVALID, "cgs_stream_sync: null context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return CGS_OK;
}

} // extern "C"
