// SPDX-License-Identifier: MIT
//
// Deterministic perfect-hash build for BVH node keys on gfx950
// (rt/PerfectHashFunction.{h,cu}, rt/BVH.cpp:24-38).
//
// Hash family h(k) = (A·k + B + D[k & mask]) mod T in 32-bit arithmetic (or,
// as an extension, the same function in exact arithmetic: slotOf below),
// buckets of at most 4 keys, (A,B) drawn by the host's std::mt19937 exactly as
// the reference draws them, up to 16 functions with the table growing by
// max(4, 0.05·n) after each failed one, at most 32 displacement attempts per
// bucket with stride (g+1).
//
// The reference resolves contention between buckets with atomics, so its
// output depends on thread timing.  Here contention is resolved in rounds:
// every pending bucket of the current size class proposes d = a·(g+1) (g =
// its rank by ascending bucket id, a = its attempt count); a proposal is valid
// when its slots are pairwise distinct and free; each contested slot goes to
// the lowest g; a bucket that wins all its slots is placed; every other
// pending bucket advances a.  The result depends only on the key set.
//
// One workgroup builds one hash function, so a scene's thousands of small
// BVHs are hashed by a single launch.  Per-job state lives in global scratch.
//
#include <hip/hip_runtime.h>

#include <mutex>

#include "cgs_device.hpp"
#include "cgs_internal.hpp"

namespace cgs
{

constexpr uint32_t kPending = 0, kPlaced = 1, kFailed = 2;
constexpr int kCap = 4, kMaxAttempts = 32, kMaxFunctions = 16;
constexpr int kPhfThreads = 256;

// The function's slot for a key and displacement: the reference's 32-bit
// arithmetic (PerfectHashFunction.h:58-63), or with WIDE the same function in
// exact arithmetic on keys of up to 64 bits (no counterpart in the
// reference): (A·(k mod T) + B + d) mod T = (A·k + B + d) mod T without the
// 32-bit wrap of A·k, which breaks the function's near-bijection on dense
// heap keys (every reference attempt fails at 2^21 keys).  T < 2^31, so the
// 64-bit sum cannot wrap.
template<bool WIDE, class KEY>
__device__ __forceinline__ uint32_t slotOf(uint32_t A, uint32_t B, KEY key, uint32_t d, uint32_t T)
{
    if constexpr (WIDE)
        return (uint32_t)(((uint64_t)A * (uint64_t)(key % T) + B + d) % T);
    else
        return (A * (uint32_t)key + B + d) % T;
}

template<class KEY>
__device__ __forceinline__ uint32_t bucketOf(KEY key, uint32_t mask)
{
    return (uint32_t)(key & mask);
}

template<class KEY, bool WIDE>
__global__ __launch_bounds__(kPhfThreads) void kPhfBuild(const PhfJob* __restrict__ jobs, const KEY* __restrict__ keys,
                                                         const uint32_t* __restrict__ ab, uint32_t* __restrict__ scratch,
                                                         uint32_t* __restrict__ displ, uint32_t* __restrict__ tables,
                                                         PhfResult* __restrict__ results)
{
    const PhfJob J = jobs[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    const uint32_t n = J.n, M = J.M, mask = J.mask;
    uint32_t* bsize = scratch + J.scratchOff;      // M
    KEY* bkeys = (KEY*)(bsize + M);                // 4M keys (scratch offsets and M are even)
    uint32_t* list = (uint32_t*)(bkeys + kCap * M); // M   (class lists, ascending bucket id)
    uint32_t* att = list + M;                      // M
    uint32_t* state = att + M;                     // M
    uint32_t* valid = state + M;                   // M
    uint32_t* slots = valid + M;                   // 4M
    uint32_t* occ = slots + kCap * M;              // Tcap
    uint32_t* claim = occ + J.tcap;                // Tcap
    uint32_t* D = displ + J.displOff;              // M
    uint32_t* table = tables + J.tableOff;         // Tcap

    __shared__ uint32_t sFlag, sFailures, sPending;
    __shared__ uint32_t sCount[kCap + 1], sBase[kCap + 1];
    __shared__ uint32_t sScan[kCap][kPhfThreads];

    if (tid == 0)
    {
        sFlag = 0;
        for (int s = 0; s <= kCap; s++) sCount[s] = 0;
    }
    for (uint32_t b = tid; b < M; b += kPhfThreads) bsize[b] = 0;
    __syncthreads();

    // 1. buckets (init_and_bucketsort, PerfectHashFunction.cu:27-55)
    for (uint32_t i = tid; i < n; i += kPhfThreads)
    {
        const KEY k = keys[J.keyOff + i];
        uint32_t j = bucketOf(k, mask);
        uint32_t s = atomicAdd(bsize + j, 1u);
        if (s < (uint32_t)kCap) bkeys[kCap * j + s] = k;
        else sFlag = 1;
    }
    __syncthreads();
    if (sFlag)
    {
        for (uint32_t b = tid; b < M; b += kPhfThreads) D[b] = 0;
        if (tid == 0) results[blockIdx.x] = {31, 19, n, 0, (uint32_t)CGS_ERR_BUCKET_OVERFLOW}; // member defaults (PerfectHashFunction.h:39-40)
        return;
    }

    // 2. size classes, each in ascending bucket id (a stable counting sort)
    const uint32_t chunk = (M + kPhfThreads - 1) / kPhfThreads;
    const uint32_t c0 = min(M, tid * chunk), c1 = min(M, c0 + chunk);
    uint32_t cnt[kCap + 1] = {0, 0, 0, 0, 0};
    for (uint32_t b = c0; b < c1; b++) cnt[bsize[b]]++;
    for (int s = 1; s <= kCap; s++) sScan[s - 1][tid] = cnt[s];
    __syncthreads();
    if (tid < (uint32_t)kCap)
    {
        uint32_t run = 0;
        for (int t = 0; t < kPhfThreads; t++)
        {
            uint32_t v = sScan[tid][t];
            sScan[tid][t] = run;
            run += v;
        }
        sCount[tid + 1] = run;
    }
    __syncthreads();
    if (tid == 0)
    {
        uint32_t base = 0;
        for (int s = kCap; s >= 1; s--)
        {
            sBase[s] = base;
            base += sCount[s];
        }
    }
    __syncthreads();
    {
        uint32_t pos[kCap + 1];
        for (int s = 1; s <= kCap; s++) pos[s] = sBase[s] + sScan[s - 1][tid];
        for (uint32_t b = c0; b < c1; b++)
        {
            uint32_t s = bsize[b];
            if (s) list[pos[s]++] = b;
        }
    }
    __syncthreads();

    // 3. functions (PerfectHashFunction.cu:383-484)
    uint32_t T = n;
    uint32_t A = 31, B = 19;
    int fn = 0;
    bool ok = false;
    for (; fn < kMaxFunctions; fn++)
    {
        A = ab[2 * (J.abOff + fn)];
        B = ab[2 * (J.abOff + fn) + 1];
        for (uint32_t i = tid; i < T; i += kPhfThreads)
        {
            occ[i] = 0;
            claim[i] = 0xFFFFFFFFu;
        }
        for (uint32_t b = tid; b < M; b += kPhfThreads) D[b] = 0;
        if (tid == 0) sFailures = 0;
        __syncthreads();
        for (int s = kCap; s >= 1; s--)
        {
            const uint32_t count = sCount[s], base = sBase[s];
            if (count == 0) continue;
            for (uint32_t g = tid; g < count; g += kPhfThreads)
            {
                att[base + g] = 0;
                state[base + g] = kPending;
                valid[base + g] = 0;
            }
            __syncthreads();
            for (int round = 0; round < kMaxAttempts; round++)
            {
                // propose
                for (uint32_t g = tid; g < count; g += kPhfThreads)
                {
                    const uint32_t e = base + g;
                    if (state[e] != kPending) continue;
                    const uint32_t b = list[e];
                    const uint32_t d = att[e] * (g + 1u);
                    uint32_t sl[kCap];
                    bool v = true;
                    for (int k = 0; k < s; k++)
                    {
                        sl[k] = slotOf<WIDE>(A, B, bkeys[kCap * b + k], d, T);
                        if (occ[sl[k]]) v = false;
                        for (int m = 0; m < k; m++)
                            if (sl[m] == sl[k]) v = false;
                        slots[kCap * e + k] = sl[k];
                    }
                    valid[e] = v;
                    if (v)
                        for (int k = 0; k < s; k++) atomicMin(claim + sl[k], g);
                }
                __syncthreads();
                // resolve
                if (tid == 0) sPending = 0;
                __syncthreads();
                for (uint32_t g = tid; g < count; g += kPhfThreads)
                {
                    const uint32_t e = base + g;
                    if (state[e] != kPending) continue;
                    bool won = valid[e] != 0;
                    if (won)
                        for (int k = 0; k < s; k++)
                            if (claim[slots[kCap * e + k]] != g) won = false;
                    if (won)
                    {
                        D[list[e]] = att[e] * (g + 1u);
                        state[e] = kPlaced + 8; // placed this round; occupy below
                    }
                    else if (++att[e] >= (uint32_t)kMaxAttempts)
                    {
                        state[e] = kFailed;
                        atomicAdd(&sFailures, 1u);
                    }
                    else
                        atomicAdd(&sPending, 1u);
                }
                __syncthreads();
                // release claims, occupy won slots
                for (uint32_t g = tid; g < count; g += kPhfThreads)
                {
                    const uint32_t e = base + g;
                    if (valid[e])
                    {
                        // every valid proposer of this round, whatever its outcome
                        for (int k = 0; k < s; k++) claim[slots[kCap * e + k]] = 0xFFFFFFFFu;
                        valid[e] = 0;
                    }
                    if (state[e] == kPlaced + 8)
                    {
                        for (int k = 0; k < s; k++) occ[slots[kCap * e + k]] = 1;
                        state[e] = kPlaced;
                    }
                }
                __syncthreads();
                if (sPending == 0) break;
            }
        }
        __syncthreads();
        if (sFailures == 0)
        {
            ok = true;
            break;
        }
        T = T + J.step;
        __syncthreads();
    }
    const uint32_t attempts = ok ? fn + 1 : kMaxFunctions;

    // 4. injectivity check and table fill (BVH::buildHashTable, BVH.cpp:24-38)
    for (uint32_t i = tid; i < T; i += kPhfThreads) table[i] = EMPTY;
    if (tid == 0) sFlag = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kPhfThreads)
    {
        const KEY k = keys[J.keyOff + i];
        uint32_t h = slotOf<WIDE>(A, B, k, D[bucketOf(k, mask)], T);
        uint32_t old = atomicCAS(table + h, EMPTY, J.nodeBase + i);
        if (old != EMPTY) sFlag = 1;
    }
    __syncthreads();
    if (tid == 0)
        results[blockIdx.x] = {A, B, T, attempts, sFlag ? (uint32_t)CGS_ERR_HASH_DUPLICATE : (uint32_t)CGS_OK};
}

// PerfectHashFunction::hash for a batch of keys.
template<class KEY, bool WIDE>
__global__ void kPhfHash(uint32_t A, uint32_t B, uint32_t mask, uint64_t Tm, uint32_t T, const uint32_t* __restrict__ D,
                         const KEY* __restrict__ keys, uint32_t n, uint32_t* __restrict__ out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const KEY k = keys[i];
    if constexpr (WIDE)
        out[i] = slotOf<true>(A, B, k, D[bucketOf(k, mask)], T);
    else
        out[i] = fastmod(A * k + B + D[k & mask], FastMod {Tm, T});
}

template<class KEY, bool WIDE>
__global__ void kPhfFill(uint32_t A, uint32_t B, uint32_t mask, uint64_t Tm, uint32_t T, const uint32_t* __restrict__ D,
                         const KEY* __restrict__ keys, uint32_t n, uint32_t base, uint32_t* __restrict__ table)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const KEY k = keys[i];
    if constexpr (WIDE)
        table[slotOf<true>(A, B, k, D[bucketOf(k, mask)], T)] = base + i;
    else
        table[fastmod(A * k + B + D[k & mask], FastMod {Tm, T})] = base + i;
}

__global__ void kFill(uint32_t* __restrict__ p, uint32_t v, size_t n)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation build (cgs_phf_build_static): one array per
// device for the life of the process, kStaticPhfWords words (17.3 MiB, room
// for one build of up to kStaticPhfMaxKeys keys).  Host code takes its
// address once per device with hipGetSymbolAddress and passes that address
// to the kernels; every build on a device shares the array, so builds that
// overlap on one device race.  kPhfBuild initialises every word it reads.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint32_t scratch[kStaticPhfWords];

constexpr int kStaticMaxDevices = 64;
static std::mutex g_staticMu;
static uint32_t* g_staticPtr[kStaticMaxDevices] = {};

hipError_t phfStaticInit(int device)
{
    if (device < 0 || device >= kStaticMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard lk(g_staticMu);
    if (g_staticPtr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_staticPtr[device] = (uint32_t*)p;
    return hipSuccess;
}

uint32_t* phfStaticScratch(int device)
{
    if (device < 0 || device >= kStaticMaxDevices) return nullptr;
    std::lock_guard lk(g_staticMu);
    return g_staticPtr[device];
}

hipError_t launchPhfBuild(const PhfJob* jobs, uint32_t njobs, const uint32_t* keys, const uint32_t* ab, uint32_t* scratch,
                          uint32_t* displ, uint32_t* tables, PhfResult* results, hipStream_t s, bool wide)
{
    if (njobs == 0) return hipSuccess;
    if (wide)
        hipLaunchKernelGGL((kPhfBuild<uint32_t, true>), dim3(njobs), dim3(kPhfThreads), 0, s, jobs, keys, ab, scratch,
                           displ, tables, results);
    else
        hipLaunchKernelGGL((kPhfBuild<uint32_t, false>), dim3(njobs), dim3(kPhfThreads), 0, s, jobs, keys, ab, scratch,
                           displ, tables, results);
    return hipGetLastError();
}

hipError_t launchPhfBuildWide64(const PhfJob* jobs, uint32_t njobs, const uint64_t* keys, const uint32_t* ab,
                                uint32_t* scratch, uint32_t* displ, uint32_t* tables, PhfResult* results, hipStream_t s)
{
    if (njobs == 0) return hipSuccess;
    hipLaunchKernelGGL((kPhfBuild<uint64_t, true>), dim3(njobs), dim3(kPhfThreads), 0, s, jobs, keys, ab, scratch, displ,
                       tables, results);
    return hipGetLastError();
}

template<class KEY>
static hipError_t phfHash(const cgs_phf& h, const uint32_t* D, const KEY* keys, uint32_t n, uint32_t* out, hipStream_t s,
                          bool wide)
{
    if (n == 0) return hipSuccess;
    uint64_t Tm = 0xFFFFFFFFFFFFFFFFull / h.table_size + 1;
    if (wide)
        hipLaunchKernelGGL((kPhfHash<KEY, true>), dim3((n + 255) / 256), dim3(256), 0, s, h.A, h.B, h.mask, Tm,
                           h.table_size, D, keys, n, out);
    else
        hipLaunchKernelGGL((kPhfHash<KEY, false>), dim3((n + 255) / 256), dim3(256), 0, s, h.A, h.B, h.mask, Tm,
                           h.table_size, D, keys, n, out);
    return hipGetLastError();
}

template<class KEY>
static hipError_t phfFill(const cgs_phf& h, const uint32_t* D, const KEY* keys, uint32_t n, uint32_t base,
                          uint32_t* table, hipStream_t s, bool wide)
{
    size_t T = h.table_size;
    hipLaunchKernelGGL(kFill, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s, table, EMPTY, T);
    if (n == 0) return hipGetLastError();
    uint64_t Tm = 0xFFFFFFFFFFFFFFFFull / h.table_size + 1;
    if (wide)
        hipLaunchKernelGGL((kPhfFill<KEY, true>), dim3((n + 255) / 256), dim3(256), 0, s, h.A, h.B, h.mask, Tm,
                           h.table_size, D, keys, n, base, table);
    else
        hipLaunchKernelGGL((kPhfFill<KEY, false>), dim3((n + 255) / 256), dim3(256), 0, s, h.A, h.B, h.mask, Tm,
                           h.table_size, D, keys, n, base, table);
    return hipGetLastError();
}

hipError_t launchPhfHash(const cgs_phf& h, const uint32_t* D, const uint32_t* keys, uint32_t n, uint32_t* out, hipStream_t s)
{
    return phfHash(h, D, keys, n, out, s, false);
}

hipError_t launchPhfHashWide64(const cgs_phf& h, const uint32_t* D, const uint64_t* keys, uint32_t n, uint32_t* out,
                               hipStream_t s)
{
    return phfHash(h, D, keys, n, out, s, true);
}

hipError_t launchPhfFill(const cgs_phf& h, const uint32_t* D, const uint32_t* keys, uint32_t n, uint32_t base,
                         uint32_t* table, hipStream_t s)
{
    return phfFill(h, D, keys, n, base, table, s, false);
}

hipError_t launchPhfFillWide64(const cgs_phf& h, const uint32_t* D, const uint64_t* keys, uint32_t n, uint32_t* table,
                               hipStream_t s)
{
    return phfFill(h, D, keys, n, 0u, table, s, true);
}

hipError_t launchFill(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kFill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

} // namespace cgs
