// SPDX-License-Identifier: MIT
//
// Size limits of the persistent launches, shared by the host library and the
// CPU tests (plain C++, no HIP headers).
//
#pragma once

#include <cstdint>

namespace cgs
{

// Persistent kernels run 256-thread workgroups (4 waves of 64).  Each wave
// takes `perFetch` work items at a time from a 32-bit counter with one
// atomicAdd, and stops at the first fetch past the end: the counter ends at
// most at items + waves × perFetch.  A launch is refused unless that stays
// below 2^32, so the counter never wraps and restarts a wave at item 0.
inline bool counterFits(uint64_t items, uint64_t grid, uint64_t perFetch)
{
    const uint64_t waves = grid * 4;
    return items < (1ull << 32) && perFetch < (1ull << 32) && waves < (1ull << 32) &&
           items + waves * perFetch < (1ull << 32);
}

// Work items of a render: one per pixel of every 8×8 tile (edge tiles whole).
inline uint64_t renderItems(uint32_t w, uint32_t h)
{
    return (((uint64_t)w + 7) / 8) * (((uint64_t)h + 7) / 8) * 64;
}

} // namespace cgs
