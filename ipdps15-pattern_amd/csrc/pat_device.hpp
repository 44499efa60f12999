// Device-side helpers shared by the engines (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pat {

constexpr int kWave = 64;
constexpr int kXcds = 8;

// Maps a hardware workgroup id to a work index so that each XCD (the
// dispatcher deals workgroups round-robin over the 8 XCDs) gets a contiguous
// range of work, keeping neighbouring work items' shared data in one L2.
__device__ __forceinline__ int xcd_swizzle(int bid, int nblocks)
{
    const int per = nblocks / kXcds;
    if (bid >= per * kXcds) return bid;
    return (bid % kXcds) * per + bid / kXcds;
}

// This thread's wavefront within its workgroup, as a wavefront-uniform value
// (a scalar register), so that per-wavefront indices, loop bounds and
// pointers derived from it stay on the scalar unit.
__device__ __forceinline__ int wave_in_block()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

// Periodic wrap of one coordinate into [0, L) after a move shorter than L.
__device__ __forceinline__ float wrap_coord(float x, float L)
{
    if (x < 0.0f) x += L;
    else if (x >= L) x -= L;
    return x;
}

// Minimum-image component (moldyn functor, kmeans/irregular.h:55-60).
__device__ __forceinline__ float min_image(float d, float L, float hL)
{
    if (d < -hL) d += L;
    if (d > hL) d -= L;
    return d;
}

}  // namespace pat
