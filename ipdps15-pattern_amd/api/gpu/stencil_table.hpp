// Device execution of an application's stencil functor table.
//
// The reference compiles each application's functor header (heat_3D/stencil.h,
// kmeans/stencil.h) with GPU_KERNEL defined and calls the functors through a
// __constant__ function-pointer table, one call per cell
// (kmeans/lib/compute_cuda.cu:30-100).  Here the application compiles the same
// header into a HIP translation unit and names its table:
//
//     #define GPU_KERNEL
//     #include "stencil.h"
//     #include "gpu/stencil_table.hpp"
//     static const int kUnit[] = {4};                  // element bytes per entry
//     static const int kSameAs[] = {PAT_STENCIL_HEAT}; // or -1
//     PAT_STENCIL_TABLE(FGPU::stencil_function_table, kUnit, kSameAs)
//
// With CONSTANT = __device__ constexpr (pattern_api.h) the table is a
// compile-time array, so each index becomes its own kernel with the functor
// called directly and inlined.  One thread per interior cell, 64 consecutive
// x cells per wavefront (coalesced rows), 64 x 4 cells per workgroup, one
// z-plane per grid row.  The library's stencil handle calls the table's
// sweep() once per step on its two halo-inclusive grids (pat_stencil_set_table).
#ifndef PAT_GPU_STENCIL_TABLE_HPP_
#define PAT_GPU_STENCIL_TABLE_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "pattern.h"

namespace pat_gpu {

struct StencilGrid {
    int size[3];       // halo-inclusive extents, dimension 0 fastest
    int lo[3], hi[3];  // interior cells [lo, hi) per dimension
};

// The reference has a 5-argument functor (with parameter,
// heat_3D/stencil.h:32) and a 4-argument one (kmeans/stencil.h:32).
template <typename F>
__device__ __forceinline__ void call_stencil(F f, void *in, void *out, int *offset, int *size, void *param)
{
    if constexpr (std::is_invocable_v<F, void *, void *, int *, int *, void *>)
        f(in, out, offset, size, param);
    else
        f(in, out, offset, size);
}

template <class Tab, int I>
__global__ __launch_bounds__(256) void stencil_sweep(const void *in, void *out, StencilGrid g, const void *param)
{
    const int x = g.lo[0] + (int)blockIdx.x * 64 + (int)threadIdx.x;
    const int y = g.lo[1] + (int)blockIdx.y * 4 + (int)threadIdx.y;
    const int z = g.lo[2] + (int)blockIdx.z;
    if (x >= g.hi[0] || y >= g.hi[1]) return;
    int offset[3] = {x, y, z};
    int size[3] = {g.size[0], g.size[1], g.size[2]};
    Tab::template call<I>(const_cast<void *>(in), out, offset, size, const_cast<void *>(param));
}

template <class Tab, int I>
int stencil_launch(const void *in, void *out, const StencilGrid &g, const void *param, hipStream_t st)
{
    if constexpr (!Tab::template valid<I>()) {
        return PAT_ERR_UNSUPPORTED;
    } else {
        const dim3 block(64, 4, 1);
        const dim3 grid((unsigned)((g.hi[0] - g.lo[0] + 63) / 64), (unsigned)((g.hi[1] - g.lo[1] + 3) / 4),
                        (unsigned)(g.hi[2] - g.lo[2]));
        hipLaunchKernelGGL((stencil_sweep<Tab, I>), grid, block, 0, st, in, out, g, param);
        return hipGetLastError() == hipSuccess ? PAT_OK : PAT_ERR_HIP;
    }
}

template <class Tab, std::size_t... I>
int stencil_dispatch(int idx, const void *in, void *out, const StencilGrid &g, const void *param, hipStream_t st,
                     std::index_sequence<I...>)
{
    int rc = PAT_ERR_UNSUPPORTED;
    ((idx == (int)I ? (void)(rc = stencil_launch<Tab, (int)I>(in, out, g, param, st)) : (void)0), ...);
    return rc;
}

// pat_stencil_table::sweep for table Tab: one sweep of functor idx over the
// interior of a grid of extents size[0..ndims-1] with halo width 'width'.
template <class Tab>
int stencil_sweep_entry(int idx, const void *in, void *out, const int32_t *size, int ndims, int width,
                        const void *param, void *stream)
{
    if (!in || !out || !size || ndims < 1 || ndims > 3 || width < 0) return PAT_ERR_INVALID;
    StencilGrid g;
    for (int d = 0; d < 3; d++) {
        const bool used = d < ndims;
        g.size[d] = used ? size[d] : 1;
        g.lo[d] = used ? width : 0;
        g.hi[d] = used ? size[d] - width : 1;
        if (g.hi[d] <= g.lo[d]) return PAT_ERR_INVALID;
    }
    if (g.hi[2] - g.lo[2] > 65535) return PAT_ERR_UNSUPPORTED;  // one grid row per plane
    return stencil_dispatch<Tab>(idx, in, out, g, param, (hipStream_t)stream,
                                 std::make_index_sequence<(std::size_t)Tab::count>{});
}

}  // namespace pat_gpu

// Defines pat_app_stencil_table() for the application's table TABLE (a
// __device__ constexpr array of stencil functors), with UNIT[i] the element
// bytes functor i reads and writes and SAME_AS[i] the library's tuned functor
// that computes the same bits (PAT_STENCIL_HEAT) or -1.
#define PAT_STENCIL_TABLE(TABLE, UNIT, SAME_AS)                                                                  \
    namespace {                                                                                                  \
    struct PatStencilTab {                                                                                       \
        static constexpr int count = (int)(sizeof(TABLE) / sizeof(TABLE[0]));                                    \
        template <int I>                                                                                         \
        static constexpr bool valid()                                                                            \
        {                                                                                                        \
            return TABLE[I] != nullptr;                                                                          \
        }                                                                                                        \
        template <int I>                                                                                         \
        __device__ static void call(void *in, void *out, int *offset, int *size, void *param)                    \
        {                                                                                                        \
            pat_gpu::call_stencil(TABLE[I], in, out, offset, size, param);                                       \
        }                                                                                                        \
    };                                                                                                           \
    static_assert(sizeof(UNIT) / sizeof(UNIT[0]) == PatStencilTab::count, "one element size per table entry");  \
    static_assert(sizeof(SAME_AS) / sizeof(SAME_AS[0]) == PatStencilTab::count, "one same-as entry per functor"); \
    }                                                                                                            \
    extern "C" const pat_stencil_table *pat_app_stencil_table(void)                                              \
    {                                                                                                            \
        static const pat_stencil_table t = {PatStencilTab::count, UNIT, SAME_AS,                                 \
                                            &pat_gpu::stencil_sweep_entry<PatStencilTab>};                       \
        return &t;                                                                                               \
    }

#endif
