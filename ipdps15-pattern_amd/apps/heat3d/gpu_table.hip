// The heat3D functor table compiled for the device: the application's
// functor header with GPU_KERNEL defined (the reference builds its GPU
// functors the same way, heat_3D/lib/compute_cuda.cu:1-6), registered with
// the stencil runtime at StencilInit through pat_app_stencil_table().
#define GPU_KERNEL
#include "stencil.h"

#include "gpu/stencil_table.hpp"

static const int kUnit[] = {4, 8};                       // heat: float, heat_ref: double
static const int kSameAs[] = {PAT_STENCIL_HEAT, -1};     // heat has the tuned engine's bits
PAT_STENCIL_TABLE(FGPU::stencil_function_table, kUnit, kSameAs)
