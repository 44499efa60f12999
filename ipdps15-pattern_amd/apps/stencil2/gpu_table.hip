// The W = 3 star table compiled for the device (the application's functor
// header with GPU_KERNEL defined), registered with the stencil runtime at
// StencilInit through pat_app_stencil_table().
#define GPU_KERNEL
#include "stencil.h"

#include "gpu/stencil_table.hpp"

static const int kUnit[] = {4, 4};
static const int kSameAs[] = {-1, -1};
PAT_STENCIL_TABLE(FGPU::stencil_function_table, kUnit, kSameAs)
