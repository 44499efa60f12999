// The moldyn functor tables compiled for the device: the application's
// functor header with GPU_KERNEL defined, registered with the irregular
// runtime at IrregularInit through pat_app_ir_table().
#define GPU_KERNEL
#include "irregular.h"

#include "gpu/irregular_table.hpp"

static const int kSameAs[] = {PAT_IR_LJ_MINIMAGE, -1};  // map has the library kernels' bits
PAT_IR_TABLE(FGPU::map_function_table, FGPU::reduce_function_table, kSameAs)
