// The generalized-reduction functor tables compiled for the device: the
// application's functor header with GPU_KERNEL defined, registered with the
// regular runtime at RegularInit through pat_app_gr_table().
#define GPU_KERNEL
#include "regular.h"

#include "gpu/regular_table.hpp"

PAT_GR_TABLE(FFGPU::map_function_table, FFGPU::reduce_function_table)
