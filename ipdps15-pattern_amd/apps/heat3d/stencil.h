// heat3D functor table in the stencil API's shape (stencil_fp with
// parameter, heat_3D/stencil.h:32), compiled for the host (CPU_KERNEL,
// namespace FCPU) or for the device (GPU_KERNEL, namespace FGPU; see
// gpu_table.hip):
//   0  heat      the workload's rule, FP32: u' = u + alpha*(sum of the 6 face
//                neighbours - 6u), alpha = *(float *)parameter, the sum in
//                the order x-, x+, y-, y+, z-, z+; the library's tuned engine
//                computes the same bits (PAT_STENCIL_HEAT);
//   1  heat_ref  heat_3D's own functor (heat_3D/stencil.h:94-117), FP64:
//                each axis adds c*(its two neighbours + c*u) with c = 0.234,
//                in the order y, x, z; the parameter is unused.
#ifndef STENCIL
#define STENCIL
#define W 1

#include "pattern_api.h"

#if defined(GPU_KERNEL)
#define DEVICE __device__
#define CONSTANT __device__ constexpr
namespace FGPU {
#elif defined(CPU_KERNEL)
#define DEVICE inline
#define CONSTANT static constexpr
namespace FCPU {
#endif

typedef void (*stencil_fp)(void *input, void *output, int *offset, int *size, void *parameter);

DEVICE void heat(void *input, void *output, int *offset, int *size, void *parameter)
{
    const int k = offset[0], j = offset[1], i = offset[2];
    const float alpha = *(float *)parameter;
    const float c = GET_FLOAT3(input, i, j, k);
    float s = GET_FLOAT3(input, i, j, k - 1) + GET_FLOAT3(input, i, j, k + 1);
    s = s + GET_FLOAT3(input, i, j - 1, k);
    s = s + GET_FLOAT3(input, i, j + 1, k);
    s = s + GET_FLOAT3(input, i - 1, j, k);
    s = s + GET_FLOAT3(input, i + 1, j, k);
    GET_FLOAT3(output, i, j, k) = c + alpha * (s - 6.0f * c);
}

DEVICE void heat_ref(void *input, void *output, int *offset, int *size, void *parameter)
{
    (void)parameter;
    const int k = offset[0], j = offset[1], i = offset[2];
    const double c = 0.234, u = GET_DOUBLE3(input, i, j, k);
    const double ty = c * (GET_DOUBLE3(input, i, j - 1, k) + GET_DOUBLE3(input, i, j + 1, k) + u * c);
    const double tx = c * (GET_DOUBLE3(input, i, j, k - 1) + GET_DOUBLE3(input, i, j, k + 1) + u * c);
    const double tz = c * (GET_DOUBLE3(input, i - 1, j, k) + GET_DOUBLE3(input, i + 1, j, k) + u * c);
    GET_DOUBLE3(output, i, j, k) = ty + tx + tz;
}

CONSTANT stencil_fp stencil_function_table[] = {heat, heat_ref};

}  // namespace FCPU / FGPU
#endif
