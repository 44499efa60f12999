// heat3D functor table in the stencil API's shape (stencil_fp with
// parameter, heat_3D/stencil.h:32): one output cell from the 7-point star.
// The workload's rule: u' = u + alpha*(sum of 6 face neighbours - 6u),
// alpha = *(float *)parameter, sum in the order x-, x+, y-, y+, z-, z+.
// Table index 0 is the GPU engine's PAT_STENCIL_HEAT.
#ifndef STENCIL
#define STENCIL
#define W 1

#include "pattern_api.h"

namespace FCPU {

typedef void (*stencil_fp)(void *input, void *output, int *offset, int *size, void *parameter);

inline void heat(void *input, void *output, int *offset, int *size, void *parameter)
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

static const stencil_fp stencil_function_table[] = {heat};

}  // namespace FCPU
#endif
