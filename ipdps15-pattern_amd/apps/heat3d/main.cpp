// heat3D driver over the stencil API: the calls of heat_3D/main.cpp:51-82
// (RuntimeInit, StencilCreator, CreateRuntime, StencilInit, fill grid(),
// StencilBegin, RuntimeFinalize) on the heat3D workload: an N^3 FP32
// interior (default 128) with a Dirichlet-0 halo, uniform [0,1) from seed 42,
// 100 steps of the 7-point rule with alpha = 0.1.  Prints the L-infinity norm
// of the final field and checks one step of the GPU engine against the host
// functor; then one step of heat_3D's own FP64 functor (table index 1,
// StencilCreator<double>, as heat_3D/main.cpp:59 runs it) against the host
// functor.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CPU_KERNEL
#include "include/runtime.h"
#include "lib/stencil_creator.h"
#include "lib/stencil_runtime.h"
#include "stencil.h"

int main(int argc, char *argv[])
{
    const int N = argc > 1 ? atoi(argv[1]) : 128;
    const int iters = argc > 2 ? atoi(argv[2]) : 100;
    float alpha = 0.1f;
    RuntimeInit(argc, argv);

    // seeded interior, kept for the host check
    std::vector<float> halo((size_t)(N + 2) * (N + 2) * (N + 2));
    PAT_API_CALL(pat_gen_heat3d(N, N, N, 42, halo.data()));
    auto at = [&](int x, int y, int z) -> float & { return halo[((size_t)z * (N + 2) + y) * (N + 2) + x]; };

    IndexArray global_size(N, N, N);
    IntArray proc_size(1, 1, 1);
    // one step on the GPU, checked against the host functor on every cell
    {
        StencilCreator<float> sc(3, global_size, 3, proc_size, W, 1, false, false, &alpha, sizeof(float));
        StencilRuntime *sr = sc.CreateRuntime();
        sr->StencilInit();
        float *g = (float *)sr->grid()->data();
        for (int z = 0; z < N; z++)
            for (int y = 0; y < N; y++)
                for (int x = 0; x < N; x++) g[((size_t)z * N + y) * N + x] = at(x + 1, y + 1, z + 1);
        sr->StencilBegin();
        std::vector<float> out(halo.size(), 0.0f);
        int size[3] = {N + 2, N + 2, N + 2};
        long bad = 0;
        for (int z = 1; z <= N; z++)
            for (int y = 1; y <= N; y++)
                for (int x = 1; x <= N; x++) {
                    int off[3] = {x, y, z};
                    FCPU::stencil_function_table[0](halo.data(), out.data(), off, size, &alpha);
                    if (out[((size_t)z * (N + 2) + y) * (N + 2) + x] != g[((size_t)(z - 1) * N + (y - 1)) * N + (x - 1)])
                        bad++;
                }
        printf("heat3d: one step, %ld of %ld cells differ from the host functor\n", bad, (long)N * N * N);
        sr->StencilFinalize();
        delete sr;
        if (bad) return 1;
    }
    // heat_3D's FP64 functor: one step on the GPU, every cell against the host functor
    {
        double params[5] = {0, 0, 0, 0, 0};
        StencilCreator<double> sc(3, global_size, 3, proc_size, W, 1, false, false, params, sizeof(params));
        StencilRuntime *sr = sc.CreateRuntime();
        sr->set_stencil_idx(1);
        sr->StencilInit();
        double *g = (double *)sr->grid()->data();
        std::vector<double> in(halo.size(), 0.0), out(halo.size(), 0.0);
        for (int z = 0; z < N; z++)
            for (int y = 0; y < N; y++)
                for (int x = 0; x < N; x++) {
                    g[((size_t)z * N + y) * N + x] = at(x + 1, y + 1, z + 1);
                    in[((size_t)(z + 1) * (N + 2) + y + 1) * (N + 2) + x + 1] = at(x + 1, y + 1, z + 1);
                }
        sr->StencilBegin();
        int size[3] = {N + 2, N + 2, N + 2};
        long bad = 0;
        for (int z = 1; z <= N; z++)
            for (int y = 1; y <= N; y++)
                for (int x = 1; x <= N; x++) {
                    int off[3] = {x, y, z};
                    FCPU::stencil_function_table[1](in.data(), out.data(), off, size, params);
                    if (out[((size_t)z * (N + 2) + y) * (N + 2) + x] != g[((size_t)(z - 1) * N + (y - 1)) * N + (x - 1)])
                        bad++;
                }
        printf("heat3d: heat_3D's FP64 functor, one step, %ld of %ld cells differ from the host functor\n", bad,
               (long)N * N * N);
        sr->StencilFinalize();
        delete sr;
        if (bad) return 1;
    }
    StencilCreator<float> sc(3, global_size, 3, proc_size, W, iters, false, false, &alpha, sizeof(float));
    StencilRuntime *sr = sc.CreateRuntime();
    sr->StencilInit();
    sr->set_stencil_idx(0);
    float *g = (float *)sr->grid()->data();
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) g[((size_t)z * N + y) * N + x] = at(x + 1, y + 1, z + 1);
    sr->StencilBegin();
    double linf = 0.0;
    for (size_t i = 0; i < (size_t)N * N * N; i++) linf = std::fmax(linf, std::fabs(g[i]));
    printf("heat3d: %d^3 x %d steps, Linf = %.9e\n", N, iters, linf);
    sr->StencilFinalize();
    RuntimeFinalize();
    delete sr;
    return 0;
}
