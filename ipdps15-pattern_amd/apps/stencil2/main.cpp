// W = 3 star driver over the stencil API: the calls of heat_3D/main.cpp:51-82
// (RuntimeInit, StencilCreator, CreateRuntime, StencilInit, fill grid(),
// StencilBegin, RuntimeFinalize) with the kmeans directory's stencil table
// (kmeans/stencil.h, W = 3).  An N^3 FP32 interior (default 128) with a
// 3-cell zero halo, uniform [0,1) from seed 42; one step is checked against
// the host functor on every cell, then 'iters' steps (default 10) run and
// the L-infinity norm of the field is printed.
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
    const int iters = argc > 2 ? atoi(argv[2]) : 10;
    RuntimeInit(argc, argv);

    // seeded interior (the generator's 1-cell halo is dropped)
    std::vector<float> gen((size_t)(N + 2) * (N + 2) * (N + 2));
    PAT_API_CALL(pat_gen_heat3d(N, N, N, 42, gen.data()));
    std::vector<float> interior((size_t)N * N * N);
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++)
                interior[((size_t)z * N + y) * N + x] = gen[(((size_t)z + 1) * (N + 2) + y + 1) * (N + 2) + x + 1];

    IndexArray global_size(N, N, N);
    IntArray proc_size(1, 1, 1);
    {
        StencilCreator<float> sc(3, global_size, 3, proc_size, W, 1, false, false, nullptr, 0);
        StencilRuntime *sr = sc.CreateRuntime();
        sr->StencilInit();
        sr->set_stencil_idx(0);
        float *g = (float *)sr->grid()->data();
        std::copy(interior.begin(), interior.end(), g);
        sr->StencilBegin();
        // host functor over the same halo-inclusive grid
        const int S = N + 2 * W;
        std::vector<float> in((size_t)S * S * S, 0.0f), out((size_t)S * S * S, 0.0f);
        for (int z = 0; z < N; z++)
            for (int y = 0; y < N; y++)
                for (int x = 0; x < N; x++)
                    in[(((size_t)z + W) * S + y + W) * S + x + W] = interior[((size_t)z * N + y) * N + x];
        int size[3] = {S, S, S};
        long bad = 0;
        for (int z = W; z < N + W; z++)
            for (int y = W; y < N + W; y++)
                for (int x = W; x < N + W; x++) {
                    int off[3] = {x, y, z};
                    FCPU::stencil_function_table[0](in.data(), out.data(), off, size);
                    if (out[((size_t)z * S + y) * S + x] != g[(((size_t)z - W) * N + y - W) * N + x - W]) bad++;
                }
        printf("stencil2: one step, %ld of %ld cells differ from the host functor\n", bad, (long)N * N * N);
        sr->StencilFinalize();
        delete sr;
        if (bad) return 1;
    }
    StencilCreator<float> sc(3, global_size, 3, proc_size, W, iters, false, false, nullptr, 0);
    StencilRuntime *sr = sc.CreateRuntime();
    sr->StencilInit();
    float *g = (float *)sr->grid()->data();
    std::copy(interior.begin(), interior.end(), g);
    sr->StencilBegin();
    double linf = 0.0;
    for (size_t i = 0; i < (size_t)N * N * N; i++) linf = std::fmax(linf, std::fabs(g[i]));
    printf("stencil2: %d^3 x %d steps, Linf = %.9e\n", N, iters, linf);
    sr->StencilFinalize();
    RuntimeFinalize();
    delete sr;
    return 0;
}
