// miniMD driver over the irregular-reduction API: the calls of
// miniMD/ljs.cu:55-184 (FCC atoms, binned half neighbour list flattened to
// EDGE[], IrregularRuntime with the cutforcesq parameter, IrregularStart x5)
// on M^3 FCC cells (default 20: 32,000 atoms) at density 0.8442 with
// neighbour cutoff 2.8 and force cutoff 2.5.  As in the reference's miniMD
// there are no periodic images in the force.  Every pass is checked against
// the host map1 functor.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "irregular.h"
#include "include/runtime.h"
#include "irregular/irregular_runtime.h"

int main(int argc, char *argv[])
{
    const int M = argc > 1 ? atoi(argv[1]) : 20;
    RuntimeInit(argc, argv);
    const long n = 4L * M * M * M;
    std::vector<float> x(3 * n);
    double box[3];
    PAT_API_CALL(pat_gen_fcc(M, M, M, 0.8442, 0.0, 1, x.data(), box));
    int64_t ne = 0;
    PAT_API_CALL(pat_gen_pairs(n, x.data(), box[0], 2.8, 0, 0, nullptr, &ne));
    std::vector<EDGE> edges(ne);
    PAT_API_CALL(pat_gen_pairs(n, x.data(), box[0], 2.8, 0, 0, (int32_t *)edges.data(), &ne));
    double cutforcesq = 2.5 * 2.5;

    IrregularRuntime runtime(ne, n, edges.data(), nullptr, x.data(), 0, 3 * sizeof(float), sizeof(VALUE), 3, nullptr,
                             0, &cutforcesq, sizeof(double), 1, 5);
    runtime.IrregularInit();
    std::vector<VALUE> f(n);
    FCPU::Cobject obj(n);
    for (long e = 0; e < ne; e++)
        FCPU::map_function_table[0](&obj, x.data(), nullptr, edges.data(), &cutforcesq, (unsigned)e, 0,
                                    FCPU::reduce_function_table[0]);
    for (int pass = 0; pass < 5; pass++) {
        runtime.IrregularStart();
        runtime.get_reduction_result(f.data());
        double fmax = 0, err = 0;
        for (long i = 0; i < n; i++) {
            const VALUE h = obj.used(i) ? obj.value(i) : VALUE{0, 0, 0};
            fmax = std::fmax(fmax, std::fabs(h.x));
            err = std::fmax(err, std::fmax(std::fabs(h.x - f[i].x), std::fmax(std::fabs(h.y - f[i].y),
                                                                               std::fabs(h.z - f[i].z))));
        }
        if (err > 1e-5 * fmax) {
            fprintf(stderr, "minimd: pass %d differs from the host functor (%.2e)\n", pass, err / fmax);
            return 1;
        }
    }
    printf("minimd: %ld atoms, %ld neighbour pairs, 5 passes match the host functor\n", n, (long)ne);
    RuntimeFinalize();
    return 0;
}
