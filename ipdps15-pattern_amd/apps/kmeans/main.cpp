// kmeans driver over the generalized-reduction API: the calls of
// kmeans/km.cu:20-87 (points + per-unit Offset array, RegularRuntime,
// RegularInit, set_map_idx(1), set_reduce_idx(1), RegularStart) followed by
// the Lloyd update km.cu leaves out, on a Gaussian mixture (default
// N = 2^20, D = 32, K = 256, 20 iterations).  The first pass's keys are
// checked against the host map1 functor on a sample of points.  A sixth
// argument gap > 0 stores the points gap floats apart (unit i at
// i * (D + gap) floats), so the Offset array describes units that are not
// packed back to back and the runtime gathers them.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "include/runtime.h"
#include "regular.h"
#include "regular/regular_runtime.h"

int main(int argc, char *argv[])
{
    const long N = argc > 1 ? atol(argv[1]) : (1L << 20);
    const int D = argc > 2 ? atoi(argv[2]) : 32;
    const int K = argc > 3 ? atoi(argv[3]) : 256;
    const int iters = argc > 4 ? atoi(argv[4]) : 20;
    const int gap = argc > 5 ? atoi(argv[5]) : 0;
    RuntimeInit(argc, argv);

    std::vector<float> packed((size_t)N * D);
    PAT_API_CALL(pat_gen_gmm(N, D, 256, 1.0, -10.0, 10.0, 2024, packed.data()));
    const size_t stride = (size_t)D + (gap > 0 ? gap : 0);  // floats per unit slot
    std::vector<float> points(gap > 0 ? (size_t)N * stride : 0, -1.0e30f);
    if (gap > 0)
        for (long i = 0; i < N; i++) std::copy_n(&packed[(size_t)i * D], D, &points[(size_t)i * stride]);
    else
        points.swap(packed);
    std::vector<Offset> offsets(N);
    for (long i = 0; i < N; i++) {
        offsets[i].offset = (size_t)i * stride * sizeof(float);
        offsets[i].size = D * sizeof(float);
    }
    std::vector<float> centers(K * (size_t)D);  // first K points
    for (int k = 0; k < K; k++) std::copy_n(&points[(size_t)k * stride], D, &centers[(size_t)k * D]);

    RegularRuntime rr(1, points.data(), points.size() * sizeof(float), offsets.data(), N, centers.data(),
                      K * D * (int)sizeof(float));
    rr.RegularInit();
    rr.set_map_idx(1);
    rr.set_reduce_idx(1);

    double sse = 0.0;
    for (int it = 0; it < iters; it++) {
        rr.RegularStart();
        struct output out = rr.RegularGetOutput();
        if (it == 0) {
            // host functor check of the first pass on every 97th point
            std::vector<int> keys(N);
            rr.get_keys(keys.data());
            FFCPU::KmParam p{centers.data(), D, K};
            long bad = 0, checked = 0;
            for (long i = 0; i < N; i += 97) {
                CO obj(K, (D + 2) * sizeof(float));
                FFCPU::map_function_table[1](&obj, points.data(), offsets[i].offset, &p, 0,
                                             FFCPU::reduce_function_table[1]);
                int hk = -1;
                for (int k = 0; k < K; k++)
                    if (obj.used(k)) hk = k;
                bad += hk != keys[i];
                checked++;
            }
            printf("kmeans: first pass, %ld of %ld sampled keys differ from the host functor\n", bad, checked);
            if (bad) return 1;
        }
        // Lloyd update from the reduction object: c = sum / count
        const float *vals = (const float *)out.output_vals;
        const unsigned int *keys = (const unsigned int *)out.output_keys;
        long total = 0;
        sse = 0.0;
        for (unsigned int i = 0; i < out.key_num; i++) {
            const float *v = vals + (size_t)i * (D + 2);
            const unsigned int k = keys[i];
            for (int d = 0; d < D; d++) centers[(size_t)k * D + d] = v[d] / v[D];
            total += (long)v[D];
            sse += v[D + 1];
        }
        if (total != N) {
            fprintf(stderr, "kmeans: counts sum to %ld, expected %ld\n", total, N);
            return 1;
        }
    }
    printf("kmeans: N=%ld D=%d K=%d, %d iterations (units %s), sum of min distances %.6e\n", N, D, K, iters,
           gap > 0 ? "gathered" : "packed", sse);
    RuntimeFinalize();
    return 0;
}
