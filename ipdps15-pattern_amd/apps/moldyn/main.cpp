// moldyn driver over the irregular-reduction API: the step loop of
// kmeans/moldyn.cu:331-403 (host coordinate update, reset_node_data,
// IrregularStart, get_reduction_result, host velocity update) as
// velocity-Verlet, on an FCC lattice of M^3 cells (default 16: 16,384
// molecules) at density 0.8442, jitter 0.05, cutoff 2.5, dt 0.005, 50 steps;
// or, with -f <parameter file> [steps], a reference dataset (.mesh + .xyz).
// Node data is double3 as moldyn's point_data.  The force is map 0 of the
// application's functor table (gpu_table.hip), which the library's LJ
// kernels compute bit for bit (same_as); the first pass is checked against
// the same functor on the host (within 1e-5: the library kernel adds each
// node's forces in neighbour order, the host in edge order), and map 1, the
// reference functor's FP64 arithmetic, runs on the GPU through the table's
// own kernel and is checked against the host likewise.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#define CPU_KERNEL
#include "irregular.h"
#include "include/runtime.h"
#include "irregular/irregular_runtime.h"
#include "regular/regular_runtime.h"

int main(int argc, char *argv[])
{
    const double dt = 0.005, hdt = 0.5 * dt;
    RuntimeInit(argc, argv);

    long n = 0;
    int steps = 50;
    double cutoff = 2.5, box[3];
    std::vector<float> pos3;
    std::vector<point_data> x;
    std::vector<EDGE> edges;
    int64_t ne = 0;
    if (argc > 2 && std::string(argv[1]) == "-f") {
        // a reference dataset: the parameter file moldyn reads from stdin
        // (kmeans/moldyn.cu:193-201) naming its .mesh graph and .xyz
        // coordinates (paths relative to the parameter file); the box side
        // from the molecule count at moldyn's density, as InitSettings
        // computes it (kmeans/moldyn.cu:493-499, DENSITY kmeans/moldyn.h:81)
        pat_mold m;
        PAT_API_CALL(pat_read_mold(argv[2], &m));
        std::string dir(argv[2]);
        dir = dir.find('/') == std::string::npos ? std::string(".") : dir.substr(0, dir.rfind('/'));
        const std::string mesh = m.mesh[0] == '/' ? m.mesh : dir + "/" + m.mesh;
        const std::string xyz = m.xyz[0] == '/' ? m.xyz : dir + "/" + m.xyz;
        int64_t nn = 0;
        PAT_API_CALL(pat_read_mesh(mesh.c_str(), 1, &nn, &ne, nullptr, 0));
        edges.resize(ne);
        PAT_API_CALL(pat_read_mesh(mesh.c_str(), 1, &nn, &ne, (int32_t *)edges.data(), ne));
        n = (long)nn;
        std::vector<double> xyzv(3 * n);
        PAT_API_CALL(pat_read_xyz(xyz.c_str(), n, xyzv.data()));
        steps = argc > 3 ? atoi(argv[3]) : m.steps;
        cutoff = m.cutoff;
        box[0] = box[1] = box[2] = std::pow((double)n / 0.83134, 0.3333333);
        pos3.resize(3 * n);
        x.resize(n);
        for (long i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) {
                x[i].points[c] = xyzv[3 * i + c];
                pos3[3 * i + c] = (float)xyzv[3 * i + c];
            }
        printf("moldyn: dataset %s: %ld molecules, %lld pairs, box %.4f, cutoff %.2f\n", argv[2], n, (long long)ne,
               box[0], cutoff);
    } else {
        const int M = argc > 1 ? atoi(argv[1]) : 16;
        steps = argc > 2 ? atoi(argv[2]) : 50;
        n = 4L * M * M * M;
        pos3.resize(3 * n);
        PAT_API_CALL(pat_gen_fcc(M, M, M, 0.8442, 0.05, 42, pos3.data(), box));
        PAT_API_CALL(pat_gen_pairs(n, pos3.data(), box[0], cutoff, 1, 43, nullptr, &ne));
        edges.resize(ne);
        PAT_API_CALL(pat_gen_pairs(n, pos3.data(), box[0], cutoff, 1, 43, (int32_t *)edges.data(), &ne));
        x.resize(n);
        for (long i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) x[i].points[c] = pos3[3 * i + c];
    }
    std::vector<double> v(3 * n, 0.0);
    std::vector<VALUE> f(n);
    struct para pa = {box[0], cutoff};

    IrregularRuntime runtime(ne, n, edges.data(), nullptr, x.data(), 0, sizeof(point_data), sizeof(VALUE), 3, nullptr,
                             0, &pa, sizeof(pa), 1, steps);
    runtime.set_map_idx(0);
    runtime.set_reduce_idx(0);
    runtime.IrregularInit();
    runtime.IrregularStart();
    runtime.get_reduction_result(f.data());

    // host functor check of the first force pass: every node owned by
    // partition 0 (the serial scatter over the edge list)
    {
        std::vector<int> part(n, 0);
        Cobject obj(n);
        for (long e = 0; e < ne; e++)
            FCPU::map_function_table[0](&obj, x.data(), part.data(), edges.data(), &pa, (unsigned)e, 0,
                                        FCPU::reduce_function_table[0]);
        double fmax = 0, err = 0;
        for (long i = 0; i < n; i++) {
            const VALUE h = obj.used(i) ? obj.value(i) : VALUE{0, 0, 0};
            fmax = std::fmax(fmax, std::fabs(h.x));
            err = std::fmax(err, std::fabs(h.x - f[i].x));
            err = std::fmax(err, std::fmax(std::fabs(h.y - f[i].y), std::fabs(h.z - f[i].z)));
        }
        printf("moldyn: %ld molecules, %ld pairs; first pass max |dF| / max |F| = %.2e\n", n, (long)ne, err / fmax);
        if (err > 1e-5 * fmax) return 1;
        // the reference functor's FP64 arithmetic (map 1, without the 48 eps
        // prefactor) on the device, against the same functor on the host
        runtime.set_map_idx(1);
        runtime.IrregularStart();
        std::vector<VALUE> fr(n);
        runtime.get_reduction_result(fr.data());
        Cobject ref(n);
        for (long e = 0; e < ne; e++)
            FCPU::map_function_table[1](&ref, x.data(), part.data(), edges.data(), &pa, (unsigned)e, 0,
                                        FCPU::reduce_function_table[0]);
        long bad = 0;
        for (long i = 0; i < n; i++) {
            const VALUE h = ref.used(i) ? ref.value(i) : VALUE{0, 0, 0};
            if (std::fabs(48.0 * h.x - f[i].x) > 1e-5 * fmax || std::fabs(48.0 * h.y - f[i].y) > 1e-5 * fmax ||
                std::fabs(48.0 * h.z - f[i].z) > 1e-5 * fmax)
                bad++;
        }
        double rerr = 0;
        for (long i = 0; i < n; i++) {
            const VALUE h = ref.used(i) ? ref.value(i) : VALUE{0, 0, 0};
            rerr = std::fmax(rerr, std::fmax(std::fabs(h.x - fr[i].x),
                                             std::fmax(std::fabs(h.y - fr[i].y), std::fabs(h.z - fr[i].z))));
        }
        printf("moldyn: reference functor (FP64, no 48): device vs host max |dF| = %.2e; 48 x it vs the FP32 force: "
               "%ld nodes beyond 1e-5\n", rerr, bad);
        if (bad || rerr > 1e-5 * fmax / 48.0) return 1;
        runtime.set_map_idx(0);
    }

    const double L = box[0];
    for (int step = 0; step < steps; step++) {
        for (long i = 0; i < n; i++) {
            const float fi[3] = {f[i].x, f[i].y, f[i].z};
            for (int c = 0; c < 3; c++) {
                v[3 * i + c] += hdt * fi[c];
                double p = x[i].points[c] + dt * v[3 * i + c];
                if (p < 0) p += L;
                else if (p >= L) p -= L;
                x[i].points[c] = p;
            }
        }
        runtime.reset_node_data(x.data());
        runtime.IrregularStart();
        runtime.get_reduction_result(f.data());
        for (long i = 0; i < n; i++) {
            v[3 * i] += hdt * f[i].x;
            v[3 * i + 1] += hdt * f[i].y;
            v[3 * i + 2] += hdt * f[i].z;
        }
        runtime.IrregularBarrier();
    }
    double ekin = 0, px = 0;
    for (long i = 0; i < n; i++) {
        ekin += 0.5 * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        px += v[3 * i];
    }
    printf("moldyn: %d steps, kinetic energy %.6e, total x momentum %.2e\n", steps, ekin, px);

    // velocity statistics through the generalized-reduction API, as moldyn
    // runs them (kmeans/moldyn.cu:440-470): map 0 emits {|v|, |v| > vaverh}
    // per molecule, reduce 0 adds them; checked against the host loop
    double vsum = 0.0;
    for (long i = 0; i < n; i++)
        vsum += std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    float vaverh = (float)(vsum / n);
    std::vector<Offset> offsets(n);
    for (long i = 0; i < n; i++) offsets[i] = Offset{(size_t)i * 3 * sizeof(double), 3 * sizeof(double)};
    RegularRuntime rr(1, v.data(), sizeof(double) * 3 * n, offsets.data(), n, &vaverh, sizeof(float));
    rr.RegularInit();
    rr.set_map_idx(0);
    rr.set_reduce_idx(0);
    rr.RegularStart();
    struct output o = rr.RegularGetOutput();
    const float *val = (const float *)o.output_vals;
    double hs = 0.0, hc = 0.0;
    for (long i = 0; i < n; i++) {
        const float sq = (float)std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        if (sq > vaverh) hc += 1.0;
        hs += sq;
    }
    printf("moldyn: velocity statistics sum |v| %.6e (host %.6e), above vaverh %.0f (host %.0f)\n", val[0], hs,
           val[1], hc);
    if (std::fabs(val[0] - hs) > 1e-6 * hs || val[1] != (float)hc) return 1;
    RuntimeFinalize();
    return 0;
}
