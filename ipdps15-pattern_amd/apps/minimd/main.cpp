// miniMD driver over the irregular-reduction API: the calls of
// miniMD/ljs.cu:55-184 (FCC atoms, binned half neighbour list flattened to
// EDGE[], IrregularRuntime with the cutforcesq parameter, IrregularStart x5)
// on M^3 FCC cells (default 20: 32,000 atoms) at density 0.8442 with
// neighbour cutoff 2.8 and force cutoff 2.5, or the sizes of an lj.in file
// (-f path); then the thermo pair energy through RegularRuntime (map2).  As in the reference's miniMD
// there are no periodic images in the force.  Node data is double[3] per
// atom, as miniMD passes it.  The force is map 0 of the application's functor
// table, compiled for the device (gpu_table.hip); every pass is checked
// against the same functor on the host.
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
    RuntimeInit(argc, argv);
    // sizes on the command line, or from an lj.in file as miniMD's input()
    // reads it (miniMD/input.cpp:39-72): problem size, density, cutoffs
    int nx = 20, ny = 20, nz = 20;
    double rho = 0.8442, cutforce = 2.5, cutneigh = 2.8;
    if (argc > 2 && std::string(argv[1]) == "-f") {
        pat_ljin in;
        PAT_API_CALL(pat_read_ljin(argv[2], &in));
        nx = in.nx, ny = in.ny, nz = in.nz;
        rho = in.rho, cutforce = in.cutforce, cutneigh = in.cutneigh;
        printf("minimd: %s: %d x %d x %d cells, density %g, cutoffs %g / %g\n", argv[2], nx, ny, nz, rho, cutforce,
               cutneigh);
    } else if (argc > 1) {
        nx = ny = nz = atoi(argv[1]);
    }
    const long n = 4L * nx * ny * nz;
    std::vector<float> x(3 * n);
    double box[3];
    PAT_API_CALL(pat_gen_fcc(nx, ny, nz, rho, 0.0, 1, x.data(), box));
    if (!(nx == ny && ny == nz)) {
        fprintf(stderr, "minimd: the pair generator takes a cubic box (%d x %d x %d cells)\n", nx, ny, nz);
        return 1;
    }
    int64_t ne = 0;
    PAT_API_CALL(pat_gen_pairs(n, x.data(), box[0], cutneigh, 0, 0, nullptr, &ne));
    std::vector<EDGE> edges(ne);
    PAT_API_CALL(pat_gen_pairs(n, x.data(), box[0], cutneigh, 0, 0, (int32_t *)edges.data(), &ne));
    double cutforcesq = cutforce * cutforce;

    std::vector<double> xd(x.begin(), x.end());
    IrregularRuntime runtime(ne, n, edges.data(), nullptr, xd.data(), 0, 3 * sizeof(double), sizeof(VALUE), 3,
                             nullptr, 0, &cutforcesq, sizeof(double), 1, 5);
    runtime.IrregularInit();
    std::vector<VALUE> f(n);
    std::vector<int> part(n, 0);
    Cobject obj(n);
    for (long e = 0; e < ne; e++)
        FCPU::map_function_table[0](&obj, xd.data(), part.data(), edges.data(), &cutforcesq, (unsigned)e, 0,
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

    // thermo pair energy through the generalized-reduction API (map2,
    // miniMD/regular.h:139-181) over the same edges, with the parameter
    // laid out as map2 reads it ({cutforcesq, x[3N]} in FP64, the nodes1
    // array of miniMD/ljs.cu:150-155); checked against a host loop
    std::vector<double> nodes1(1 + 3 * n);
    nodes1[0] = cutforcesq;
    for (long i = 0; i < 3 * n; i++) nodes1[1 + i] = x[i];
    std::vector<Offset> offsets(ne);
    for (long e = 0; e < ne; e++) offsets[e] = Offset{(size_t)e * sizeof(EDGE), sizeof(EDGE)};
    RegularRuntime gr(1, edges.data(), sizeof(EDGE) * ne, offsets.data(), ne, nodes1.data(),
                      (int)(sizeof(double) * nodes1.size()));
    gr.RegularInit();
    gr.set_map_idx(2);
    gr.set_reduce_idx(2);
    gr.RegularStart();
    struct output o = gr.RegularGetOutput();
    const double epair = *(const double *)o.output_vals;
    double host = 0.0;
    const double *v = nodes1.data() + 1;
    for (long e = 0; e < ne; e++) {
        const long i = ((const int *)edges.data())[2 * e], j = ((const int *)edges.data())[2 * e + 1];
        const double dx = v[3 * i] - v[3 * j], dy = v[3 * i + 1] - v[3 * j + 1], dz = v[3 * i + 2] - v[3 * j + 2];
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < cutforcesq) {
            const double sr2 = 1.0 / rsq, sr6 = sr2 * sr2 * sr2;
            host += sr6 * (sr6 - 1.0);
        }
    }
    printf("minimd: pair energy sum %.12e (host %.12e)\n", epair, host);
    if (std::fabs(epair - host) > 1e-10 * std::fabs(host)) return 1;
    RuntimeFinalize();
    return 0;
}
