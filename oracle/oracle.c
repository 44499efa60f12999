/*
 * oracle.c -- CPU restatement of the IPDPS'15 pattern framework's three
 * per-partition compute engines (stencil sweep, irregular edge reduction,
 * generalized reduction for kmeans).
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  It is the oracle the HIP engines are
 * compared against.  Only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark may load it; the product library never links or calls
 * it, and nothing here is a fallback for the GPU path.
 *
 * Every function cites the reference file:line whose arithmetic it restates.
 * Build rules: -O2 -ffp-contract=off (no fused multiply-add, so each float
 * expression rounds exactly as written, like the reference's x86 CPU path).
 *
 * Pinning (what checks each function beyond this file; see tests/):
 *   heat3d_*      closed-form sine-eigenmode decay, linearity, axis symmetry,
 *                 brute-force FP64 on tiny grids (test_oracle_stencil.py)
 *   lj_*          two-particle closed form, perfect-FCC zero force, Newton's
 *                 third law, edge-order invariance, and miniMD's own
 *                 Force::compute built from the reference sources
 *                 (oracle/_ref, test_oracle_lj.py)
 *   neighbor_*    miniMD's own Neighbor::build (oracle/_ref) and O(N^2) brute
 *                 force (test_oracle_lj.py)
 *   kmeans_*      brute-force FP64 argmin, hand-built ties, label recovery,
 *                 SSE monotonicity (test_oracle_kmeans.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define IDX3(x, y, z, sx, sy) ((int64_t)(x) + (int64_t)(sx) * ((int64_t)(y) + (int64_t)(sy) * (int64_t)(z)))

/* ======================================================================
 * Stencil engine: heat3D
 * ====================================================================== */

/*
 * One explicit heat step over the interior of a halo-inclusive grid.
 * Layout follows GET_FLOAT3 (heat_3D/lib/macro.h:70): element (i=z, j=y, k=x)
 * lives at in + size[0]*j + i*size[0]*size[1] + k, size = halo-inclusive
 * extents, so x is the fastest index.  The functor shape (one output cell from
 * the 7-point star) is heat_3D/stencil.h:94-117; the update rule is the heat3D
 * workload's own: u' = u + alpha*(s - 6u) with
 * s = ((((x- + x+) + y-) + y+) + z-) + z+, evaluated in that order.
 * The halo is Dirichlet: it is copied through unchanged (stays 0).
 */
void orc_heat3d_step_f32(const float *in, float *out, int nx, int ny, int nz, float alpha)
{
    const int sx = nx + 2, sy = ny + 2;
    const int64_t sxy = (int64_t)sx * sy;
    const int64_t total = sxy * (nz + 2);
    /* halo stays as in the input (Dirichlet) */
    memcpy(out, in, sizeof(float) * total);
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                int64_t id = IDX3(x, y, z, sx, sy);
                float c = in[id];
                float s = in[id - 1] + in[id + 1];
                s = s + in[id - sx];
                s = s + in[id + sx];
                s = s + in[id - sxy];
                s = s + in[id + sxy];
                out[id] = c + alpha * (s - 6.0f * c);
            }
}

void orc_heat3d_step_f64(const double *in, double *out, int nx, int ny, int nz, double alpha)
{
    const int sx = nx + 2, sy = ny + 2;
    const int64_t sxy = (int64_t)sx * sy;
    const int64_t total = sxy * (nz + 2);
    memcpy(out, in, sizeof(double) * total);
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                int64_t id = IDX3(x, y, z, sx, sy);
                double c = in[id];
                double s = in[id - 1] + in[id + 1];
                s = s + in[id - sx];
                s = s + in[id + sx];
                s = s + in[id - sxy];
                s = s + in[id + sxy];
                out[id] = c + alpha * (s - 6.0 * c);
            }
}

/*
 * nsteps of ping-pong (the double buffer of kmeans/lib/grid.h:7-73 with the
 * Swap() that the reference never calls, grid.h:61-65).  Returns 0 when the
 * result is in a, 1 when it is in b.
 */
int orc_heat3d_run_f32(float *a, float *b, int nx, int ny, int nz, int nsteps, float alpha)
{
    float *src = a, *dst = b;
    for (int s = 0; s < nsteps; s++) {
        orc_heat3d_step_f32(src, dst, nx, ny, nz, alpha);
        float *t = src; src = dst; dst = t;
    }
    return src == a ? 0 : 1;
}

int orc_heat3d_run_f64(double *a, double *b, int nx, int ny, int nz, int nsteps, double alpha)
{
    double *src = a, *dst = b;
    for (int s = 0; s < nsteps; s++) {
        orc_heat3d_step_f64(src, dst, nx, ny, nz, alpha);
        double *t = src; src = dst; dst = t;
    }
    return src == a ? 0 : 1;
}

/*
 * The W = 3 star of the reference's kmeans directory (stencil2,
 * kmeans/stencil.h:93-163, W from :3): over a halo-inclusive FP32 grid of
 * halo width w (interior nx x ny x nz, x fastest), each interior cell
 * becomes the sum of the cells 1..w-1 away along each axis, accumulated
 * from 0 in the order x+, x-, y+, y-, z+, z- (each side by increasing
 * distance), divided by 36.  The halo is copied through unchanged.
 */
void orc_star_step_f32(const float *in, float *out, int nx, int ny, int nz, int w)
{
    const int64_t sx = nx + 2 * w, sy = ny + 2 * w;
    const int64_t sxy = sx * sy;
    memcpy(out, in, sizeof(float) * (size_t)(sxy * (nz + 2 * w)));
    for (int z = w; z < nz + w; z++)
        for (int y = w; y < ny + w; y++)
            for (int x = w; x < nx + w; x++) {
                const int64_t id = (int64_t)z * sxy + (int64_t)y * sx + x;
                float t = 0.0f;
                for (int m = 1; m < w; m++) t += in[id + m];
                for (int m = 1; m < w; m++) t += in[id - m];
                for (int m = 1; m < w; m++) t += in[id + m * sx];
                for (int m = 1; m < w; m++) t += in[id - m * sx];
                for (int m = 1; m < w; m++) t += in[id + m * sxy];
                for (int m = 1; m < w; m++) t += in[id - m * sxy];
                out[id] = t / 36.0f;
            }
}

int orc_star_run_f32(float *a, float *b, int nx, int ny, int nz, int w, int nsteps)
{
    float *src = a, *dst = b;
    for (int s = 0; s < nsteps; s++) {
        orc_star_step_f32(src, dst, nx, ny, nz, w);
        float *t = src; src = dst; dst = t;
    }
    return src == a ? 0 : 1;
}

/*
 * heat_3D's own functor (heat_3D/stencil.h:94-117), FP64, halo width 1: each
 * axis contributes c * (its two neighbours + c * u) with c = 0.234, and the
 * contributions are added in the order y, x, z.
 */
void orc_heat_ref_step_f64(const double *in, double *out, int nx, int ny, int nz)
{
    const int64_t sx = nx + 2, sy = ny + 2;
    const int64_t sxy = sx * sy;
    const double c = 0.234;
    memcpy(out, in, sizeof(double) * (size_t)(sxy * (nz + 2)));
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                const int64_t id = (int64_t)z * sxy + (int64_t)y * sx + x;
                const double u = in[id];
                const double ty = c * (in[id - sx] + in[id + sx] + u * c);
                const double tx = c * (in[id - 1] + in[id + 1] + u * c);
                const double tz = c * (in[id - sxy] + in[id + sxy] + u * c);
                out[id] = ty + tx + tz;
            }
}

int orc_heat_ref_run_f64(double *a, double *b, int nx, int ny, int nz, int nsteps)
{
    double *src = a, *dst = b;
    for (int s = 0; s < nsteps; s++) {
        orc_heat_ref_step_f64(src, dst, nx, ny, nz);
        double *t = src; src = dst; dst = t;
    }
    return src == a ? 0 : 1;
}

/*
 * Threaded CPU path in the shape of the reference's CPU stencil engine:
 * the interior is cut into 16^3 tiles (CPU tile size, kmeans/lib/common.h:17;
 * tiling in kmeans/lib/data_util.cpp:224-307) and the tile list is split
 * statically over the threads (compute_tiles, kmeans/lib/compute_cpu.cpp:
 * 69-126), each tile swept cell by cell (process_a_tile, :35-67, with the
 * inner offsets reset per row as the reference intends).  Same arithmetic as
 * orc_heat3d_step_f32, so results are bit-identical to the serial path.
 */
int orc_heat3d_run_f32_tiled(float *a, float *b, int nx, int ny, int nz, int nsteps, float alpha, int nthreads)
{
    const int T = 16;
    const int sx = nx + 2, sy = ny + 2;
    const int64_t sxy = (int64_t)sx * sy;
    const int64_t total = sxy * (nz + 2);
    const int tx = (nx + T - 1) / T, ty = (ny + T - 1) / T, tz = (nz + T - 1) / T;
    const int ntiles = tx * ty * tz;
    float *src = a, *dst = b;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    /* halo planes of both buffers are the same Dirichlet zeros */
    memcpy(dst, src, sizeof(float) * total);
    for (int s = 0; s < nsteps; s++) {
        #pragma omp parallel num_threads(nthreads)
        {
            int tid = omp_get_thread_num(), nt = omp_get_num_threads();
            int per = (ntiles + nt - 1) / nt;
            int t0 = tid * per, t1 = t0 + per < ntiles ? t0 + per : ntiles;
            for (int t = t0; t < t1; t++) {
                int bx = t % tx, by = (t / tx) % ty, bz = t / (tx * ty);
                int x0 = 1 + bx * T, y0 = 1 + by * T, z0 = 1 + bz * T;
                int x1 = x0 + T <= nx + 1 ? x0 + T : nx + 1;
                int y1 = y0 + T <= ny + 1 ? y0 + T : ny + 1;
                int z1 = z0 + T <= nz + 1 ? z0 + T : nz + 1;
                for (int z = z0; z < z1; z++)
                    for (int y = y0; y < y1; y++)
                        for (int x = x0; x < x1; x++) {
                            int64_t id = IDX3(x, y, z, sx, sy);
                            float c = src[id];
                            float sm = src[id - 1] + src[id + 1];
                            sm = sm + src[id - sx];
                            sm = sm + src[id + sx];
                            sm = sm + src[id - sxy];
                            sm = sm + src[id + sxy];
                            dst[id] = c + alpha * (sm - 6.0f * c);
                        }
            }
        }
        float *t = src; src = dst; dst = t;
    }
    return src == a ? 0 : 1;
}

/* ======================================================================
 * Irregular engine: Lennard-Jones edge reduction
 * ====================================================================== */

/*
 * Pair force of the moldyn functor (kmeans/irregular.h:51-72): minimum-image
 * displacement d = x_i - x_j with box side L (:55-60), rd = |d|^2, and for
 * rd < rc^2: rrd = 1/rd, rrd2 = rrd^2, rrd4 = rrd2^2, rrd6 = rrd2*rrd4,
 * rrd7 = rrd6*rrd, r148 = rrd7 - 0.5*rrd4.  The reference omits the 48*eps
 * prefactor (it folds it into the integrator constants, kmeans/moldyn.cu:
 * 396-398); here the force is written out in full, F = 48*r148*d for
 * eps = sigma = 1.  pbc = 0 drops the minimum image (miniMD map1,
 * miniMD/irregular.h:127-139, which has no periodic images).
 * Returns 1 and the force-on-i components when the pair is inside the cutoff.
 */
static inline int pair_force_f32(const float *pi, const float *pj, float L, float hL, int pbc, float rc2,
                                 float *fx, float *fy, float *fz, float *phi)
{
    float dx = pi[0] - pj[0];
    float dy = pi[1] - pj[1];
    float dz = pi[2] - pj[2];
    if (pbc) {
        if (dx < -hL) dx += L;
        if (dy < -hL) dy += L;
        if (dz < -hL) dz += L;
        if (dx > hL) dx -= L;
        if (dy > hL) dy -= L;
        if (dz > hL) dz -= L;
    }
    float rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    if (!(rd < rc2)) return 0;
    float rrd = 1.0f / rd;
    float rrd2 = rrd * rrd;
    float rrd4 = rrd2 * rrd2;
    float rrd6 = rrd2 * rrd4;
    float rrd7 = rrd6 * rrd;
    float r148 = rrd7 - 0.5f * rrd4;
    float F = 48.0f * r148;
    *fx = dx * F;
    *fy = dy * F;
    *fz = dz * F;
    if (phi) {
        float rrd3 = rrd2 * rrd;
        *phi = 4.0f * (rrd6 - rrd3);
    }
    return 1;
}

static inline int pair_force_f64(const double *pi, const double *pj, double L, double hL, int pbc, double rc2,
                                 double *fx, double *fy, double *fz, double *phi)
{
    double dx = pi[0] - pj[0];
    double dy = pi[1] - pj[1];
    double dz = pi[2] - pj[2];
    if (pbc) {
        if (dx < -hL) dx += L;
        if (dy < -hL) dy += L;
        if (dz < -hL) dz += L;
        if (dx > hL) dx -= L;
        if (dy > hL) dy -= L;
        if (dz > hL) dz -= L;
    }
    double rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    if (!(rd < rc2)) return 0;
    double rrd = 1.0 / rd;
    double rrd2 = rrd * rrd;
    double rrd4 = rrd2 * rrd2;
    double rrd6 = rrd2 * rrd4;
    double rrd7 = rrd6 * rrd;
    double r148 = rrd7 - 0.5 * rrd4;
    double F = 48.0 * r148;
    *fx = dx * F;
    *fy = dy * F;
    *fz = dz * F;
    if (phi) {
        double rrd3 = rrd2 * rrd;
        *phi = 4.0 * (rrd6 - rrd3);
    }
    return 1;
}

/*
 * The reference's LJ functors in their own arithmetic, applied serially over
 * the edge list with the reduction object's semantics (the first value for a
 * node copied, later ones added, in FP32): FP64 positions (N x 3), FP64 pair
 * arithmetic, each force component rounded to float where it is inserted,
 * the j side the negated float.
 *   form 0: moldyn's map (kmeans/irregular.h:26-92): minimum image in a box
 *           of side L, rd < rc^2 with rc^2 = cutoff^2, rrd = 1/rd,
 *           rrd2 = rrd^2, rrd4 = rrd2^2, rrd6 = rrd2 rrd4, rrd7 = rrd6 rrd,
 *           r148 = rrd7 - 0.5 rrd4, force = d r148 (no 48 eps prefactor);
 *   form 1: miniMD's map1 (miniMD/irregular.h:109-163): no images,
 *           rd < cutforcesq, rrd3 = rrd2 rrd, force = d (rrd3 (rrd3 - 0.5) rrd).
 * cut2 is rc^2 (form 0) or cutforcesq (form 1).  f: N x 3 FP32 (zero where
 * nothing was inserted).
 */
void orc_ir_lj_ref(int64_t N, int64_t E, const int32_t *edges, const double *pos, double L, double cut2, int form,
                   float *f)
{
    const double hL = 0.5 * L;
    unsigned char *used = (unsigned char *)calloc((size_t)(N > 0 ? N : 1), 1);
    memset(f, 0, sizeof(float) * 3 * (size_t)N);
    for (int64_t e = 0; e < E; e++) {
        const int i = edges[2 * e], j = edges[2 * e + 1];
        double d[3];
        for (int c = 0; c < 3; c++) d[c] = pos[3 * (int64_t)i + c] - pos[3 * (int64_t)j + c];
        if (form == 0) {
            for (int c = 0; c < 3; c++)
                if (d[c] < -hL) d[c] += L;
            for (int c = 0; c < 3; c++)
                if (d[c] > hL) d[c] -= L;
        }
        const double rd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(rd < cut2)) continue;
        const double rrd = 1.0 / rd, rrd2 = rrd * rrd;
        double s;
        if (form == 0) {
            const double rrd4 = rrd2 * rrd2, rrd6 = rrd2 * rrd4, rrd7 = rrd6 * rrd;
            s = rrd7 - 0.5 * rrd4;
        } else {
            const double rrd3 = rrd2 * rrd;
            s = rrd3 * (rrd3 - 0.5) * rrd;
        }
        float v[3];
        for (int c = 0; c < 3; c++) v[c] = (float)(d[c] * s);
        for (int side = 0; side < 2; side++) {
            const int64_t n = side ? j : i;
            if (side)
                for (int c = 0; c < 3; c++) v[c] = -v[c];
            if (!used[n]) {
                for (int c = 0; c < 3; c++) f[3 * n + c] = v[c];
                used[n] = 1;
            } else {
                for (int c = 0; c < 3; c++) f[3 * n + c] += v[c];
            }
        }
    }
    free(used);
}

/*
 * Serial scatter over a half edge list, f[i] += F*d, f[j] -= F*d, in edge
 * order: the serial semantics of miniMD Force::compute (miniMD/force.cpp:
 * 88-116) and of moldyn's commented-out scatter (kmeans/irregular.h:94-100),
 * which the GPU engine's LDS hash reduction is meant to equal.
 * pos, f: N x 3 row-major.  Returns the potential energy sum_pairs 4(r^-12 - r^-6).
 */
double orc_lj_force_f32(int64_t N, int64_t E, const int32_t *edges, const float *pos, float L, int pbc,
                        float rc2, float *f)
{
    const float hL = 0.5f * L;
    double epot = 0.0;
    memset(f, 0, sizeof(float) * 3 * N);
    for (int64_t e = 0; e < E; e++) {
        int i = edges[2 * e], j = edges[2 * e + 1];
        float fx, fy, fz, phi;
        if (pair_force_f32(pos + 3 * (int64_t)i, pos + 3 * (int64_t)j, L, hL, pbc, rc2, &fx, &fy, &fz, &phi)) {
            f[3 * (int64_t)i + 0] += fx;
            f[3 * (int64_t)i + 1] += fy;
            f[3 * (int64_t)i + 2] += fz;
            f[3 * (int64_t)j + 0] -= fx;
            f[3 * (int64_t)j + 1] -= fy;
            f[3 * (int64_t)j + 2] -= fz;
            epot += phi;
        }
    }
    return epot;
}

double orc_lj_force_f64(int64_t N, int64_t E, const int32_t *edges, const double *pos, double L, int pbc,
                        double rc2, double *f)
{
    const double hL = 0.5 * L;
    double epot = 0.0;
    memset(f, 0, sizeof(double) * 3 * N);
    for (int64_t e = 0; e < E; e++) {
        int i = edges[2 * e], j = edges[2 * e + 1];
        double fx, fy, fz, phi;
        if (pair_force_f64(pos + 3 * (int64_t)i, pos + 3 * (int64_t)j, L, 0.5 * L, pbc, rc2, &fx, &fy, &fz, &phi)) {
            f[3 * (int64_t)i + 0] += fx;
            f[3 * (int64_t)i + 1] += fy;
            f[3 * (int64_t)i + 2] += fz;
            f[3 * (int64_t)j + 0] -= fx;
            f[3 * (int64_t)j + 1] -= fy;
            f[3 * (int64_t)j + 2] -= fz;
            epot += phi;
        }
    }
    (void)hL;
    return epot;
}

/*
 * Threaded CPU path in the shape of the reference's CPU irregular engine
 * (compute_cpu, kmeans/irregular/cpu_kernel.cu:13-70): threads take chunks
 * of edges (CPU_BLOCK_SIZE, kmeans/irregular/parameters.h:8, here dynamic
 * chunks of 2^16) into thread-private node accumulators, which are merged
 * after the join (the reference's merge is racy and partition-0-only,
 * cpu_kernel.cu:65 / irregular_runtime.cu:418-425; fixed here).
 * scratch: nthreads * 3N floats.
 */
void orc_lj_force_f32_threads(int64_t N, int64_t E, const int32_t *edges, const float *pos, float L, int pbc,
                              float rc2, float *f, float *scratch, int nthreads)
{
    const float hL = 0.5f * L;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num();
        float *acc = scratch + (int64_t)tid * 3 * N;
        memset(acc, 0, sizeof(float) * 3 * N);
        #pragma omp for schedule(dynamic, 65536)
        for (int64_t e = 0; e < E; e++) {
            int i = edges[2 * e], j = edges[2 * e + 1];
            float fx, fy, fz;
            if (pair_force_f32(pos + 3 * (int64_t)i, pos + 3 * (int64_t)j, L, hL, pbc, rc2, &fx, &fy, &fz, 0)) {
                acc[3 * (int64_t)i + 0] += fx;
                acc[3 * (int64_t)i + 1] += fy;
                acc[3 * (int64_t)i + 2] += fz;
                acc[3 * (int64_t)j + 0] -= fx;
                acc[3 * (int64_t)j + 1] -= fy;
                acc[3 * (int64_t)j + 2] -= fz;
            }
        }
        int nt = omp_get_num_threads();
        #pragma omp for schedule(static)
        for (int64_t m = 0; m < 3 * N; m++) {
            float s = 0.0f;
            for (int t = 0; t < nt; t++) s += scratch[(int64_t)t * 3 * N + m];
            f[m] = s;
        }
    }
}

/*
 * Gather form of the serial scatter: for each node, the contributions of its
 * incident edges are added in edge order, exactly as orc_lj_force_f32 adds
 * them, so the result is bit-identical to the serial scatter while nodes are
 * processed in parallel.  The incidence lists are built once per call by a
 * stable counting sort of (node, edge).  Used where the serial order must be
 * kept at full size.
 */
void orc_lj_force_f32_gather(int64_t N, int64_t E, const int32_t *edges, const float *pos, float L, int pbc,
                             float rc2, float *f, int nthreads)
{
    const float hL = 0.5f * L;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int64_t *start = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    int64_t *inc = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(E > 0 ? E : 1));
    for (int64_t e = 0; e < E; e++) {
        start[edges[2 * e] + 1]++;
        start[edges[2 * e + 1] + 1]++;
    }
    for (int64_t i = 0; i < N; i++) start[i + 1] += start[i];
    {
        int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
        memcpy(cur, start, sizeof(int64_t) * N);
        for (int64_t e = 0; e < E; e++) {  /* edge order within each node */
            inc[cur[edges[2 * e]]++] = 2 * e;          /* node is the i side */
            inc[cur[edges[2 * e + 1]]++] = 2 * e + 1;  /* node is the j side */
        }
        free(cur);
    }
    #pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
    for (int64_t r = 0; r < N; r++) {
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for (int64_t k = start[r]; k < start[r + 1]; k++) {
            const int64_t e = inc[k] >> 1;
            const int side = (int)(inc[k] & 1);
            const int i = edges[2 * e], j = edges[2 * e + 1];
            float fx, fy, fz;
            if (pair_force_f32(pos + 3 * (int64_t)i, pos + 3 * (int64_t)j, L, hL, pbc, rc2, &fx, &fy, &fz, 0)) {
                if (side == 0) {
                    ax += fx;
                    ay += fy;
                    az += fz;
                } else {
                    ax -= fx;
                    ay -= fy;
                    az -= fz;
                }
            }
        }
        f[3 * r] = ax;
        f[3 * r + 1] = ay;
        f[3 * r + 2] = az;
    }
    free(start);
    free(inc);
}

/*
 * FP64 form of orc_lj_force_f32_gather (same incidence order, pair_force_f64
 * arithmetic): the FP64 shadow of the full-size runs, threaded over nodes.
 */
void orc_lj_force_f64_gather(int64_t N, int64_t E, const int32_t *edges, const double *pos, double L, int pbc,
                             double rc2, double *f, int nthreads)
{
    const double hL = 0.5 * L;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int64_t *start = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    int64_t *inc = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(E > 0 ? E : 1));
    for (int64_t e = 0; e < E; e++) {
        start[edges[2 * e] + 1]++;
        start[edges[2 * e + 1] + 1]++;
    }
    for (int64_t i = 0; i < N; i++) start[i + 1] += start[i];
    {
        int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N > 0 ? N : 1));
        memcpy(cur, start, sizeof(int64_t) * N);
        for (int64_t e = 0; e < E; e++) {
            inc[cur[edges[2 * e]]++] = 2 * e;
            inc[cur[edges[2 * e + 1]]++] = 2 * e + 1;
        }
        free(cur);
    }
    #pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
    for (int64_t r = 0; r < N; r++) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t k = start[r]; k < start[r + 1]; k++) {
            const int64_t e = inc[k] >> 1;
            const int side = (int)(inc[k] & 1);
            const int i = edges[2 * e], j = edges[2 * e + 1];
            double fx, fy, fz;
            if (pair_force_f64(pos + 3 * (int64_t)i, pos + 3 * (int64_t)j, L, hL, pbc, rc2, &fx, &fy, &fz, 0)) {
                if (side == 0) {
                    ax += fx;
                    ay += fy;
                    az += fz;
                } else {
                    ax -= fx;
                    ay -= fy;
                    az -= fz;
                }
            }
        }
        f[3 * r] = ax;
        f[3 * r + 1] = ay;
        f[3 * r + 2] = az;
    }
    free(start);
    free(inc);
}

/*
 * Cell-list half pair list: every pair i<j whose minimum-image distance
 * squared (same float expression as orc_pairs_brute) is below cut2, in
 * (i, j) lexicographic order.  Cells are at least one cutoff wide (needs
 * L >= 3 * cutoff).  The binned build of miniMD's Neighbor::build
 * (miniMD/neighbor.cpp:76-205), with periodic images.  Returns the count;
 * writes at most cap pairs.
 */
static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_pairs_cells(int64_t N, const float *pos, float L, float cut, int32_t *edges, int64_t cap)
{
    const float hL = 0.5f * L, cut2 = cut * cut;
    int nb = (int)floor((double)L / ((double)cut * 1.0001));
    if (nb < 3) return -1;
    const int64_t ncell = (int64_t)nb * nb * nb;
    int64_t *head = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t *cellof = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int64_t i = 0; i < N; i++) {
        int c[3];
        for (int d = 0; d < 3; d++) {
            int q = (int)floor((double)pos[3 * i + d] * nb / L);
            c[d] = q < 0 ? 0 : (q >= nb ? nb - 1 : q);
        }
        cellof[i] = c[0] + (int64_t)nb * (c[1] + (int64_t)nb * c[2]);
        head[cellof[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) head[c + 1] += head[c];
    {
        int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
        memcpy(cur, head, sizeof(int64_t) * ncell);
        for (int64_t i = 0; i < N; i++) order[cur[cellof[i]]++] = (int32_t)i;
        free(cur);
    }
    int64_t n = 0;
    int32_t *row = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    for (int64_t i = 0; i < N; i++) {
        const int64_t c = cellof[i];
        const int cx = (int)(c % nb), cy = (int)((c / nb) % nb), cz = (int)(c / ((int64_t)nb * nb));
        int64_t m = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int qx = (cx + dx + nb) % nb, qy = (cy + dy + nb) % nb, qz = (cz + dz + nb) % nb;
                    const int64_t q = qx + (int64_t)nb * (qy + (int64_t)nb * qz);
                    for (int64_t s = head[q]; s < head[q + 1]; s++) {
                        const int32_t j = order[s];
                        if (j <= i) continue;
                        float ddx = pos[3 * i] - pos[3 * j];
                        float ddy = pos[3 * i + 1] - pos[3 * j + 1];
                        float ddz = pos[3 * i + 2] - pos[3 * j + 2];
                        if (ddx < -hL) ddx += L;
                        if (ddy < -hL) ddy += L;
                        if (ddz < -hL) ddz += L;
                        if (ddx > hL) ddx -= L;
                        if (ddy > hL) ddy -= L;
                        if (ddz > hL) ddz -= L;
                        float r2 = ddx * ddx + ddy * ddy;
                        r2 = r2 + ddz * ddz;
                        if (r2 < cut2) row[m++] = j;
                    }
                }
        qsort(row, (size_t)m, sizeof(int32_t), cmp_i32);
        for (int64_t t = 0; t < m; t++) {
            if (edges && n < cap) {
                edges[2 * n] = (int32_t)i;
                edges[2 * n + 1] = row[t];
            }
            n++;
        }
    }
    free(row);
    free(head);
    free(cellof);
    free(order);
    return n;
}

int64_t orc_pairs_brute(int64_t N, const float *pos, float L, int pbc, float cut2, int32_t *edges, int64_t cap);

/*
 * The rebuild loops' pair list: the cell list where the box holds three
 * cells of the cutoff, else (boxes of two cells, where every pair has at
 * most one image within the cutoff) the O(N^2) list -- the same pairs, in
 * the same (i, j) order.  Returns -1 for boxes under two cells.
 */
int64_t orc_pairs_any(int64_t N, const float *pos, float L, float cut, int32_t *edges, int64_t cap)
{
    const int nb = (int)floor((double)L / ((double)cut * 1.0001));
    if (nb >= 3) return orc_pairs_cells(N, pos, L, cut, edges, cap);
    if (nb < 2) return -1;
    return orc_pairs_brute(N, pos, L, 1, cut * cut, edges, cap);
}

/*
 * Velocity-Verlet with unit mass and periodic wrap into [0, L), the
 * integrator the moldyn and miniMD workloads run around the force
 * reduction (moldyn's own update is a leapfrog with folded constants,
 * kmeans/moldyn.cu:339-403; miniMD's is commented out, miniMD/ljs.cu:178-184).
 * Per node and per step, in this order:
 *   v = v + hdt*f;  x = x + dt*v;  wrap(x);  f = F(x);  v = v + hdt*f
 * with hdt = 0.5f*dt rounded once.  Starts by computing f = F(x0).
 * pos, vel updated in place; f returns the last force.  Returns epot of the
 * last force evaluation.
 */
static inline float wrap_f32(float x, float L)
{
    if (x < 0.0f) x += L;
    else if (x >= L) x -= L;
    return x;
}

double orc_md_run_f32(int64_t N, int64_t E, const int32_t *edges, float *pos, float *vel, float *f, float L,
                      int pbc, float rc2, float dt, int nsteps, int nthreads, float *scratch)
{
    /* Forces: serial order (gather form, any thread count) unless scratch is
     * given, in which case the reference CPU engine's chunked thread-private
     * accumulation is used (its sums round in another order). */
    const float hdt = 0.5f * dt;
    double epot = 0.0;
#define ORC_FORCE()                                                                                                    \
    do {                                                                                                               \
        if (scratch) orc_lj_force_f32_threads(N, E, edges, pos, L, pbc, rc2, f, scratch, nthreads);                    \
        else if (nthreads > 1) orc_lj_force_f32_gather(N, E, edges, pos, L, pbc, rc2, f, nthreads);                    \
        else epot = orc_lj_force_f32(N, E, edges, pos, L, pbc, rc2, f);                                                \
    } while (0)
    ORC_FORCE();
    for (int s = 0; s < nsteps; s++) {
        for (int64_t m = 0; m < 3 * N; m++) {
            vel[m] = vel[m] + hdt * f[m];
            float x = pos[m] + dt * vel[m];
            pos[m] = pbc ? wrap_f32(x, L) : x;
        }
        ORC_FORCE();
        for (int64_t m = 0; m < 3 * N; m++) vel[m] = vel[m] + hdt * f[m];
    }
#undef ORC_FORCE
    return epot;
}

/*
 * The MD loop with a neighbour list rebuilt from the positions before the
 * force of step 0 and of every step whose index is a multiple of 'every'
 * (miniMD's reneighbouring, miniMD/lj.in:9-10): pairs within cutneigh by
 * orc_pairs_cells, forces within rc by the gather form (serial order).
 */
double orc_md_run_rebuild_f32(int64_t N, float *pos, float *vel, float *f, float L, float rc2, float cutneigh,
                              int every, float dt, int nsteps, int nthreads, int64_t *npairs_out)
{
    const float hdt = 0.5f * dt;
    int64_t cap = 0, np = 0;
    int32_t *edges = NULL;
    for (int s = 0; s <= nsteps; s++) {
        if (s > 0) {
            for (int64_t m = 0; m < 3 * N; m++) {
                vel[m] = vel[m] + hdt * f[m];
                pos[m] = wrap_f32(pos[m] + dt * vel[m], L);
            }
        }
        if (s == 0 || s % every == 0) {
            np = orc_pairs_any(N, pos, L, cutneigh, NULL, 0);
            if (np < 0) {
                free(edges);
                return NAN;
            }
            if (np > cap) {
                free(edges);
                cap = np;
                edges = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(cap > 0 ? cap : 1));
            }
            orc_pairs_any(N, pos, L, cutneigh, edges, cap);
        }
        orc_lj_force_f32_gather(N, np, edges, pos, L, 1, rc2, f, nthreads);
        if (s > 0)
            for (int64_t m = 0; m < 3 * N; m++) vel[m] = vel[m] + hdt * f[m];
    }
    if (npairs_out) *npairs_out = np;
    free(edges);
    return 0.0;
}

static inline double wrap_f64(double x, double L)
{
    if (x < 0.0) x += L;
    else if (x >= L) x -= L;
    return x;
}

/*
 * FP64 form of orc_md_run_rebuild_f32: the same velocity-Verlet loop and
 * rebuild schedule, positions, velocities and forces in FP64.  The pair list
 * is built from the positions rounded to FP32 with the neighbour cutoff
 * (cutneigh > rc by the skin, so every pair within rc is listed); forces
 * use the FP64 minimum image and cutoff test.
 */
double orc_md_run_rebuild_f64(int64_t N, double *pos, double *vel, double *f, double L, double rc2, double cutneigh,
                              int every, double dt, int nsteps, int nthreads, int64_t *npairs_out)
{
    const double hdt = 0.5 * dt;
    int64_t cap = 0, np = 0;
    int32_t *edges = NULL;
    float *p32 = (float *)malloc(sizeof(float) * 3 * (size_t)(N > 0 ? N : 1));
    for (int s = 0; s <= nsteps; s++) {
        if (s > 0) {
            for (int64_t m = 0; m < 3 * N; m++) {
                vel[m] = vel[m] + hdt * f[m];
                pos[m] = wrap_f64(pos[m] + dt * vel[m], L);
            }
        }
        if (s == 0 || s % every == 0) {
            for (int64_t m = 0; m < 3 * N; m++) {
                float x = (float)pos[m];
                p32[m] = x >= (float)L ? 0.0f : x; /* rounding may reach L */
            }
            np = orc_pairs_any(N, p32, (float)L, (float)cutneigh, NULL, 0);
            if (np < 0) {
                free(edges);
                free(p32);
                return NAN;
            }
            if (np > cap) {
                free(edges);
                cap = np;
                edges = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(cap > 0 ? cap : 1));
            }
            orc_pairs_any(N, p32, (float)L, (float)cutneigh, edges, cap);
        }
        orc_lj_force_f64_gather(N, np, edges, pos, L, 1, rc2, f, nthreads);
        if (s > 0)
            for (int64_t m = 0; m < 3 * N; m++) vel[m] = vel[m] + hdt * f[m];
    }
    if (npairs_out) *npairs_out = np;
    free(edges);
    free(p32);
    return 0.0;
}

double orc_md_run_f64(int64_t N, int64_t E, const int32_t *edges, double *pos, double *vel, double *f, double L,
                      int pbc, double rc2, double dt, int nsteps)
{
    const double hdt = 0.5 * dt;
    double epot = orc_lj_force_f64(N, E, edges, pos, L, pbc, rc2, f);
    for (int s = 0; s < nsteps; s++) {
        for (int64_t m = 0; m < 3 * N; m++) {
            vel[m] = vel[m] + hdt * f[m];
            double x = pos[m] + dt * vel[m];
            pos[m] = pbc ? wrap_f64(x, L) : x;
        }
        epot = orc_lj_force_f64(N, E, edges, pos, L, pbc, rc2, f);
        for (int64_t m = 0; m < 3 * N; m++) vel[m] = vel[m] + hdt * f[m];
    }
    return epot;
}

/*
 * Brute-force O(N^2) half list: every pair i<j with minimum-image distance
 * squared below cut2, emitted in (i, j) lexicographic order.  The moldyn
 * program's RebuildNeighbor (kmeans/moldyn.cu:552-594) is this O(N^2) loop.
 * Returns the pair count; writes at most cap pairs when edges is non-NULL.
 */
int64_t orc_pairs_brute(int64_t N, const float *pos, float L, int pbc, float cut2, int32_t *edges, int64_t cap)
{
    const float hL = 0.5f * L;
    int64_t n = 0;
    for (int64_t i = 0; i < N; i++)
        for (int64_t j = i + 1; j < N; j++) {
            float dx = pos[3 * i] - pos[3 * j];
            float dy = pos[3 * i + 1] - pos[3 * j + 1];
            float dz = pos[3 * i + 2] - pos[3 * j + 2];
            if (pbc) {
                if (dx < -hL) dx += L;
                if (dy < -hL) dy += L;
                if (dz < -hL) dz += L;
                if (dx > hL) dx -= L;
                if (dy > hL) dy -= L;
                if (dz > hL) dz -= L;
            }
            float r2 = dx * dx + dy * dy;
            r2 = r2 + dz * dz;
            if (r2 < cut2) {
                if (edges && n < cap) {
                    edges[2 * n] = (int32_t)i;
                    edges[2 * n + 1] = (int32_t)j;
                }
                n++;
            }
        }
    return n;
}

/* ======================================================================
 * Generalized-reduction engine: kmeans
 * ====================================================================== */

/*
 * Assignment of kmeans map1 (kmeans/regular.h:84-103), generalised from
 * D=3/K=40 to any D, K: for each centre i in order,
 * dist = sum_d (c_d - x_d)^2 accumulated in dimension order in FP32,
 * dist = sqrt(dist), and the strict comparison dist < min_dist starting from
 * min_dist = 65536*65, so the lowest index wins ties.  Writes the key and
 * the emitted min_dist (value[4] of map1, :110-115).
 */
void orc_kmeans_assign_f32(int64_t N, int D, int K, const float *x, const float *c, int32_t *assign,
                           float *mindist, int nthreads)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel for schedule(dynamic, 10000) num_threads(nthreads)
    for (int64_t p = 0; p < N; p++) {
        const float *xp = x + p * D;
        unsigned int key = 0;
        float min_dist = 65536 * 65;
        for (int i = 0; i < K; i++) {
            const float *ci = c + (int64_t)i * D;
            float dist = 0;
            for (int d = 0; d < D; d++) {
                float t = ci[d] - xp[d];
                dist = dist + t * t;
            }
            dist = sqrtf(dist);
            if (dist < min_dist) {
                min_dist = dist;
                key = i;
            }
        }
        assign[p] = (int32_t)key;
        if (mindist) mindist[p] = min_dist;
    }
}

/* FP64 twin of the assignment, for error statements. */
void orc_kmeans_assign_f64(int64_t N, int D, int K, const float *x, const double *c, int32_t *assign,
                           double *mindist, int nthreads)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel for schedule(dynamic, 10000) num_threads(nthreads)
    for (int64_t p = 0; p < N; p++) {
        const float *xp = x + p * D;
        int key = 0;
        double min_dist = 65536.0 * 65;
        for (int i = 0; i < K; i++) {
            const double *ci = c + (int64_t)i * D;
            double dist = 0;
            for (int d = 0; d < D; d++) {
                double t = ci[d] - (double)xp[d];
                dist = dist + t * t;
            }
            dist = sqrt(dist);
            if (dist < min_dist) {
                min_dist = dist;
                key = i;
            }
        }
        assign[p] = key;
        if (mindist) mindist[p] = min_dist;
    }
}

/*
 * Fixed-point scale for order-independent centroid sums.  The reference's
 * reduce1 (kmeans/regular.h:121-136) adds FP32 values in whatever order the
 * locked hash inserts happen, which is not reproducible; here each coordinate
 * is rounded once to an integer multiple of 2^-s and summed as int64, so any
 * summation order gives the same bits.  s is chosen so that max|x| * 2^s
 * stays within 2^30 (each rounded value fits in 31 bits) and N * max|x| * 2^s
 * below 2^62.  Returns s.
 */
int orc_kmeans_fixed_shift(int64_t N, float maxabs)
{
    int eN = 0, eM = 0;
    while ((1LL << eN) < N) eN++;
    double m = maxabs > 0 ? maxabs : 1.0f;
    frexp(m, &eM); /* m < 2^eM */
    int s = 62 - eN - eM;
    if (30 - eM < s) s = 30 - eM;
    if (s > 60) s = 60;
    return s;
}

static inline int64_t to_fixed(float v, double scale)
{
    return (int64_t)rint((double)v * scale);
}

/*
 * Reduction object of kmeans map1/reduce1 (values {x_0..x_{D-1}, 1, min_dist},
 * kmeans/regular.h:110-115, summed per key by reduce1, :121-136) with an
 * integer count and fixed-point sums.  In the shape of the reference's CPU
 * regular engine (compute_cpu_reg, kmeans/regular/cpu_kernel.cu:17-81):
 * threads take chunks of units under a shared counter into thread-private
 * reduction objects, merged after the join (the reference merges racily
 * while other threads still map, cpu_kernel.cu:77-78; fixed here).
 * sums: K*D int64, counts: K int64, dsum: K int64 (min_dist in fixed point
 * with dshift).
 */
void orc_kmeans_accumulate(int64_t N, int D, int K, const float *x, const int32_t *assign,
                           const float *mindist, int shift, int dshift, int64_t *sums, int64_t *counts,
                           int64_t *dsum, int nthreads)
{
    const double scale = ldexp(1.0, shift), dscale = ldexp(1.0, dshift);
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    memset(sums, 0, sizeof(int64_t) * K * D);
    memset(counts, 0, sizeof(int64_t) * K);
    if (dsum) memset(dsum, 0, sizeof(int64_t) * K);
    #pragma omp parallel num_threads(nthreads)
    {
        int64_t *ls = (int64_t *)calloc((size_t)K * D, sizeof(int64_t));
        int64_t *lc = (int64_t *)calloc((size_t)K, sizeof(int64_t));
        int64_t *ld = (int64_t *)calloc((size_t)K, sizeof(int64_t));
        #pragma omp for schedule(dynamic, 10000)
        for (int64_t p = 0; p < N; p++) {
            int k = assign[p];
            const float *xp = x + p * D;
            for (int d = 0; d < D; d++) ls[(int64_t)k * D + d] += to_fixed(xp[d], scale);
            lc[k] += 1;
            if (mindist) ld[k] += to_fixed(mindist[p], dscale);
        }
        #pragma omp critical
        {
            for (int64_t m = 0; m < (int64_t)K * D; m++) sums[m] += ls[m];
            for (int k = 0; k < K; k++) counts[k] += lc[k];
            if (dsum)
                for (int k = 0; k < K; k++) dsum[k] += ld[k];
        }
        free(ls);
        free(lc);
        free(ld);
    }
}

/*
 * Lloyd update, which the reference leaves out (km.cu runs one map pass and
 * RegularGetOutput is empty, kmeans/regular/regular_runtime.cu:288-291):
 * c_k = (float)(((double)sum_k * 2^-s) / (double)count_k); a cluster with no
 * points keeps its previous centre.
 */
void orc_kmeans_update(int D, int K, const int64_t *sums, const int64_t *counts, int shift, float *c)
{
    const double inv = ldexp(1.0, -shift);
    for (int k = 0; k < K; k++) {
        if (counts[k] == 0) continue;
        double n = (double)counts[k];
        for (int d = 0; d < D; d++) c[(int64_t)k * D + d] = (float)(((double)sums[(int64_t)k * D + d] * inv) / n);
    }
}

/*
 * Full Lloyd loop: centres start as the first K points (the workload's
 * initialisation), then iters x (assign, accumulate, update).  Final
 * assignment is of the last iteration (made against the centres that
 * iteration started from).  sse_out (optional, iters entries) receives
 * sum_p min_dist_p^2 in FP64 per iteration.
 */
void orc_kmeans_lloyd(int64_t N, int D, int K, const float *x, int iters, int shift, float *c, int32_t *assign,
                      double *sse_out, int nthreads)
{
    int64_t *sums = (int64_t *)malloc(sizeof(int64_t) * K * D);
    int64_t *counts = (int64_t *)malloc(sizeof(int64_t) * K);
    float *md = (float *)malloc(sizeof(float) * N);
    for (int it = 0; it < iters; it++) {
        orc_kmeans_assign_f32(N, D, K, x, c, assign, md, nthreads);
        if (sse_out) {
            double s = 0;
            for (int64_t p = 0; p < N; p++) s += (double)md[p] * md[p];
            sse_out[it] = s;
        }
        orc_kmeans_accumulate(N, D, K, x, assign, 0, shift, 0, sums, counts, 0, nthreads);
        orc_kmeans_update(D, K, sums, counts, shift, c);
    }
    free(sums);
    free(counts);
    free(md);
}

/* FP64 Lloyd loop with plain double sums, for error statements. */
void orc_kmeans_lloyd_f64(int64_t N, int D, int K, const float *x, int iters, double *c, int32_t *assign,
                          int nthreads)
{
    double *sums = (double *)malloc(sizeof(double) * K * D);
    int64_t *counts = (int64_t *)malloc(sizeof(int64_t) * K);
    for (int it = 0; it < iters; it++) {
        orc_kmeans_assign_f64(N, D, K, x, c, assign, 0, nthreads);
        memset(sums, 0, sizeof(double) * K * D);
        memset(counts, 0, sizeof(int64_t) * K);
        for (int64_t p = 0; p < N; p++) {
            int k = assign[p];
            for (int d = 0; d < D; d++) sums[(int64_t)k * D + d] += x[p * D + d];
            counts[k]++;
        }
        for (int k = 0; k < K; k++)
            if (counts[k])
                for (int d = 0; d < D; d++) c[(int64_t)k * D + d] = sums[(int64_t)k * D + d] / counts[k];
    }
    free(sums);
    free(counts);
}

/* ======================================================================
 * Generalized reduction with additive functors (map/reduce pairs whose
 * reduce adds values), the work the reference runs through RegularRuntime
 * beside kmeans.  Keys: the reference spreads emits over ANY_KEY buckets
 * (threadIdx % 41 on the GPU, 0 on the CPU, kmeans/regular/
 * gpu_kernel_tools.cu:68-71) and the caller sums them; here every emit goes
 * to key 0, and the sums are FP64 in unit order.
 * ====================================================================== */

/*
 * Velocity statistics, map/reduce of kmeans/regular.h:30-73 as moldyn calls
 * it (kmeans/moldyn.cu:440-470): per unit (double[3] velocity) value[0] =
 * (float)sqrt(vx*vx + vy*vy + vz*vz) in FP64, value[1] = value[0] > vaverh
 * (FP32 compare); reduce adds both.  out[0] = sum of value[0] (FP64, in unit
 * order), out[1] = the count; per-unit values go to per (float, 2 per unit)
 * when not NULL.
 */
void orc_gr_vstats(int64_t n, const double *vel, float vaverh, double *out, float *per)
{
    double s = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; i++) {
        const double *v = vel + 3 * i;
        float value0 = (float)sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        float value1 = value0 > vaverh ? 1.0f : 0.0f;
        if (per) {
            per[2 * i] = value0;
            per[2 * i + 1] = value1;
        }
        s += (double)value0;
        c += (double)value1;
    }
    out[0] = s;
    out[1] = c;
}

/*
 * Pair energy, map2/reduce2 of miniMD/regular.h:139-181: per unit (int[2]
 * edge i, j) with parameter {cutsq, x[3N]} (FP64): rsq = dx*dx + dy*dy +
 * dz*dz (no minimum image), and when rsq < cutsq emit phi = sr6*(sr6 - 1),
 * sr2 = 1/rsq, sr6 = sr2*sr2*sr2; reduce adds.  Returns the FP64 sum in unit
 * order; *emits = the number of pairs within the cutoff.
 */
double orc_gr_lj_energy(int64_t e, const int32_t *edges, const double *x, double cutsq, int64_t *emits)
{
    double s = 0.0;
    int64_t c = 0;
    for (int64_t k = 0; k < e; k++) {
        const int64_t i = edges[2 * k], j = edges[2 * k + 1];
        double delx = x[3 * i] - x[3 * j];
        double dely = x[3 * i + 1] - x[3 * j + 1];
        double delz = x[3 * i + 2] - x[3 * j + 2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq < cutsq) {
            double sr2 = 1.0 / rsq;
            double sr6 = sr2 * sr2 * sr2;
            double phi = sr6 * (sr6 - 1.0);
            s += phi;
            c++;
        }
    }
    *emits = c;
    return s;
}
