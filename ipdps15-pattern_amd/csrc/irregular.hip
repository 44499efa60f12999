// Irregular-reduction engine for gfx950: the Lennard-Jones edge loop of the
// moldyn and miniMD workloads, reducing per-edge forces into node arrays.
//
// Replaces the reference's irregular compute_gpu (kmeans/irregular/
// gpu_kernel.cu:21-103): 14 persistent blocks, each grabbing one LDS-sized
// partition of edges (built on the host by Partitioner/Reorder,
// kmeans/irregular/partitioner.cpp:20-499, reorder.cpp:13-281), calling the
// edge functor (kmeans/irregular.h:26-107) which inserts +F and -F into an
// LDS open-addressing hash (ros.cu:10-68), then merging the hash into a
// device hash (rog.cu:11-100).
//
// Here the edges are reordered once on the device into a by-destination CSR
// (every half edge (i, j) appears in row i and in row j): a degree histogram,
// an exclusive scan, a scatter, and a per-row sort that makes the layout --
// and so every later force sum -- deterministic.  The force kernel then gives
// each row to a group of LPR lanes of a wavefront: the lanes stride the row's
// neighbour list, each computing the pair force with exactly the functor's
// arithmetic, and the group reduces its partial sums with cross-lane
// butterflies, so a node's update is a wavefront-segmented reduction with no
// atomics and one coalesced write per node.  The velocity-Verlet update is
// fused into that write.  The plain edge-parallel kernel with float atomics
// (the shape of the reference's scatter) is kept as the baseline.
//
// Pair arithmetic (per pair, identical in the oracle): d = x_i - x_j,
// minimum image (moldyn; miniMD's map1 has none), rd = (dx^2 + dy^2) + dz^2,
// for rd < rc^2: rrd = 1/rd, rrd2 = rrd^2, rrd4 = rrd2^2, rrd6 = rrd2*rrd4,
// rrd7 = rrd6*rrd, F = 48*(rrd7 - 0.5*rrd4) (the 48*eps factor the
// reference folds into its integrator, kmeans/moldyn.cu:396-398), force on
// i = d*F.  For the reversed pair the displacement is exactly -d, so each
// row entry is exactly the negation the serial scatter applies.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pat_common.hpp"
#include "pat_device.hpp"

namespace {

using pat::kWave;

// ------------------------------------------------------------------ CSR build

__global__ void csr_degree(const int2 *__restrict__ edges, int64_t E, int *__restrict__ deg)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        int2 ij = edges[e];
        atomicAdd(&deg[ij.x], 1);
        atomicAdd(&deg[ij.y], 1);
    }
}

constexpr int kScanBlock = 1024;

// Block-level exclusive scan of in[0..n) into out, block totals into sums.
__global__ __launch_bounds__(kScanBlock) void scan_blocks(const int *__restrict__ in, int *__restrict__ out, int64_t n,
                                                          int *__restrict__ sums)
{
    __shared__ int wsum[kScanBlock / kWave];
    const int64_t g = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int v = g < n ? in[g] : 0;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    int x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int s = lane < kScanBlock / kWave ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            int y = __shfl_up(s, o, kWave);
            if (lane >= o) s += y;
        }
        if (lane < kScanBlock / kWave) wsum[lane] = s;
    }
    __syncthreads();
    const int base = w > 0 ? wsum[w - 1] : 0;
    if (g < n) out[g] = base + x - v;  // exclusive
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = base + x;
}

__global__ __launch_bounds__(kScanBlock) void scan_add(int *__restrict__ out, int64_t n, const int *__restrict__ offs)
{
    const int64_t g = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    if (g < n) out[g] += offs[blockIdx.x];
}

__global__ void csr_scatter(const int2 *__restrict__ edges, int64_t E, int *__restrict__ cursor, int *__restrict__ tmp,
                            int *__restrict__ tmpe)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        int2 ij = edges[e];
        int p = atomicAdd(&cursor[ij.x], 1);
        tmp[p] = ij.y;
        tmpe[p] = (int)e;
        int q = atomicAdd(&cursor[ij.y], 1);
        tmp[q] = ij.x;
        tmpe[q] = (int)e;
    }
}

// One wavefront per row: places each entry at the rank of its edge id, so a
// row lists its neighbours in the order their edges appear in the input list
// -- the order in which a serial scatter over that list adds them.
__global__ __launch_bounds__(256) void csr_row_sort_by_edge(const int *__restrict__ rowptr, int64_t N,
                                                            const int *__restrict__ tmp, const int *__restrict__ tmpe,
                                                            int *__restrict__ cols)
{
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= N) return;
    const int b = rowptr[row], e = rowptr[row + 1];
    for (int k = b + lane; k < e; k += kWave) {
        const int v = tmpe[k];
        int rank = 0;
        for (int m = b; m < e; m++) rank += tmpe[m] < v;
        cols[b + rank] = tmp[k];
    }
}

// One wavefront per row: rank sort of the row's columns (ties broken by
// position), writing the sorted row from tmp into cols.
__global__ __launch_bounds__(256) void csr_row_sort(const int *__restrict__ rowptr, int64_t N, const int *__restrict__ tmp,
                                                    int *__restrict__ cols)
{
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= N) return;
    const int b = rowptr[row], e = rowptr[row + 1];
    for (int k = b + lane; k < e; k += kWave) {
        const int v = tmp[k];
        int rank = 0;
        for (int m = b; m < e; m++) {
            const int u = tmp[m];
            rank += (u < v) || (u == v && m < k);
        }
        cols[b + rank] = v;
    }
}

// ------------------------------------------------------------ neighbour build
//
// Binned rebuild of the full neighbour CSR from positions, with periodic
// images: the device form of miniMD's Neighbor::binatoms / build
// (miniMD/neighbor.cpp:76-250), which bins atoms into cells at least one
// cutoff wide and searches a stencil of neighbouring bins.  Here every pair
// within the cutoff lands in both rows (so the force needs no scatter) and
// rows are sorted, which makes the CSR identical to the one built from a
// host pair list with the same cutoff.

struct BinGeom {
    float L, hL, cut2, inv;  // box, half box, cutoff^2, nb / L
    int nb;                  // bins per axis (>= 3)
};

__device__ __forceinline__ int bin_of(float x, const BinGeom &g)
{
    int c = (int)floorf(x * g.inv);
    return c < 0 ? 0 : (c >= g.nb ? g.nb - 1 : c);
}

__global__ void nb_bin_atoms(const float4 *__restrict__ pos, int N, BinGeom g, int *__restrict__ cell,
                             int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 p = pos[i];
    const int c = bin_of(p.x, g) + g.nb * (bin_of(p.y, g) + g.nb * bin_of(p.z, g));
    cell[i] = c;
    atomicAdd(&count[c], 1);
}

__global__ void nb_scatter_atoms(const int *__restrict__ cell, int N, int *__restrict__ cursor, int *__restrict__ tmp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    tmp[atomicAdd(&cursor[cell[i]], 1)] = i;
}

// Minimum-image pair test with the pair generator's float expression.
__device__ __forceinline__ bool nb_within(float4 a, float4 b, const BinGeom &g)
{
    float dx = pat::min_image(a.x - b.x, g.L, g.hL);
    float dy = pat::min_image(a.y - b.y, g.L, g.hL);
    float dz = pat::min_image(a.z - b.z, g.L, g.hL);
    float r2 = dx * dx + dy * dy;
    r2 = r2 + dz * dz;
    return r2 < g.cut2;
}

// One thread per atom over the 27 neighbouring bins: with fill == 0 counts
// the row length into deg, with fill != 0 writes the row into tmp.
__global__ void nb_rows(const float4 *__restrict__ pos, int N, BinGeom g, const int *__restrict__ cell,
                        const int *__restrict__ cstart, const int *__restrict__ catoms, int *__restrict__ deg,
                        const int *__restrict__ rowptr, int *__restrict__ tmp, int fill)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 pi = pos[i];
    const int c = cell[i];
    const int cx = c % g.nb, cy = (c / g.nb) % g.nb, cz = c / (g.nb * g.nb);
    int n = 0;
    int w = fill ? rowptr[i] : 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int qx = (cx + dx + g.nb) % g.nb, qy = (cy + dy + g.nb) % g.nb, qz = (cz + dz + g.nb) % g.nb;
                const int q = qx + g.nb * (qy + g.nb * qz);
                for (int s = cstart[q]; s < cstart[q + 1]; s++) {
                    const int j = catoms[s];
                    if (j != i && nb_within(pi, pos[j], g)) {
                        if (fill) tmp[w++] = j;
                        n++;
                    }
                }
            }
    if (!fill) deg[i] = n;
}

// ------------------------------------------------------------------ force

struct LJParams {
    float L, hL;   // box side and half side (pbc)
    float rc2;     // cutoff squared
    float dt, hdt; // time step and half step (hdt = 0.5f*dt rounded once)
    int pbc;
};

// 1.0f / b, correctly rounded.  For b in [2^-126, 2^125] one Newton step on
// v_rcp_f32 (e = fma(-b, y, 1); y + e*y) gives the IEEE quotient bit for bit
// -- checked for every FP32 value in that range on gfx950 (the
// pat_lj_rcp_selftest entry point repeats the check) -- at 3 instructions
// instead of the 11 of the division sequence; outside it (0, subnormals,
// above 2^125) the division itself runs.
__device__ __forceinline__ float exact_rcp(float b)
{
    const float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    float r = __builtin_fmaf(e, y, y);
    if (__builtin_expect(!(b >= 0x1p-126f && b <= 0x1p125f), 0)) r = 1.0f / b;
    return r;
}

// Minimum image of one component, the functor's two conditional moves
// (kmeans/irregular.h:55-60) as one subtraction: for |d| < L at most one of
// them applies -- d < -hL or d > hL, i.e. |d| > hL -- and the move is
// d - copysign(L, d) (d - (-L) is exactly d + L).
__device__ __forceinline__ float min_image_sel(float d, float L, float hL)
{
    const float s = fabsf(d) > hL ? __builtin_copysignf(L, d) : 0.0f;
    return d - s;
}

// Force on i from j (both float4 positions); returns 0 beyond the cutoff.
__device__ __forceinline__ void pair_force(float4 pi, float4 pj, const LJParams &p, float &fx, float &fy, float &fz)
{
    float dx = pi.x - pj.x;
    float dy = pi.y - pj.y;
    float dz = pi.z - pj.z;
    if (p.pbc) {
        dx = min_image_sel(dx, p.L, p.hL);
        dy = min_image_sel(dy, p.L, p.hL);
        dz = min_image_sel(dz, p.L, p.hL);
    }
    float rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    if (rd < p.rc2) {
        const float rrd = exact_rcp(rd);
        const float rrd2 = rrd * rrd;
        const float rrd4 = rrd2 * rrd2;
        const float rrd6 = rrd2 * rrd4;
        const float rrd7 = rrd6 * rrd;
        const float F = 48.0f * (rrd7 - 0.5f * rrd4);
        fx = dx * F;
        fy = dy * F;
        fz = dz * F;
    } else {
        fx = fy = fz = 0.0f;
    }
}

// Potential of one pair (4(r^-12 - r^-6)), 0 beyond the cutoff.
__device__ __forceinline__ float pair_phi(float4 pi, float4 pj, const LJParams &p)
{
    float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    if (p.pbc) {
        dx = min_image_sel(dx, p.L, p.hL);
        dy = min_image_sel(dy, p.L, p.hL);
        dz = min_image_sel(dz, p.L, p.hL);
    }
    float rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    if (!(rd < p.rc2)) return 0.0f;
    const float rrd = exact_rcp(rd);
    const float rrd2 = rrd * rrd;
    const float rrd3 = rrd2 * rrd;
    const float rrd4 = rrd2 * rrd2;
    const float rrd6 = rrd2 * rrd4;
    return 4.0f * (rrd6 - rrd3);
}

// Epilogue modes of the CSR force kernel.
enum : int {
    kForceOnly = 0,   // write f
    kKick = 1,        // f; v += hdt*f
    kKickDrift = 2,   // f; v += hdt*f; v += hdt*f; x_out = wrap(x + dt*v)
};

// Tree-order variant: rows handled by groups of LPR lanes that stride the
// row and combine partial sums with butterflies.  Its sums round in a
// different order from the serial scatter, so its results agree with the
// oracle to FP32 tolerance rather than bit for bit.  A 256-thread block holds
// 256/LPR rows; rows are dealt to blocks XCD-contiguously.
template <int LPR>
__global__ __launch_bounds__(256) void lj_csr_tree(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                              const int *__restrict__ cols, int N, LJParams p, float4 *__restrict__ force,
                                              float4 *__restrict__ vel, float4 *__restrict__ pos_out, int mode)
{
    constexpr int RPB = 256 / LPR;
    const int blk = pat::xcd_swizzle(blockIdx.x, gridDim.x);
    const int row = blk * RPB + threadIdx.x / LPR;
    const int l = threadIdx.x % LPR;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (row < N) {
        pi = pos[row];
        if (p.pbc == 2) {  // batched systems: each node carries its box side in w
            p.L = pi.w;
            p.hL = 0.5f * pi.w;
        }
        const int b = rowptr[row], e = rowptr[row + 1];
        for (int k = b + l; k < e; k += LPR) {
            const float4 pj = pos[cols[k]];
            float ax, ay, az;
            pair_force(pi, pj, p, ax, ay, az);
            fx = fx + ax;
            fy = fy + ay;
            fz = fz + az;
        }
    }
    // butterfly reduction inside the LPR-lane group
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
        fx = fx + __shfl_xor(fx, o, LPR);
        fy = fy + __shfl_xor(fy, o, LPR);
        fz = fz + __shfl_xor(fz, o, LPR);
    }
    if (row < N && l == 0) {
        force[row] = make_float4(fx, fy, fz, 0.0f);
        if (mode != kForceOnly) {
            float4 v = vel[row];
            v.x = v.x + p.hdt * fx;
            v.y = v.y + p.hdt * fy;
            v.z = v.z + p.hdt * fz;
            if (mode == kKickDrift) {
                v.x = v.x + p.hdt * fx;
                v.y = v.y + p.hdt * fy;
                v.z = v.z + p.hdt * fz;
                float4 x;
                x.x = pi.x + p.dt * v.x;
                x.y = pi.y + p.dt * v.y;
                x.z = pi.z + p.dt * v.z;
                x.w = pi.w;
                if (p.pbc) {
                    x.x = pat::wrap_coord(x.x, p.L);
                    x.y = pat::wrap_coord(x.y, p.L);
                    x.z = pat::wrap_coord(x.z, p.L);
                }
                pos_out[row] = x;
            }
            vel[row] = v;
        }
    }
}

// Shared epilogue: force write and the fused velocity-Verlet update.
__device__ __forceinline__ void lj_epilogue(int row, float4 pi, float fx, float fy, float fz, const LJParams &p,
                                            float4 *__restrict__ force, float4 *__restrict__ vel,
                                            float4 *__restrict__ pos_out, int mode)
{
    force[row] = make_float4(fx, fy, fz, 0.0f);
    if (mode == kForceOnly) return;
    float4 v = vel[row];
    v.x = v.x + p.hdt * fx;
    v.y = v.y + p.hdt * fy;
    v.z = v.z + p.hdt * fz;
    if (mode == kKickDrift) {
        v.x = v.x + p.hdt * fx;
        v.y = v.y + p.hdt * fy;
        v.z = v.z + p.hdt * fz;
        float4 x;
        x.x = pi.x + p.dt * v.x;
        x.y = pi.y + p.dt * v.y;
        x.z = pi.z + p.dt * v.z;
        x.w = pi.w;
        if (p.pbc) {
            x.x = pat::wrap_coord(x.x, p.L);
            x.y = pat::wrap_coord(x.y, p.L);
            x.z = pat::wrap_coord(x.z, p.L);
        }
        pos_out[row] = x;
    }
    vel[row] = v;
}

// In-order row sums, one lane per row: the lane walks its row in CSR order
// (edge order), adding each pair force as the serial scatter does, with the
// next four neighbours' loads issued ahead.  Bit-identical to the oracle.
__global__ __launch_bounds__(256) void lj_csr_seq(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                                  const int *__restrict__ cols, int N, LJParams p,
                                                  float4 *__restrict__ force, float4 *__restrict__ vel,
                                                  float4 *__restrict__ pos_out, int mode)
{
    const int blk = pat::xcd_swizzle(blockIdx.x, gridDim.x);
    const int row = blk * 256 + threadIdx.x;
    if (row >= N) return;
    const float4 pi = pos[row];
    if (p.pbc == 2) {
        p.L = pi.w;
        p.hL = 0.5f * pi.w;
    }
    const int b = rowptr[row], e = rowptr[row + 1];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    int k = b;
    for (; k + 4 <= e; k += 4) {
        const int j0 = cols[k], j1 = cols[k + 1], j2 = cols[k + 2], j3 = cols[k + 3];
        const float4 q0 = pos[j0], q1 = pos[j1], q2 = pos[j2], q3 = pos[j3];
        float ax, ay, az;
        pair_force(pi, q0, p, ax, ay, az);
        fx = fx + ax; fy = fy + ay; fz = fz + az;
        pair_force(pi, q1, p, ax, ay, az);
        fx = fx + ax; fy = fy + ay; fz = fz + az;
        pair_force(pi, q2, p, ax, ay, az);
        fx = fx + ax; fy = fy + ay; fz = fz + az;
        pair_force(pi, q3, p, ax, ay, az);
        fx = fx + ax; fy = fy + ay; fz = fz + az;
    }
    for (; k < e; k++) {
        float ax, ay, az;
        pair_force(pi, pos[cols[k]], p, ax, ay, az);
        fx = fx + ax; fy = fy + ay; fz = fz + az;
    }
    lj_epilogue(row, pi, fx, fy, fz, p, force, vel, pos_out, mode);
}

// In-order row sums with LPR lanes per row: each chunk of LPR consecutive
// entries is evaluated in parallel (one pair per lane, coalesced column
// reads), then every lane of the group folds the chunk's values into the
// running sum in entry order through cross-lane reads.  Same sums, bit for
// bit, as lj_csr_seq.
template <int LPR>
__global__ __launch_bounds__(256) void lj_csr_ord(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                                  const int *__restrict__ cols, int N, LJParams p,
                                                  float4 *__restrict__ force, float4 *__restrict__ vel,
                                                  float4 *__restrict__ pos_out, int mode)
{
    constexpr int RPB = 256 / LPR;
    const int blk = pat::xcd_swizzle(blockIdx.x, gridDim.x);
    const int row = blk * RPB + threadIdx.x / LPR;
    const int l = threadIdx.x % LPR;
    if (row >= N) return;  // whole groups leave together
    const float4 pi = pos[row];
    if (p.pbc == 2) {
        p.L = pi.w;
        p.hL = 0.5f * pi.w;
    }
    const int b = rowptr[row], e = rowptr[row + 1];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (int k0 = b; k0 < e; k0 += LPR) {
        const int k = k0 + l;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (k < e) pair_force(pi, pos[cols[k]], p, ax, ay, az);
        const int n = min(LPR, e - k0);
#pragma unroll
        for (int t = 0; t < LPR; t++) {
            const float bx = __shfl(ax, t, LPR), by = __shfl(ay, t, LPR), bz = __shfl(az, t, LPR);
            if (t < n) {
                fx = fx + bx;
                fy = fy + by;
                fz = fz + bz;
            }
        }
    }
    if (l == 0) lj_epilogue(row, pi, fx, fy, fz, p, force, vel, pos_out, mode);
}

// ------------------------------------------------------------ reciprocal check
//
// Exhaustive comparison, over every FP32 value in [lo, hi] (bit patterns),
// of the IEEE quotient 1.0f / b with v_rcp_f32 refined by one and by two
// Newton steps (e = fma(-b, y, 1); y = fma(e, y, y)).  A refined form may
// stand in for the division only where this reports no difference.
__device__ __forceinline__ float rcp_newton(float b, int steps)
{
    float y = __builtin_amdgcn_rcpf(b);
    for (int i = 0; i < steps; i++) {
        const float e = __builtin_fmaf(-b, y, 1.0f);
        y = __builtin_fmaf(e, y, y);
    }
    return y;
}

__global__ __launch_bounds__(256) void rcp_check(uint32_t lo, uint32_t hi, unsigned long long *__restrict__ bad)
{
    unsigned long long b1 = 0, b2 = 0;
    for (uint64_t i = (uint64_t)lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= hi;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float b = __uint_as_float((uint32_t)i);
        const uint32_t ref = __float_as_uint(1.0f / b);
        b1 += __float_as_uint(rcp_newton(b, 1)) != ref;
        b2 += __float_as_uint(rcp_newton(b, 2)) != ref;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        b1 += __shfl_xor(b1, o, kWave);
        b2 += __shfl_xor(b2, o, kWave);
    }
    if (threadIdx.x % kWave == 0) {
        if (b1) atomicAdd(&bad[0], b1);
        if (b2) atomicAdd(&bad[1], b2);
    }
}

// ------------------------------------------------------------ sliced layout
//
// SELL-64 (sliced ELLPACK with 64-row slices, one row per wavefront lane):
// slice s covers rows 64s .. 64s+63 and is width_s = max degree in the slice
// entries wide; entry k of row 64s+l sits at sptr[s] + 64k + l, so the k-th
// column read of a wavefront is one contiguous 256-byte load.  Padding
// entries repeat the row's own index (rows past N repeat node N-1) and are
// masked out of the sum.

__global__ void sell_width(const int *__restrict__ rowptr, int N, int nsl, int *__restrict__ width)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsl) return;
    const int r0 = s * kWave, r1 = min(r0 + kWave, N);
    int w = 0;
    for (int r = r0; r < r1; r++) w = max(w, rowptr[r + 1] - rowptr[r]);
    width[s] = w;
}

// pad < 0: padding repeats the row's own index; otherwise it is 'pad' (a
// sentinel node placed beyond every cutoff).
__global__ __launch_bounds__(256) void sell_fill(const int *__restrict__ rowptr, const int *__restrict__ cols, int N,
                                                 int nsl, const int *__restrict__ sptr, int *__restrict__ scol, int pad)
{
    const int s = blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nsl) return;
    const int row = s * kWave + lane;
    const int b = row < N ? rowptr[row] : 0, e = row < N ? rowptr[row + 1] : 0;
    const int fill = pad < 0 ? min(row, N - 1) : pad;
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    for (int k = 0; k < w; k++) scol[base + k * kWave + lane] = b + k < e ? cols[b + k] : fill;
}

// Neighbour position j.  SENT kernels address with a 32-bit byte offset
// (the handle keeps fewer than 2^28 nodes), the others with 64-bit indices.
template <bool SENT>
__device__ __forceinline__ float4 gather_pos(const float4 *__restrict__ pos, int j)
{
    if (SENT) return *(const float4 *)((const char *)pos + ((uint32_t)j << 4));
    return pos[j];
}

// In-order row sums over the sliced layout, one wavefront per slice and one
// lane per row: the lane adds its row's pair forces in CSR order, U
// entries' column reads and position gathers issued ahead of their
// arithmetic.  Padding entries either repeat the row's own index and are
// masked (SENT = false), or point at a sentinel node N whose position is
// beyond every cutoff (SENT = true): its pair force is exactly +0, and
// adding +0 to a sum that started at +0 leaves every bit unchanged, so no
// mask is needed.  Bit-identical to lj_csr_seq either way.
template <int U, bool SENT>
__global__ __launch_bounds__(256) void lj_sell(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                               const int *__restrict__ sptr, const int *__restrict__ scol, int N,
                                               LJParams p, float4 *__restrict__ force, float4 *__restrict__ vel,
                                               float4 *__restrict__ pos_out, int mode)
{
    const int nsl = (N + kWave - 1) / kWave;
    const int s = pat::xcd_swizzle(blockIdx.x, gridDim.x) * 4 + threadIdx.x / kWave;
    if (s >= nsl) return;  // whole wavefronts
    const int lane = threadIdx.x % kWave;
    const int row = s * kWave + lane;
    const float4 pi = pos[min(row, N - 1)];
    if (p.pbc == 2) {
        p.L = pi.w;
        p.hL = 0.5f * pi.w;
    }
    const int deg = SENT ? 0 : (row < N ? rowptr[row + 1] - rowptr[row] : 0);
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    const int *__restrict__ c = scol + base + lane;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    int k = 0;
    for (; k + U <= w; k += U) {
        int j[U];
        float4 q[U];
#pragma unroll
        for (int u = 0; u < U; u++) j[u] = c[(k + u) * kWave];
#pragma unroll
        for (int u = 0; u < U; u++) q[u] = gather_pos<SENT>(pos, j[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            float ax, ay, az;
            pair_force(pi, q[u], p, ax, ay, az);
            if (SENT || k + u < deg) {
                fx = fx + ax;
                fy = fy + ay;
                fz = fz + az;
            }
        }
    }
    for (; k < w; k++) {
        float ax, ay, az;
        pair_force(pi, gather_pos<SENT>(pos, c[k * kWave]), p, ax, ay, az);
        if (SENT || k < deg) {
            fx = fx + ax;
            fy = fy + ay;
            fz = fz + az;
        }
    }
    if (row < N) lj_epilogue(row, pi, fx, fy, fz, p, force, vel, pos_out, mode);
}

// The same sums, software-pipelined: while a group of U pairs is evaluated,
// the positions of the next group and the column indices of the one after
// are already in flight (two register sets used alternately), so a lane
// always has a group's memory latency to hide behind arithmetic.
// Bit-identical to lj_sell.
template <int U, bool SENT>
__global__ __launch_bounds__(256) void lj_sell_pipe(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                                    const int *__restrict__ sptr, const int *__restrict__ scol, int N,
                                                    LJParams p, float4 *__restrict__ force, float4 *__restrict__ vel,
                                                    float4 *__restrict__ pos_out, int mode)
{
    const int nsl = (N + kWave - 1) / kWave;
    const int s = pat::xcd_swizzle(blockIdx.x, gridDim.x) * 4 + threadIdx.x / kWave;
    if (s >= nsl) return;  // whole wavefronts
    const int lane = threadIdx.x % kWave;
    const int row = s * kWave + lane;
    const int self = min(row, N - 1);
    const float4 pi = pos[self];
    if (p.pbc == 2) {
        p.L = pi.w;
        p.hL = 0.5f * pi.w;
    }
    const int deg = SENT ? 0 : (row < N ? rowptr[row + 1] - rowptr[row] : 0);
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    const int *__restrict__ c = scol + base + lane;
    // column of entry k; past the slice width a valid dummy whose force is
    // masked (or zero, for the sentinel)
    const int dummy = SENT ? N : self;
    auto col_at = [&](int k) { return k < w ? c[k * kWave] : dummy; };
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    // evaluates group k from qc, loads the positions of group k+U (from jx)
    // into qx and the columns of group k+2U into jy
    auto step = [&](int k, const float4 (&qc)[U], const int (&jx)[U], float4 (&qx)[U], int (&jy)[U]) {
#pragma unroll
        for (int u = 0; u < U; u++) qx[u] = gather_pos<SENT>(pos, jx[u]);
#pragma unroll
        for (int u = 0; u < U; u++) jy[u] = col_at(k + 2 * U + u);
#pragma unroll
        for (int u = 0; u < U; u++) {
            float ax, ay, az;
            pair_force(pi, qc[u], p, ax, ay, az);
            if (SENT || k + u < deg) {
                fx = fx + ax;
                fy = fy + ay;
                fz = fz + az;
            }
        }
    };
    int jA[U], jB[U];
    float4 qA[U], qB[U];
#pragma unroll
    for (int u = 0; u < U; u++) jA[u] = col_at(u);
#pragma unroll
    for (int u = 0; u < U; u++) jB[u] = col_at(U + u);
#pragma unroll
    for (int u = 0; u < U; u++) qA[u] = gather_pos<SENT>(pos, jA[u]);
    for (int k = 0; k < w; k += 2 * U) {
        step(k, qA, jB, qB, jA);
        if (k + U < w) step(k + U, qB, jA, qA, jB);
    }
    if (row < N) lj_epilogue(row, pi, fx, fy, fz, p, force, vel, pos_out, mode);
}

// Per-node half potential 0.5*sum_j phi(i, j) over the full CSR row.
__global__ __launch_bounds__(256) void lj_csr_energy(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                                     const int *__restrict__ cols, int N, LJParams p,
                                                     double *__restrict__ node_e)
{
    const int row = blockIdx.x * 4 + threadIdx.x / 64;
    const int l = threadIdx.x % 64;
    double acc = 0.0;
    if (row < N) {
        const float4 pi = pos[row];
        if (p.pbc == 2) {
            p.L = pi.w;
            p.hL = 0.5f * pi.w;
        }
        for (int k = rowptr[row] + l; k < rowptr[row + 1]; k += 64) acc += (double)pair_phi(pi, pos[cols[k]], p);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (row < N && l == 0) node_e[row] = 0.5 * acc;
}

// Plain baseline: one thread per half edge, float atomics into both nodes.
__global__ __launch_bounds__(256) void lj_atomic(const float4 *__restrict__ pos, const int2 *__restrict__ edges,
                                                 int64_t E, LJParams p, float4 *__restrict__ force)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const int2 ij = edges[e];
        float fx, fy, fz;
        pair_force(pos[ij.x], pos[ij.y], p, fx, fy, fz);
        if (fx != 0.0f || fy != 0.0f || fz != 0.0f) {
            float *fi = reinterpret_cast<float *>(force + ij.x);
            float *fj = reinterpret_cast<float *>(force + ij.y);
            atomicAdd(fi + 0, fx);
            atomicAdd(fi + 1, fy);
            atomicAdd(fi + 2, fz);
            atomicAdd(fj + 0, -fx);
            atomicAdd(fj + 1, -fy);
            atomicAdd(fj + 2, -fz);
        }
    }
}

// Copies one system's CSR into the batched CSR, shifting node and entry
// indices by the system's offsets.
__global__ void csr_place(const int *__restrict__ rp, const int *__restrict__ cl, int n, int node_off, int entry_off,
                          int *__restrict__ grp, int *__restrict__ gcl)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= n) grp[node_off + t] = rp[t] + entry_off;
    const int ne = rp[n];
    for (int64_t k = t; k < ne; k += (int64_t)gridDim.x * blockDim.x) gcl[entry_off + k] = cl[k] + node_off;
}

__global__ void fill_zero4(float4 *__restrict__ a, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Node-wise velocity-Verlet pieces for the plain path.
__global__ void vv_kick_drift(float4 *__restrict__ pos, float4 *__restrict__ vel, const float4 *__restrict__ force,
                              int N, LJParams p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 v = vel[i], f = force[i], x = pos[i];
    if (p.pbc == 2) p.L = x.w;
    v.x = v.x + p.hdt * f.x;
    v.y = v.y + p.hdt * f.y;
    v.z = v.z + p.hdt * f.z;
    x.x = x.x + p.dt * v.x;
    x.y = x.y + p.dt * v.y;
    x.z = x.z + p.dt * v.z;
    if (p.pbc) {
        x.x = pat::wrap_coord(x.x, p.L);
        x.y = pat::wrap_coord(x.y, p.L);
        x.z = pat::wrap_coord(x.z, p.L);
    }
    vel[i] = v;
    pos[i] = x;
}

__global__ void vv_kick(float4 *__restrict__ vel, const float4 *__restrict__ force, int N, LJParams p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 v = vel[i], f = force[i];
    v.x = v.x + p.hdt * f.x;
    v.y = v.y + p.hdt * f.y;
    v.z = v.z + p.hdt * f.z;
    vel[i] = v;
}

}  // namespace

namespace pat {

struct LJConfig {
    float box;
    int pbc;
    float rc2;
    float dt;
};

static LJParams make_params(float box, int pbc, float rc2, float dt)
{
    LJParams p;
    p.L = box;
    p.hL = 0.5f * box;
    p.rc2 = rc2;
    p.dt = dt;
    p.hdt = 0.5f * dt;
    p.pbc = pbc;
    return p;
}

int64_t csr_workspace_bytes(int64_t N, int64_t E)
{
    const int64_t nb = (N + 1 + kScanBlock - 1) / kScanBlock;
    const int64_t nb2 = (nb + kScanBlock - 1) / kScanBlock;
    // deg/cursor (N+1), block sums (nb), their scan (nb), second level (nb2 x 2), tmp cols and edge ids (2 x 2E)
    return (int64_t)sizeof(int) * ((N + 1) + 2 * nb + 2 * nb2 + 2 + 4 * E) + 256;
}

// Exclusive scan of n ints (n <= kScanBlock^3) using ws for block sums.
static int scan_exclusive(const int *in, int *out, int64_t n, int *ws, hipStream_t st)
{
    const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    int *sums = ws, *sums_scan = ws + nb;
    hipLaunchKernelGGL(scan_blocks, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, out, n, sums);
    PAT_LAUNCH_CHECK();
    if (nb > 1) {
        const int64_t nb2 = (nb + kScanBlock - 1) / kScanBlock;
        PAT_REQUIRE(nb2 <= kScanBlock, PAT_ERR_INVALID, "scan: %lld elements exceed the two-level scan",
                    (long long)n);
        int *s2 = sums_scan + nb, *s2s = s2 + nb2;
        hipLaunchKernelGGL(scan_blocks, dim3((unsigned)nb2), dim3(kScanBlock), 0, st, sums, sums_scan, nb, s2);
        PAT_LAUNCH_CHECK();
        if (nb2 > 1) {
            hipLaunchKernelGGL(scan_blocks, dim3(1), dim3(kScanBlock), 0, st, s2, s2s, nb2, s2s + nb2);
            PAT_LAUNCH_CHECK();
            hipLaunchKernelGGL(scan_add, dim3((unsigned)nb2), dim3(kScanBlock), 0, st, sums_scan, nb, s2s);
            PAT_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(scan_add, dim3((unsigned)nb), dim3(kScanBlock), 0, st, out, n, sums_scan);
        PAT_LAUNCH_CHECK();
    }
    return PAT_OK;
}

// order 0: rows in edge-list order; 1: rows sorted by neighbour index.
int csr_build(const int32_t *edges, int64_t E, int64_t N, int order, int32_t *rowptr, int32_t *cols, void *work,
              int64_t work_bytes, hipStream_t st)
{
    PAT_REQUIRE(order == 0 || order == 1, PAT_ERR_INVALID, "csr_build: row order %d not in {0, 1}", order);
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - 1 && E >= 0 && 2 * E < (1LL << 31), PAT_ERR_INVALID,
                "csr_build: sizes out of range (N=%lld, E=%lld)", (long long)N, (long long)E);
    PAT_REQUIRE(work_bytes >= csr_workspace_bytes(N, E), PAT_ERR_INVALID, "csr_build: workspace too small");
    const int64_t nb = (N + 1 + kScanBlock - 1) / kScanBlock;
    int *deg = (int *)work;
    int *ws = deg + (N + 1);
    int *tmp = ws + 2 * nb + 2 * ((nb + kScanBlock - 1) / kScanBlock) + 2;
    PAT_HIP(hipMemsetAsync(deg, 0, sizeof(int) * (N + 1), st));
    const unsigned grid = (unsigned)std::min<int64_t>((E + 255) / 256 + 1, 65536);
    if (E > 0) {
        hipLaunchKernelGGL(csr_degree, dim3(grid), dim3(256), 0, st, (const int2 *)edges, E, deg);
        PAT_LAUNCH_CHECK();
    }
    PAT_TRY(scan_exclusive(deg, rowptr, N + 1, ws, st));
    // cursor = rowptr (reuse deg)
    PAT_HIP(hipMemcpyAsync(deg, rowptr, sizeof(int) * (N + 1), hipMemcpyDeviceToDevice, st));
    if (E > 0) {
        int *tmpe = tmp + 2 * E;
        hipLaunchKernelGGL(csr_scatter, dim3(grid), dim3(256), 0, st, (const int2 *)edges, E, deg, tmp, tmpe);
        PAT_LAUNCH_CHECK();
        const unsigned g2 = (unsigned)((N * kWave + 255) / 256);
        if (order == 0)
            hipLaunchKernelGGL(csr_row_sort_by_edge, dim3(g2), dim3(256), 0, st, rowptr, N, tmp, tmpe, cols);
        else
            hipLaunchKernelGGL(csr_row_sort, dim3(g2), dim3(256), 0, st, rowptr, N, tmp, cols);
        PAT_LAUNCH_CHECK();
    }
    return PAT_OK;
}

using CsrKernel = void (*)(const float4 *, const int *, const int *, int, LJParams, float4 *, float4 *, float4 *, int);

// lpr: 1 = one lane per row (in order); 4..32 = lanes per row (in order);
// tree != 0 selects the butterfly variant (4..64 lanes).
static CsrKernel csr_kernel_for(int lpr, int tree)
{
    if (tree) {
        switch (lpr) {
        case 4: return lj_csr_tree<4>;
        case 8: return lj_csr_tree<8>;
        case 16: return lj_csr_tree<16>;
        case 32: return lj_csr_tree<32>;
        case 64: return lj_csr_tree<64>;
        default: return nullptr;
        }
    }
    switch (lpr) {
    case 1: return lj_csr_seq;
    case 4: return lj_csr_ord<4>;
    case 8: return lj_csr_ord<8>;
    case 16: return lj_csr_ord<16>;
    case 32: return lj_csr_ord<32>;
    default: return nullptr;
    }
}

int lj_force_csr(const float4 *pos, const int32_t *rowptr, const int32_t *cols, int64_t N, float box, int pbc,
                 float rc2, float dt, int lpr, int mode, float4 *force, float4 *vel, float4 *pos_out, hipStream_t st)
{
    // negative lpr selects the tree-order variant
    const int tree = lpr < 0;
    if (tree) lpr = -lpr;
    CsrKernel k = csr_kernel_for(lpr, tree);
    PAT_REQUIRE(k, PAT_ERR_UNSUPPORTED, "lj_force_csr: lanes per row %d not available (%s)", lpr,
                tree ? "tree: 4..64" : "in order: 1, 4, 8, 16, 32");
    PAT_REQUIRE(N > 0 && N < (1LL << 31), PAT_ERR_INVALID, "lj_force_csr: bad node count");
    const int rpb = 256 / lpr;
    const int64_t blocks = (N + rpb - 1) / rpb;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(256), 0, st, pos, (const int *)rowptr, (const int *)cols, (int)N,
                       make_params(box, pbc, rc2, dt), force, vel, pos_out, mode);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// Sliced layout of a CSR: widths on the device, slice offsets scanned on the
// host in 64-bit (the padded total must stay below 2^31), then the fill.
// width_h / sptr_h: host scratch of nsl and nsl+1 ints; *scol grows to the
// padded total (capacity *cap).  Refuses (PAT_ERR_UNSUPPORTED, nothing
// changed) when padding would exceed max_pad times the entry count plus one
// slice of 64 entries per slice.
int sell_build(const int32_t *rowptr, const int32_t *cols, int64_t N, double max_pad, int32_t *width_d,
               int32_t *sptr, int32_t **scol, int64_t *cap, int64_t *total, int pad, hipStream_t st)
{
    const int64_t nsl = (N + kWave - 1) / kWave;
    hipLaunchKernelGGL(sell_width, dim3((unsigned)((nsl + 255) / 256)), dim3(256), 0, st, (const int *)rowptr, (int)N,
                       (int)nsl, (int *)width_d);
    PAT_LAUNCH_CHECK();
    std::vector<int32_t> w((size_t)nsl);
    std::vector<int32_t> sp((size_t)nsl + 1);
    int32_t nnz = 0;
    PAT_HIP(hipMemcpyAsync(w.data(), width_d, sizeof(int32_t) * nsl, hipMemcpyDeviceToHost, st));
    PAT_HIP(hipMemcpyAsync(&nnz, rowptr + N, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    int64_t acc = 0;
    for (int64_t s = 0; s < nsl; s++) {
        sp[s] = (int32_t)acc;
        acc += (int64_t)kWave * w[s];
        PAT_REQUIRE(acc < (1LL << 31), PAT_ERR_UNSUPPORTED, "sell_build: padded layout exceeds 2^31 entries");
    }
    sp[nsl] = (int32_t)acc;
    PAT_REQUIRE((double)acc <= max_pad * (double)nnz + (double)kWave * nsl, PAT_ERR_UNSUPPORTED,
                "sell_build: padding %lld for %d entries exceeds the limit", (long long)acc, nnz);
    if (acc > *cap) {
        const int64_t want = std::min<int64_t>((int64_t)(acc * 1.25) + 4096, (1LL << 31) - 1);
        int32_t *nc = nullptr;
        if (pat::dev_malloc((void **)&nc, sizeof(int32_t) * (size_t)want) != hipSuccess) {
            (void)hipGetLastError();
            pat::set_error("sell_build: hipMalloc of %lld entries failed", (long long)want);
            return PAT_ERR_NOMEM;
        }
        if (*scol) (void)pat::dev_free(*scol);
        *scol = nc;
        *cap = want;
    }
    PAT_HIP(hipMemcpyAsync(sptr, sp.data(), sizeof(int32_t) * (nsl + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(sell_fill, dim3((unsigned)((nsl + 3) / 4)), dim3(256), 0, st, (const int *)rowptr,
                       (const int *)cols, (int)N, (int)nsl, (const int *)sptr, (int *)*scol, pad);
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipStreamSynchronize(st));  // sp leaves scope
    *total = acc;
    return PAT_OK;
}

// Sliced layout into caller buffers (sptr: N/64+1 ints, also the width
// scratch; scol: cap entries); *total = padded entries, or the needed
// capacity with PAT_ERR_INVALID when cap is short.
int sell_build_into(const int32_t *rowptr, const int32_t *cols, int64_t N, int32_t *sptr, int32_t *scol, int64_t cap,
                    int64_t *total, hipStream_t st)
{
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - kWave, PAT_ERR_INVALID, "sell_build: bad node count");
    const int64_t nsl = (N + kWave - 1) / kWave;
    hipLaunchKernelGGL(sell_width, dim3((unsigned)((nsl + 255) / 256)), dim3(256), 0, st, (const int *)rowptr, (int)N,
                       (int)nsl, (int *)sptr);
    PAT_LAUNCH_CHECK();
    std::vector<int32_t> sp((size_t)nsl + 1);
    PAT_HIP(hipMemcpyAsync(sp.data(), sptr, sizeof(int32_t) * nsl, hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    int64_t acc = 0;
    for (int64_t i = 0; i < nsl; i++) {
        const int64_t wi = sp[i];
        sp[i] = (int32_t)acc;
        acc += (int64_t)kWave * wi;
        PAT_REQUIRE(acc < (1LL << 31), PAT_ERR_UNSUPPORTED, "sell_build: padded layout exceeds 2^31 entries");
    }
    sp[nsl] = (int32_t)acc;
    *total = acc;
    PAT_REQUIRE(acc <= cap, PAT_ERR_INVALID, "sell_build: %lld entries needed, capacity %lld", (long long)acc,
                (long long)cap);
    PAT_HIP(hipMemcpyAsync(sptr, sp.data(), sizeof(int32_t) * (nsl + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(sell_fill, dim3((unsigned)((nsl + 3) / 4)), dim3(256), 0, st, (const int *)rowptr,
                       (const int *)cols, (int)N, (int)nsl, (const int *)sptr, (int *)scol, -1);
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipStreamSynchronize(st));  // sp leaves scope
    return PAT_OK;
}

// kind: 0 = lj_sell<4>, 1 = lj_sell_pipe<4>, 2 = lj_sell_pipe<8> over a
// layout whose padding repeats the row; 3, 4, 5 = the same over a layout
// padded with the sentinel node N (pos then holds N+1 entries, the last
// beyond every cutoff, and N < 2^28).  Same sums for all.
int lj_force_sell(const float4 *pos, const int32_t *rowptr, const int32_t *sptr, const int32_t *scol, int64_t N,
                  float box, int pbc, float rc2, float dt, int mode, float4 *force, float4 *vel, float4 *pos_out,
                  int kind, hipStream_t st)
{
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - kWave, PAT_ERR_INVALID, "lj_force_sell: bad node count");
    using K = void (*)(const float4 *, const int *, const int *, const int *, int, LJParams, float4 *, float4 *,
                       float4 *, int);
    const K ks[] = {lj_sell<4, false>, lj_sell_pipe<4, false>, lj_sell_pipe<8, false>,
                    lj_sell<4, true>,  lj_sell_pipe<4, true>,  lj_sell_pipe<8, true>};
    PAT_REQUIRE(kind >= 0 && kind < 6, PAT_ERR_INVALID, "lj_force_sell: kind %d", kind);
    PAT_REQUIRE(kind < 3 || N < (1LL << 28) - 1, PAT_ERR_INVALID, "lj_force_sell: sentinel kinds need N < 2^28 - 1");
    const int64_t nsl = (N + kWave - 1) / kWave;
    hipLaunchKernelGGL(ks[kind], dim3((unsigned)((nsl + 3) / 4)), dim3(256), 0, st, pos, (const int *)rowptr,
                       (const int *)sptr, (const int *)scol, (int)N, make_params(box, pbc, rc2, dt), force, vel,
                       pos_out, mode);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int rcp_selftest(uint32_t lo, uint32_t hi, int64_t *bad_one, int64_t *bad_two)
{
    unsigned long long *d = nullptr;
    PAT_HIP(hipMalloc((void **)&d, 16));
    unsigned long long h[2] = {0, 0};
    int rc = PAT_OK;
    if (hipMemset(d, 0, 16) != hipSuccess) rc = PAT_ERR_HIP;
    if (rc == PAT_OK) {
        hipLaunchKernelGGL(rcp_check, dim3(65536), dim3(256), 0, 0, lo, hi, d);
        if (hipGetLastError() != hipSuccess || hipMemcpy(h, d, 16, hipMemcpyDeviceToHost) != hipSuccess)
            rc = PAT_ERR_HIP;
    }
    (void)hipFree(d);
    if (rc != PAT_OK) {
        pat::set_error("rcp_selftest: HIP call failed");
        return rc;
    }
    *bad_one = (int64_t)h[0];
    *bad_two = (int64_t)h[1];
    return PAT_OK;
}

int csr_place_dev(const int32_t *rp, const int32_t *cl, int64_t n, int64_t node_off, int64_t entry_off,
                  int32_t *grp, int32_t *gcl, hipStream_t st)
{
    const unsigned grid = (unsigned)std::min<int64_t>((n + 256) / 256 + 1, 4096);
    hipLaunchKernelGGL(csr_place, dim3(grid), dim3(256), 0, st, (const int *)rp, (const int *)cl, (int)n, (int)node_off,
                       (int)entry_off, (int *)grp, (int *)gcl);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int lj_energy_csr(const float4 *pos, const int32_t *rowptr, const int32_t *cols, int64_t N, float box, int pbc,
                  float rc2, double *node_e, hipStream_t st)
{
    hipLaunchKernelGGL(lj_csr_energy, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, pos, (const int *)rowptr,
                       (const int *)cols, (int)N, make_params(box, pbc, rc2, 0.f), node_e);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int lj_force_atomic(const float4 *pos, const int32_t *edges, int64_t E, int64_t N, float box, int pbc, float rc2,
                    float4 *force, hipStream_t st)
{
    hipLaunchKernelGGL(fill_zero4, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, force, (int)N);
    PAT_LAUNCH_CHECK();
    if (E > 0) {
        const unsigned grid = (unsigned)std::min<int64_t>((E + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(lj_atomic, dim3(grid), dim3(256), 0, st, pos, (const int2 *)edges, E,
                           make_params(box, pbc, rc2, 0.f), force);
        PAT_LAUNCH_CHECK();
    }
    return PAT_OK;
}

int vv_kick_drift_dev(float4 *pos, float4 *vel, const float4 *force, int64_t N, float box, int pbc, float dt,
                      hipStream_t st)
{
    hipLaunchKernelGGL(vv_kick_drift, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, pos, vel, force, (int)N,
                       make_params(box, pbc, 0.f, dt));
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int vv_kick_dev(float4 *vel, const float4 *force, int64_t N, float dt, hipStream_t st)
{
    hipLaunchKernelGGL(vv_kick, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, vel, force, (int)N,
                       make_params(1.f, 0, 0.f, dt));
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

}  // namespace pat

namespace pat {

// Device buffers of the binned neighbour build, kept between rebuilds.
struct NbWork {
    int64_t N = 0, ncell = 0, cap = 0;  // cap: capacity of cols/coltmp
    int nb = 0;
    int *cell = nullptr, *cstart = nullptr, *cursor = nullptr, *atmp = nullptr, *catoms = nullptr, *deg = nullptr,
        *ws = nullptr, *coltmp = nullptr;
};

static void nb_free(NbWork &w)
{
    for (int *p : {w.cell, w.cstart, w.cursor, w.atmp, w.catoms, w.deg, w.ws, w.coltmp})
        if (p) (void)pat::dev_free(p);
    w = NbWork{};
}

NbWork *nb_create() { return new NbWork(); }
void nb_destroy(NbWork *w)
{
    if (w) {
        nb_free(*w);
        delete w;
    }
}

static int nb_alloc(int **p, int64_t n)
{
    if (pat::dev_malloc((void **)p, sizeof(int) * (size_t)(n > 0 ? n : 1)) != hipSuccess) {
        (void)hipGetLastError();
        pat::set_error("neighbour build: hipMalloc of %lld ints failed", (long long)n);
        return PAT_ERR_NOMEM;
    }
    return PAT_OK;
}

// Rebuilds the full, row-sorted neighbour CSR of N float4 positions in a
// periodic cube of side L for pairs closer than cut.  *cols (capacity
// *cap entries) is reallocated when too small.  Synchronises the stream once
// (to size the column array).
int neighbor_build(const float4 *pos, int64_t N, double L, double cut, NbWork *w, int32_t *rowptr, int32_t **cols,
                   int64_t *cap, int64_t *total, hipStream_t st)
{
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - 1 && L > 0 && cut > 0, PAT_ERR_INVALID, "neighbour build: bad arguments");
    const int nb = (int)std::floor(L / (cut * 1.0001));
    PAT_REQUIRE(nb >= 3, PAT_ERR_UNSUPPORTED, "neighbour build: box %.3f holds fewer than 3 bins of %.3f", L, cut);
    PAT_REQUIRE((int64_t)nb * nb * nb < (1LL << 30), PAT_ERR_INVALID, "neighbour build: too many bins");
    const int64_t ncell = (int64_t)nb * nb * nb;
    if (w->N != N || w->ncell != ncell) {
        const int64_t cap_keep = w->cap;
        int *coltmp = w->coltmp;
        w->coltmp = nullptr;
        nb_free(*w);
        w->coltmp = coltmp;
        w->cap = cap_keep;
        const int64_t m = std::max<int64_t>(N, ncell) + 1;
        const int64_t scan_ws = 2 * ((m + kScanBlock - 1) / kScanBlock) + 2 * kScanBlock + 4;
        PAT_TRY(nb_alloc(&w->cell, N));
        PAT_TRY(nb_alloc(&w->cstart, ncell + 1));
        PAT_TRY(nb_alloc(&w->cursor, ncell + 1));
        PAT_TRY(nb_alloc(&w->atmp, N));
        PAT_TRY(nb_alloc(&w->catoms, N));
        PAT_TRY(nb_alloc(&w->deg, N + 1));
        PAT_TRY(nb_alloc(&w->ws, scan_ws));
        w->N = N;
        w->ncell = ncell;
        w->nb = nb;
    }
    BinGeom g;
    g.L = (float)L;
    g.hL = 0.5f * g.L;
    g.cut2 = (float)(cut * cut);
    g.nb = nb;
    g.inv = (float)nb / g.L;
    const unsigned gN = (unsigned)((N + 255) / 256);
    // bin the atoms and sort each bin's list by atom index
    PAT_HIP(hipMemsetAsync(w->cursor, 0, sizeof(int) * (ncell + 1), st));
    hipLaunchKernelGGL(nb_bin_atoms, dim3(gN), dim3(256), 0, st, pos, (int)N, g, w->cell, w->cursor);
    PAT_LAUNCH_CHECK();
    PAT_TRY(scan_exclusive(w->cursor, w->cstart, ncell + 1, w->ws, st));
    PAT_HIP(hipMemcpyAsync(w->cursor, w->cstart, sizeof(int) * (ncell + 1), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(nb_scatter_atoms, dim3(gN), dim3(256), 0, st, w->cell, (int)N, w->cursor, w->atmp);
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(csr_row_sort, dim3((unsigned)((ncell * kWave + 255) / 256)), dim3(256), 0, st, w->cstart, ncell,
                       w->atmp, w->catoms);
    PAT_LAUNCH_CHECK();
    // row lengths, then row pointers
    hipLaunchKernelGGL(nb_rows, dim3(gN), dim3(256), 0, st, pos, (int)N, g, w->cell, w->cstart, w->catoms, w->deg,
                       (const int *)nullptr, (int *)nullptr, 0);
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipMemsetAsync(w->deg + N, 0, sizeof(int), st));
    PAT_TRY(scan_exclusive(w->deg, rowptr, N + 1, w->ws, st));
    int tot = 0;
    PAT_HIP(hipMemcpyAsync(&tot, rowptr + N, sizeof(int), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    PAT_REQUIRE(tot >= 0, PAT_ERR_INVALID, "neighbour build: pair count overflows int32");
    if (tot > *cap || tot > w->cap) {
        const int64_t want = (int64_t)(tot * 1.25) + 1024;
        if (*cols) (void)pat::dev_free(*cols);
        if (w->coltmp) (void)pat::dev_free(w->coltmp);
        *cols = nullptr;
        w->coltmp = nullptr;
        PAT_TRY(nb_alloc(cols, want));
        PAT_TRY(nb_alloc(&w->coltmp, want));
        *cap = want;
        w->cap = want;
    }
    hipLaunchKernelGGL(nb_rows, dim3(gN), dim3(256), 0, st, pos, (int)N, g, w->cell, w->cstart, w->catoms, w->deg,
                       (const int *)rowptr, w->coltmp, 1);
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(csr_row_sort, dim3((unsigned)((N * kWave + 255) / 256)), dim3(256), 0, st, (const int *)rowptr,
                       N, w->coltmp, *cols);
    PAT_LAUNCH_CHECK();
    *total = tot;
    return PAT_OK;
}

}  // namespace pat
