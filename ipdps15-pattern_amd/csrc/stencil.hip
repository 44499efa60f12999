// Stencil engine for gfx950: the 7-point heat sweep of the heat3D workload.
//
// Replaces the reference's compute_cuda_internal (kmeans/lib/compute_cuda.cu:
// 30-100; parameterised in heat_3D/lib/compute_cuda.cu:30-105), where 14
// persistent blocks stride over 32^3 tiles, decode each cell with three
// div/mods and call the functor through a __constant__ function-pointer
// table.  Here the heat functor (heat_3D/stencil.h:94-117 shape; update
// u' = u + alpha*(s - 6u), s = ((((x- + x+) + y-) + y+) + z-) + z+) is bound
// at compile time and swept by LDS-staged 2.5D tiles:
//   * a workgroup owns a BX x (BY*RY) tile of the x-y plane and marches a
//     chunk of zc planes; each thread keeps its column's z-1, z, z+1 values
//     in registers (RY rows per thread) so every input cell is read from HBM
//     once per sweep;
//   * the current plane (plus a one-cell x/y halo) is staged in LDS so the
//     four in-plane neighbours come from LDS;
//   * rows of 64 consecutive x cells per wavefront give fully coalesced loads
//     and stores;
//   * tiles are ordered so that consecutive tile indices land on the same XCD
//     (the dispatcher deals workgroups round-robin over the 8 XCDs), keeping
//     neighbouring tiles' shared halo rows in one XCD's L2.
// Grid layout is the reference's (GET_FLOAT3, heat_3D/lib/macro.h:70):
// halo-inclusive (nx+2)(ny+2)(nz+2), x fastest.
// Compiled with -ffp-contract=off: each cell's arithmetic rounds exactly as
// written, so results are bit-identical to a plain C sweep of the same rule.
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "pat_common.hpp"
#include "stencil_dev.hpp"

namespace {

constexpr int kXcds = 8;

// Maps a hardware workgroup id to a tile id so that the tiles one XCD gets
// are contiguous in tile order.
__device__ __forceinline__ int xcd_swizzle(int bid, int nblocks)
{
    const int per = nblocks / kXcds;  // full rounds of kXcds
    if (bid >= per * kXcds) return bid;  // ragged tail keeps its id
    return (bid % kXcds) * per + bid / kXcds;
}

template <int BX, int BY, int RY>
__global__ __launch_bounds__(BX *BY) void heat_lds25d(const float *__restrict__ in, float *__restrict__ out, int nx,
                                                      int ny, int nz, int zc, float alpha, int tiles_x, int tiles_y,
                                                      int tiles_z)
{
    constexpr int TH = BY * RY;
    __shared__ float plane[TH + 2][BX + 2];

    const int nblocks = tiles_x * tiles_y * tiles_z;
    const int t = xcd_swizzle(blockIdx.x, nblocks);
    const int bx = t % tiles_x;
    const int by = (t / tiles_x) % tiles_y;
    const int bz = t / (tiles_x * tiles_y);

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int sx = nx + 2;
    const int64_t sxy = (int64_t)sx * (ny + 2);

    const int x = bx * BX + tx + 1;
    const bool xin = x <= nx;
    const int xl = xin ? x : nx + 1;  // out-of-range lanes read the halo cell

    const int z0 = bz * zc + 1;
    const int z1 = min(z0 + zc, nz + 1);

    static_assert(BX >= 2 && BY >= 2, "edge lanes stage one halo side each");
    bool act[RY];
    int64_t id[RY];  // this thread's rows in the plane being swept
    // Four plane registers per row (planes z-1 .. z+2) and two sets of halo
    // registers (planes z and z+1) for the cells this thread stages: its
    // rows' left (tx == 0) or right (tx == BX-1) neighbours, the row above
    // (ty == 0) or below (ty == BY-1) the tile.  Each load is issued two
    // planes before its use, and the z loop is unrolled by four so that the
    // roles rotate over the registers and none is copied into another (a
    // copy would wait for its load at once, exposing the memory latency on
    // every plane).
    float P[4][RY], H[2][RY], G[2] = {0.0f, 0.0f};
    // Every lane loads (no branch, so no register merges): lanes off the
    // edges read their own cell again, a cache hit.
    const int64_t hx_off = tx == 0 ? -1 : (tx == BX - 1 ? 1 : 0);
    const int hy_row = ty == 0 ? 0 : RY - 1;
    const int64_t hy_off = ty == 0 ? -(int64_t)sx : (ty == BY - 1 ? (int64_t)sx : 0);
    // halo cells k planes past the rows' current plane
    auto load_halo = [&](int64_t k, float (&h)[RY], float &g) {
#pragma unroll
        for (int r = 0; r < RY; r++) h[r] = in[id[r] + hx_off + k * sxy];
        g = in[id[hy_row] + hy_off + k * sxy];
    };
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const int y = by * TH + ty * RY + r + 1;
        act[r] = xin && y <= ny;
        id[r] = xl + (int64_t)sx * ((y <= ny ? y : ny + 1) + (int64_t)(ny + 2) * z0);
        P[0][r] = in[id[r] - sxy];
        P[1][r] = in[id[r]];
        P[2][r] = in[id[r] + sxy];
        P[3][r] = 0.0f;
        H[0][r] = 0.0f;
        H[1][r] = 0.0f;
    }
    load_halo(0, H[0], G[0]);
    // planes past the chunk's last read (z1 for the rows, z1-1 for the
    // halo) read that plane again
    const int two = z0 + 1 < z1 ? 1 : 0;
#pragma unroll
    for (int r = 0; r < RY; r++) P[3][r] = in[id[r] + (1 + two) * sxy];
    load_halo(two, H[1], G[1]);

    // Plane z: m, c, a hold planes z-1, z, z+1 and h, g plane z's halo
    // cells.  Afterwards m takes plane z+3 and h, g plane z+2's halo, both
    // first read two steps later.
    auto step = [&](int z, float (&m)[RY], const float (&c)[RY], const float (&a)[RY], float (&h)[RY], float &g) {
#pragma unroll
        for (int r = 0; r < RY; r++) plane[ty * RY + r + 1][tx + 1] = c[r];
        // x halo columns of this tile's rows
        if (tx == 0) {
#pragma unroll
            for (int r = 0; r < RY; r++) plane[ty * RY + r + 1][0] = h[r];
        }
        if (tx == BX - 1) {
#pragma unroll
            for (int r = 0; r < RY; r++) plane[ty * RY + r + 1][BX + 1] = h[r];
        }
        // y halo rows
        if (ty == 0) plane[0][tx + 1] = g;
        if (ty == BY - 1) plane[TH + 1][tx + 1] = g;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RY; r++) {
            if (act[r]) {
                const int ly = ty * RY + r + 1;
                float s = plane[ly][tx] + plane[ly][tx + 2];
                s = s + plane[ly - 1][tx + 1];
                s = s + plane[ly + 1][tx + 1];
                s = s + m[r];
                s = s + a[r];
                const float cc = c[r];
                out[id[r]] = cc + alpha * (s - 6.0f * cc);
            }
        }
        __syncthreads();
        // the rows move to plane z+1; plane z+3 is read at step z+2 (if
        // z+2 < z1: else plane z1 again) and so is plane z+2's halo (else
        // plane z1-1's)
#pragma unroll
        for (int r = 0; r < RY; r++) id[r] += sxy;
        const int64_t k = min(z + 3, z1) - (z + 1);
#pragma unroll
        for (int r = 0; r < RY; r++) m[r] = in[id[r] + k * sxy];
        load_halo(k - 1, h, g);
    };
    for (int z = z0; z < z1; z += 4) {
        step(z, P[0], P[1], P[2], H[0], G[0]);
        if (z + 1 >= z1) break;
        step(z + 1, P[1], P[2], P[3], H[1], G[1]);
        if (z + 2 >= z1) break;
        step(z + 2, P[2], P[3], P[0], H[0], G[0]);
        if (z + 3 >= z1) break;
        step(z + 3, P[3], P[0], P[1], H[1], G[1]);
    }
}

// Two sweeps fused per pass (temporal blocking).  A workgroup of 64 x 4
// threads covers a 64 x 4*RY window of the x-y plane and marches a z-chunk;
// the window's outer two cells on each side are its halo, so it writes a
// 60 x (4*RY - 4) tile of the second sweep's field.  Per plane it loads the
// input (level 0) once, computes the first sweep (level 1) on the window
// minus one cell -- kept in registers and in an LDS plane, never in HBM --
// and from three level-1 planes the second sweep (level 2) of the plane
// below.  Level 1 at the grid's halo cells is the fixed zero.  Each cell's
// arithmetic is the single sweep's, so two passes' worth of output is
// bit-identical to two sweeps; the field crosses HBM once per two steps.
// The window (bx, by) of z-chunk bz, for a 64 x 4 workgroup.
template <int RY>
__device__ __forceinline__ void tb2_window(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                           int nz, int zc, float alpha, int bx, int by, int bz)
{
    constexpr int WX = 64, WY = 4 * RY, OX = WX - 4, OY = WY - 4;
    __shared__ float p0[WY][WX];     // level-0 plane z
    __shared__ float p1[2][WY][WX];  // level-1 planes (ping-pong)

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int sx = nx + 2;
    const int64_t sxy = (int64_t)sx * (ny + 2);

    // column of this thread: global x = bx*OX - 1 + tx (window starts two
    // cells left of the tile's first output x, which is bx*OX + 1)
    const int gx = bx * OX - 1 + tx;
    const int xl = min(max(gx, 0), nx + 1);  // clamped for loads
    const bool xin1 = tx >= 1 && tx <= WX - 2;   // level-1 columns of the window
    const bool xin2 = tx >= 2 && tx <= WX - 3 && gx <= nx;  // output columns
    const bool xhalo = gx <= 0 || gx >= nx + 1;  // level 1 is the fixed zero here

    const int z0 = bz * zc + 1;
    const int z1 = min(z0 + zc, nz + 1);  // output planes [z0, z1)

    int64_t col[RY];       // offset of (x, y) in a plane
    bool act1[RY], act2[RY], zero1[RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        const int row = ty * RY + r;
        const int gy = by * OY - 1 + row;
        const int yl = min(max(gy, 0), ny + 1);
        col[r] = xl + (int64_t)sx * yl;
        act1[r] = xin1 && row >= 1 && row <= WY - 2;
        act2[r] = xin2 && row >= 2 && row <= WY - 3 && gy <= ny;
        zero1[r] = xhalo || gy <= 0 || gy >= ny + 1;
    }
    auto load = [&](int r, int z) { return in[col[r] + sxy * (int64_t)min(max(z, 0), nz + 1)]; };

    // Level 0 at planes z-1 .. z+2 (A) and level 1 at planes z-2, z-1 (B),
    // z being the level-1 plane computed next, starting one below the first
    // output.  Level 0 is loaded two planes before its use, and the z loop
    // is unrolled by four so that the roles rotate over the registers and
    // none is copied into another (a copy would wait for its load at once,
    // exposing the memory latency on every plane).
    float A[4][RY], B[2][RY];
#pragma unroll
    for (int r = 0; r < RY; r++) {
        A[0][r] = load(r, z0 - 2);
        A[1][r] = load(r, z0 - 1);
        A[2][r] = load(r, z0);
        A[3][r] = load(r, z0 + 1);
        B[0][r] = 0.0f;
        B[1][r] = 0.0f;
    }
    // Plane z: am, ac, ap hold level 0 at planes z-1, z, z+1; bm, bc level 1
    // at z-2, z-1; level 1 of plane z goes to LDS buffer cb.  Afterwards am
    // takes level 0 at plane z+3 (read two steps later) and bm level 1 at z.
    auto step = [&](int z, float (&am)[RY], const float (&ac)[RY], const float (&ap)[RY], float (&bm)[RY],
                    const float (&bc)[RY], int cb) {
        // level 0 plane z to LDS for the in-plane neighbours
#pragma unroll
        for (int r = 0; r < RY; r++) p0[ty * RY + r][tx] = ac[r];
        __syncthreads();
        // level 1 at plane z (zero on the grid's halo)
        const bool zhalo = z <= 0 || z >= nz + 1;
        float l1[RY];
#pragma unroll
        for (int r = 0; r < RY; r++) {
            float v = 0.0f;
            if (act1[r] && !zero1[r] && !zhalo) {
                const int ly = ty * RY + r;
                float sum = p0[ly][tx - 1] + p0[ly][tx + 1];
                sum = sum + p0[ly - 1][tx];
                sum = sum + p0[ly + 1][tx];
                sum = sum + am[r];
                sum = sum + ap[r];
                const float c = ac[r];
                v = c + alpha * (sum - 6.0f * c);
            }
            l1[r] = v;
            p1[cb][ty * RY + r][tx] = v;
        }
        // level 2 at plane z-1 from level-1 planes z-2 (bm), z-1 (LDS, bc), z (l1)
        if (z - 1 >= z0) {
#pragma unroll
            for (int r = 0; r < RY; r++) {
                if (act2[r]) {
                    const int ly = ty * RY + r;
                    const float(*q)[WX] = p1[1 - cb];
                    float sum = q[ly][tx - 1] + q[ly][tx + 1];
                    sum = sum + q[ly - 1][tx];
                    sum = sum + q[ly + 1][tx];
                    sum = sum + bm[r];
                    sum = sum + l1[r];
                    const float c = bc[r];
                    out[col[r] + sxy * (int64_t)(z - 1)] = c + alpha * (sum - 6.0f * c);
                }
            }
        }
        __syncthreads();
        // level 0 at plane z+3 (read at step z+2), or plane z1+1 again, the
        // chunk's last read, when no step z+2 follows
#pragma unroll
        for (int r = 0; r < RY; r++) {
            am[r] = load(r, min(z + 3, z1 + 1));
            bm[r] = l1[r];
        }
    };
    for (int z = z0 - 1; z <= z1; z += 4) {
        step(z, A[0], A[1], A[2], B[0], B[1], 0);
        if (z + 1 > z1) break;
        step(z + 1, A[1], A[2], A[3], B[1], B[0], 1);
        if (z + 2 > z1) break;
        step(z + 2, A[2], A[3], A[0], B[0], B[1], 0);
        if (z + 3 > z1) break;
        step(z + 3, A[3], A[0], A[1], B[1], B[0], 1);
    }
}

template <int RY>
__global__ __launch_bounds__(256) void heat_tb2(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                                int nz, int zc, float alpha, int tiles_x, int tiles_y, int tiles_z)
{
    const int nblocks = tiles_x * tiles_y * tiles_z;
    const int t = xcd_swizzle(blockIdx.x, nblocks);
    tb2_window<RY>(in, out, nx, ny, nz, zc, alpha, t % tiles_x, (t / tiles_x) % tiles_y, t / (tiles_x * tiles_y));
}

// The same two fused sweeps over several grids in one launch: workgroup t
// takes window t - first of the last grid whose first workgroup is <= t.
template <int RY>
__global__ __launch_bounds__(256) void heat_tb2_batch(const pat::HeatBatchGrid *__restrict__ grids, int ngrids,
                                                      int nblocks, float alpha)
{
    const int t = xcd_swizzle(blockIdx.x, nblocks);
    int lo = 0, hi = ngrids - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (grids[mid].first <= t) lo = mid;
        else hi = mid - 1;
    }
    const pat::HeatBatchGrid g = grids[lo];
    const int u = t - g.first;
    tb2_window<RY>(g.in, g.out, g.nx, g.ny, g.nz, g.zc, alpha, u % g.tiles_x, (u / g.tiles_x) % g.tiles_y,
                   u / (g.tiles_x * g.tiles_y));
}

// Several sweeps per launch on blocks of the grid held in LDS, for small
// grids, whose single sweeps are bound by a few memory latencies per launch
// (and the fused window by its halo work) rather than by bandwidth.  A
// workgroup loads its BX x BY x BZ block of cells with a T-cell margin,
// sweeps that copy T times between two LDS buffers (sweep s updates the
// block grown by T - s cells: the cells whose remaining dependence stays
// inside the copy) and writes the block back.  The grid's halo cells keep
// their values; positions beyond the halo are read only by halo cells,
// which are copied, not updated.  Each cell's arithmetic is the single
// sweep's, so the result is bit-identical to T sweeps.
constexpr int kCubeThreads = 1024;

// Sweep S: the region grown by T - S cells, by (x, y) columns (split into
// ZS segments of z when the columns are fewer than the threads), each
// thread marching z with its column's z-1, z, z+1 values in registers (the
// march unrolled, so the values rotate through fixed registers), so that a
// cell reads four neighbours from LDS and the x-y bounds are tested once.
template <int BX, int BY, int BZ, int T, int S>
__device__ __forceinline__ void cube_sweep(const float *__restrict__ a, float *__restrict__ b, int gx0, int gy0,
                                           int gz0, int nx, int ny, int nz, float alpha)
{
    constexpr int EX = BX + 2 * T, EXY = EX * (BY + 2 * T);
    constexpr int RX = BX + 2 * (T - S), RY = BY + 2 * (T - S), RZ = BZ + 2 * (T - S);
    constexpr int NC = RX * RY;
    constexpr int ZS = NC >= kCubeThreads ? 1 : (kCubeThreads / NC < RZ ? kCubeThreads / NC : RZ);
    constexpr int LZ = (RZ + ZS - 1) / ZS;
#pragma unroll
    for (int i = 0; i < (NC * ZS + kCubeThreads - 1) / kCubeThreads; i++) {
        const int w = (int)threadIdx.x + i * kCubeThreads;
        if (w < NC * ZS) {
            const int col = w % NC, seg = ZS == 1 ? 0 : w / NC;  // one segment: z uniform over the wavefront
            const int x = S + col % RX, y = S + col / RX;
            const int z0 = S + seg * LZ, nzs = min(LZ, S + RZ - z0);
            const int gx = gx0 + x, gy = gy0 + y;
            const bool inxy = gx >= 1 && gx <= nx && gy >= 1 && gy <= ny;
            int l = x + EX * y + EXY * z0;
            float zm = a[l - EXY], zc = a[l];
#pragma unroll
            for (int k = 0; k < LZ; k++) {
                if (k < nzs) {
                    const float zp = a[l + EXY];
                    const int gz = gz0 + z0 + k;
                    // computed everywhere (every read lies in the copy) and
                    // kept only inside the grid: a select, no branch
                    float sum = a[l - 1] + a[l + 1];
                    sum = sum + a[l - EX];
                    sum = sum + a[l + EX];
                    sum = sum + zm;
                    sum = sum + zp;
                    const float v = zc + alpha * (sum - 6.0f * zc);
                    b[l] = inxy && gz >= 1 && gz <= nz ? v : zc;
                    zm = zc;
                    zc = zp;
                    l += EXY;
                }
            }
        }
    }
}

// Sweeps S .. T: sweep S reads buffer (S - 1) % 2 and writes S % 2.
template <int BX, int BY, int BZ, int T, int S = 1>
__device__ __forceinline__ void cube_sweeps(float *c0, float *c1, int gx0, int gy0, int gz0, int nx, int ny, int nz,
                                            float alpha)
{
    if constexpr (S <= T) {
        cube_sweep<BX, BY, BZ, T, S>((S & 1) ? c0 : c1, (S & 1) ? c1 : c0, gx0, gy0, gz0, nx, ny, nz, alpha);
        __syncthreads();
        cube_sweeps<BX, BY, BZ, T, S + 1>(c0, c1, gx0, gy0, gz0, nx, ny, nz, alpha);
    }
}

template <int BX, int BY, int BZ, int T>
__global__ __launch_bounds__(kCubeThreads) void heat_cube(const float *__restrict__ in, float *__restrict__ out, int nx,
                                                  int ny, int nz, float alpha, int tiles_x, int tiles_y, int tiles_z)
{
    constexpr int EX = BX + 2 * T, EY = BY + 2 * T, EZ = BZ + 2 * T, EXY = EX * EY, E3 = EXY * EZ;
    extern __shared__ float cube[];  // two buffers of E3 cells
    float *c0 = cube, *c1 = cube + E3;

    const int nblocks = tiles_x * tiles_y * tiles_z;
    const int t = xcd_swizzle(blockIdx.x, nblocks);
    const int bx = t % tiles_x;
    const int by = (t / tiles_x) % tiles_y;
    const int bz = t / (tiles_x * tiles_y);
    // array coordinates (halo-inclusive) of the copy's first cell
    const int gx0 = bx * BX + 1 - T, gy0 = by * BY + 1 - T, gz0 = bz * BZ + 1 - T;
    const int sx = nx + 2;
    const int64_t sxy = (int64_t)sx * (ny + 2);

    // the copy, every load issued before the first LDS write; cells beyond
    // the array read as 0 (never used by a written cell)
    constexpr int NL = (E3 + kCubeThreads - 1) / kCubeThreads;
    float v[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
        const int c = (int)threadIdx.x + i * kCubeThreads;
        const int gx = gx0 + c % EX, gy = gy0 + (c / EX) % EY, gz = gz0 + c / EXY;
        v[i] = 0.0f;
        if (c < E3 && gx >= 0 && gx <= nx + 1 && gy >= 0 && gy <= ny + 1 && gz >= 0 && gz <= nz + 1)
            v[i] = in[gx + (int64_t)sx * gy + sxy * gz];
    }
#pragma unroll
    for (int i = 0; i < NL; i++) {
        const int c = (int)threadIdx.x + i * kCubeThreads;
        if (c < E3) c0[c] = v[i];
    }
    __syncthreads();
    cube_sweeps<BX, BY, BZ, T>(c0, c1, gx0, gy0, gz0, nx, ny, nz, alpha);
    const float *r = (T & 1) ? c1 : c0;
    // the block's cells inside the grid
#pragma unroll
    for (int i = 0; i < (BX * BY * BZ + kCubeThreads - 1) / kCubeThreads; i++) {
        const int c = (int)threadIdx.x + i * kCubeThreads;
        const int x = T + c % BX, y = T + (c / BX) % BY, z = T + c / (BX * BY);
        const int gx = gx0 + x, gy = gy0 + y, gz = gz0 + z;
        if (c < BX * BY * BZ && gx <= nx && gy <= ny && gz <= nz)
            out[gx + (int64_t)sx * gy + sxy * gz] = r[x + EX * y + EXY * z];
    }
}

// One thread per interior cell, all seven loads from global memory (through
// the caches): the shape of the reference's per-cell kernel, kept as the
// plain baseline and as an independent check of the tiled kernel.
__global__ __launch_bounds__(256) void heat_naive(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                                                  int nz, float alpha)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int x = (int)(g % nx) + 1;
    const int y = (int)((g / nx) % ny) + 1;
    const int z = (int)(g / ((int64_t)nx * ny)) + 1;
    const int sx = nx + 2;
    const int64_t sxy = (int64_t)sx * (ny + 2);
    const int64_t id = x + (int64_t)sx * (y + (int64_t)(ny + 2) * z);
    const float c = in[id];
    float s = in[id - 1] + in[id + 1];
    s = s + in[id - sx];
    s = s + in[id + sx];
    s = s + in[id - sxy];
    s = s + in[id + sxy];
    out[id] = c + alpha * (s - 6.0f * c);
}

struct TileShape {
    int bx, by, ry;
};

// Instantiated tile shapes (BX threads in x, BY thread rows, RY rows per thread).
constexpr TileShape kShapes[] = {{64, 4, 1}, {64, 4, 2}, {64, 4, 4}, {64, 8, 1}, {64, 8, 2}, {64, 2, 4}, {64, 16, 1},
                                 {32, 8, 2}, {128, 2, 2}, {128, 4, 1}};

using HeatKernel = void (*)(const float *, float *, int, int, int, int, float, int, int, int);

HeatKernel heat_kernel_for(int bx, int by, int ry)
{
#define PAT_SHAPE(X, Y, R)                                                                                             \
    if (bx == X && by == Y && ry == R) return heat_lds25d<X, Y, R>;
    PAT_SHAPE(64, 4, 1)
    PAT_SHAPE(64, 4, 2)
    PAT_SHAPE(64, 4, 4)
    PAT_SHAPE(64, 8, 1)
    PAT_SHAPE(64, 8, 2)
    PAT_SHAPE(64, 2, 4)
    PAT_SHAPE(64, 16, 1)
    PAT_SHAPE(32, 8, 2)
    PAT_SHAPE(128, 2, 2)
    PAT_SHAPE(128, 4, 1)
#undef PAT_SHAPE
    return nullptr;
}

// LDS cube shapes (block cells per axis, sweeps per launch); two buffers of
// (BX+2T)(BY+2T)(BZ+2T) floats each, 41-111 KB.
struct CubeShape {
    int bx, by, bz, t;
};
constexpr CubeShape kCubes[] = {{16, 16, 16, 2}, {16, 16, 16, 3}, {16, 16, 16, 4}, {32, 16, 8, 2}, {32, 8, 8, 2}};
constexpr int kNumCubes = sizeof(kCubes) / sizeof(kCubes[0]);

using CubeKernel = void (*)(const float *, float *, int, int, int, float, int, int, int);

CubeKernel cube_kernel(int shape)
{
    switch (shape) {
    case 0: return heat_cube<16, 16, 16, 2>;
    case 1: return heat_cube<16, 16, 16, 3>;
    case 2: return heat_cube<16, 16, 16, 4>;
    case 3: return heat_cube<32, 16, 8, 2>;
    case 4: return heat_cube<32, 8, 8, 2>;
    default: return nullptr;
    }
}

size_t cube_lds_bytes(int shape)
{
    const CubeShape &c = kCubes[shape];
    return 2 * sizeof(float) * (size_t)(c.bx + 2 * c.t) * (c.by + 2 * c.t) * (c.bz + 2 * c.t);
}

// Lets each cube kernel use its dynamic LDS (above the default 64 KB);
// once per process, before any launch or capture.
int cube_prepare()
{
    static std::once_flag once;
    static int rc = PAT_OK;
    std::call_once(once, [] {
        for (int s = 0; s < kNumCubes; s++)
            if (hipFuncSetAttribute((const void *)cube_kernel(s), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)cube_lds_bytes(s)) != hipSuccess) {
                (void)hipGetLastError();
                pat::set_error("heat3d: LDS size for cube shape %d not accepted", s);
                rc = PAT_ERR_HIP;
            }
    });
    return rc;
}

// Default tile for a grid: 64-wide rows, 4x2 thread rows, and a z-chunk that
// yields enough workgroups (>= ~4 per CU on 256 CUs) without making chunks so
// short that the two extra halo planes per chunk dominate.
void default_tile(int64_t nx, int64_t ny, int64_t nz, int *bx, int *by, int *ry, int *zc)
{
    *bx = 64;
    *by = 4;
    *ry = 2;
    const int64_t txy = ((nx + 63) / 64) * ((ny + 7) / 8);
    int64_t want_z = (2048 + txy - 1) / txy;  // chunks needed for ~2048 workgroups
    int64_t c = (nz + want_z - 1) / want_z;
    if (c < 8) c = 8;
    if (c > nz) c = nz;
    *zc = (int)c;
}

}  // namespace

namespace pat {

int heat3d_sweep(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, int variant, int bx,
                 int by, int ry, int zc, hipStream_t stream)
{
    if (variant == PAT_STENCIL_GENERIC) {
        const int64_t n = nx * ny * nz;
        hipLaunchKernelGGL(heat_naive, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, in, out, (int)nx,
                           (int)ny, (int)nz, alpha);
        PAT_LAUNCH_CHECK();
        return PAT_OK;
    }
    HeatKernel k = heat_kernel_for(bx, by, ry);
    PAT_REQUIRE(k, PAT_ERR_UNSUPPORTED, "heat3d: tile %dx%dx%d not instantiated", bx, by, ry);
    PAT_REQUIRE(zc > 0, PAT_ERR_INVALID, "heat3d: z-chunk must be positive");
    const int tiles_x = (int)((nx + bx - 1) / bx);
    const int tiles_y = (int)((ny + by * ry - 1) / (by * ry));
    const int tiles_z = (int)((nz + zc - 1) / zc);
    const int64_t nblocks = (int64_t)tiles_x * tiles_y * tiles_z;
    PAT_REQUIRE(nblocks < (1LL << 31), PAT_ERR_INVALID, "heat3d: grid too large");
    hipLaunchKernelGGL(k, dim3((unsigned)nblocks), dim3(bx, by), 0, stream, in, out, (int)nx, (int)ny, (int)nz, zc,
                       alpha, tiles_x, tiles_y, tiles_z);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

void heat3d_default_tile(int64_t nx, int64_t ny, int64_t nz, int *bx, int *by, int *ry, int *zc)
{
    default_tile(nx, ny, nz, bx, by, ry, zc);
}

// Two fused sweeps in -> out (heat_tb2): window height 4*ry rows (ry 4 or 8),
// z-chunk zc.
int heat3d_sweep2(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, int ry, int zc,
                  hipStream_t stream)
{
    PAT_REQUIRE(ry == 2 || ry == 4 || ry == 8, PAT_ERR_UNSUPPORTED,
                "heat3d: fused window of %d rows per thread not instantiated", ry);
    PAT_REQUIRE(zc > 0, PAT_ERR_INVALID, "heat3d: z-chunk must be positive");
    const int oy = 4 * ry - 4;
    const int tiles_x = (int)((nx + 59) / 60);
    const int tiles_y = (int)((ny + oy - 1) / oy);
    const int tiles_z = (int)((nz + zc - 1) / zc);
    const int64_t nblocks = (int64_t)tiles_x * tiles_y * tiles_z;
    PAT_REQUIRE(nblocks < (1LL << 31), PAT_ERR_INVALID, "heat3d: grid too large");
    auto k = ry == 2 ? heat_tb2<2> : (ry == 4 ? heat_tb2<4> : heat_tb2<8>);
    hipLaunchKernelGGL(k, dim3((unsigned)nblocks), dim3(64, 4), 0, stream, in, out, (int)nx, (int)ny, (int)nz, zc,
                       alpha, tiles_x, tiles_y, tiles_z);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// kCubes[shape].t sweeps in -> out on LDS cubes.
int heat3d_cube(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, int shape,
                hipStream_t stream)
{
    PAT_REQUIRE(shape >= 0 && shape < kNumCubes, PAT_ERR_UNSUPPORTED, "heat3d: cube shape %d not instantiated",
                shape);
    PAT_TRY(cube_prepare());
    const CubeShape &c = kCubes[shape];
    const int tiles_x = (int)((nx + c.bx - 1) / c.bx);
    const int tiles_y = (int)((ny + c.by - 1) / c.by);
    const int tiles_z = (int)((nz + c.bz - 1) / c.bz);
    const int64_t nblocks = (int64_t)tiles_x * tiles_y * tiles_z;
    PAT_REQUIRE(nblocks < (1LL << 31), PAT_ERR_INVALID, "heat3d: grid too large");
    hipLaunchKernelGGL(cube_kernel(shape), dim3((unsigned)nblocks), dim3(kCubeThreads), cube_lds_bytes(shape), stream, in,
                       out, (int)nx, (int)ny, (int)nz, alpha, tiles_x, tiles_y, tiles_z);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// One launch of tile t: a single sweep (fused == 0), two (fused == 1) or
// t.ry on LDS cubes (fused == 2).
int heat3d_launch(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, const HeatTile &t,
                  int variant, hipStream_t stream)
{
    if (t.fused == 2) return heat3d_cube(in, out, nx, ny, nz, alpha, t.bx, stream);
    if (t.fused) return heat3d_sweep2(in, out, nx, ny, nz, alpha, t.ry, t.zc, stream);
    return heat3d_sweep(in, out, nx, ny, nz, alpha, variant, t.bx, t.by, t.ry, t.zc, stream);
}

// nsteps sweeps ping-ponging buf[start] -> buf[1-start] ...: fused pairs
// while two or more steps remain, a plain sweep for an odd last one.
// Returns the buffer index holding the result in *end.
int heat3d_batch_table(int n, float *const *a, float *const *b, const int64_t *nx, const int64_t *ny,
                       const int64_t *nz, HeatBatchGrid *table, hipStream_t st, int *nblocks)
{
    PAT_REQUIRE(n > 0 && a && b && nx && ny && nz && table && nblocks, PAT_ERR_INVALID, "heat3d batch: bad arguments");
    constexpr int OX = 60, OY = 12;  // output window of the 64 x 16 fused window (4 rows per thread)
    std::vector<HeatBatchGrid> h(2 * (size_t)n);
    int64_t total = 0;
    for (int i = 0; i < n; i++) {
        PAT_REQUIRE(a[i] && b[i] && a[i] != b[i] && nx[i] > 0 && ny[i] > 0 && nz[i] > 0 &&
                        (nx[i] + 2) * (ny[i] + 2) * (nz[i] + 2) < (1LL << 40),
                    PAT_ERR_INVALID, "heat3d batch: grid %d has bad buffers or sizes", i);
        HeatBatchGrid g;
        g.in = a[i];
        g.out = b[i];
        g.nx = (int)nx[i];
        g.ny = (int)ny[i];
        g.nz = (int)nz[i];
        g.zc = (int)std::min<int64_t>(8, nz[i]);
        g.tiles_x = (int)((nx[i] + OX - 1) / OX);
        g.tiles_y = (int)((ny[i] + OY - 1) / OY);
        g.first = (int)total;
        g.pad = 0;
        total += (int64_t)g.tiles_x * g.tiles_y * ((nz[i] + g.zc - 1) / g.zc);
        PAT_REQUIRE(total < (1LL << 31), PAT_ERR_INVALID, "heat3d batch: too many workgroups");
        h[i] = g;
        g.in = b[i];
        g.out = a[i];
        h[n + i] = g;
    }
    PAT_HIP(hipMemcpyAsync(table, h.data(), sizeof(HeatBatchGrid) * h.size(), hipMemcpyHostToDevice, st));
    PAT_HIP(hipStreamSynchronize(st));  // the host copy goes out of scope
    *nblocks = (int)total;
    return PAT_OK;
}

int heat3d_batch_enqueue(int n, float *const *a, float *const *b, const int64_t *nx, const int64_t *ny,
                         const int64_t *nz, float alpha, int nsteps, const HeatBatchGrid *table, int nblocks,
                         hipStream_t st, int *in_b)
{
    PAT_REQUIRE(n > 0 && nblocks > 0 && nsteps >= 0 && table && in_b, PAT_ERR_INVALID, "heat3d batch: bad arguments");
    int c = 0;
    for (int k = 0; k + 2 <= nsteps; k += 2) {
        hipLaunchKernelGGL(heat_tb2_batch<4>, dim3((unsigned)nblocks), dim3(64, 4), 0, st, table + (size_t)c * n, n,
                           nblocks, alpha);
        PAT_LAUNCH_CHECK();
        c = 1 - c;
    }
    if (nsteps % 2) {
        for (int i = 0; i < n; i++) {
            int bx, by, ry, zc;
            default_tile(nx[i], ny[i], nz[i], &bx, &by, &ry, &zc);
            PAT_TRY(heat3d_sweep(c ? b[i] : a[i], c ? a[i] : b[i], nx[i], ny[i], nz[i], alpha, PAT_STENCIL_LDS25D, bx,
                                 by, ry, zc, st));
        }
        c = 1 - c;
    }
    *in_b = c;
    return PAT_OK;
}

int heat3d_steps(float *const buf[2], int start, int64_t nx, int64_t ny, int64_t nz, float alpha, const HeatTile &t,
                 int variant, int nsteps, hipStream_t stream, int *end)
{
    int c = start;
    int k = 0;
    const int spl = heat3d_steps_per_launch(t);
    if (spl > 1) {
        for (; k + spl <= nsteps; k += spl) {
            PAT_TRY(heat3d_launch(buf[c], buf[1 - c], nx, ny, nz, alpha, t, variant, stream));
            c = 1 - c;
        }
    }
    if (k < nsteps) {
        HeatTile p = t;
        if (t.fused) {  // the steps left over run on the plain kernel
            p.fused = 0;
            default_tile(nx, ny, nz, &p.bx, &p.by, &p.ry, &p.zc);
        }
        for (; k < nsteps; k++) {
            PAT_TRY(heat3d_sweep(buf[c], buf[1 - c], nx, ny, nz, alpha, variant, p.bx, p.by, p.ry, p.zc, stream));
            c = 1 - c;
        }
    }
    *end = c;
    return PAT_OK;
}

// Autotuner: times every instantiated tile shape with z-chunks of 2..64
// planes on the caller's grid (in -> out; out is scratch until the first
// real sweep writes it, and no kernel writes the halo), keeps the fastest,
// and caches it for the process per size class: each extent's half octave
// (floor(2 log2 n)), so a stream of grids of many sizes (a serving mix)
// tunes once per class; every tile gives the same bits on any grid.
static std::map<std::tuple<int, int, int, int>, HeatTile> g_tile_cache;
static std::mutex g_tile_mutex;

static int size_class(int64_t n) { return (int)std::floor(2.0 * std::log2((double)n)); }

// allow_fused: whether two-sweep tiles may be chosen (the caller runs
// pairs of steps); their time is compared per step.
int heat3d_tuned(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, hipStream_t st,
                 bool allow_fused, HeatTile *t)
{
    const auto key = std::make_tuple(size_class(nx), size_class(ny), size_class(nz), (int)allow_fused);
    {
        std::lock_guard<std::mutex> g(g_tile_mutex);
        auto it = g_tile_cache.find(key);
        if (it != g_tile_cache.end()) {
            *t = it->second;
            t->zc = (int)std::min<int64_t>(t->zc, nz);
            return PAT_OK;
        }
    }
    t->fused = 0;
    default_tile(nx, ny, nz, &t->bx, &t->by, &t->ry, &t->zc);
    if (!pat::tuning_enabled()) {
        // untuned: the two-sweep window the tuner picks at 128^3 and 384^3
        // where two sweeps per launch are allowed (same values either way)
        if (allow_fused) *t = HeatTile{1, 64, 4, 4, (int)std::min<int64_t>(8, nz)};
        return PAT_OK;
    }
    std::vector<HeatTile> cands;
    for (const auto &sh : kShapes)
        for (int z : {2, 4, 8, 16, 32, 64}) {
            if (z > nz && z != 2) continue;
            cands.push_back(HeatTile{0, sh.bx, sh.by, sh.ry, (int)std::min<int64_t>(z, nz)});
        }
    if (allow_fused) {
        for (int r : {2, 4, 8})
            for (int z : {2, 4, 8, 16, 32, 64, 128}) {
                if (z > nz && z != 2) continue;
                cands.push_back(HeatTile{1, 64, 4, r, (int)std::min<int64_t>(z, nz)});
            }
        PAT_TRY(cube_prepare());  // before any capture
        for (int c = 0; c < kNumCubes; c++) cands.push_back(HeatTile{2, c, 0, kCubes[c].t, 0});
    }
    pat::EventPair ev;
    PAT_HIP(ev.create());
    hipEvent_t a = ev.a, b = ev.b;
    // a few milliseconds of warm-up on the default tile (clocks ramp up
    // under load), then rounds that time every candidate in turn; each
    // candidate keeps its best round, multi-sweep ones counted per step
    const HeatTile def = *t;
    auto w0 = std::chrono::steady_clock::now();
    int warm = 0;
    for (; warm < 2000; warm++) {
        PAT_TRY(heat3d_launch(in, out, nx, ny, nz, alpha, def, PAT_STENCIL_LDS25D, st));
        if (warm % 8 == 7) {
            PAT_HIP(hipStreamSynchronize(st));
            if (std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(5)) {
                warm++;
                break;
            }
        }
    }
    const double launch_us =
        std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - w0).count() / warm;
    // Each timing covers kReps launches of the same sweep.  They run as a
    // captured graph where the stream allows it (steps run as graph
    // replays in use), since back-to-back launches from the host cost
    // several microseconds each, which hides every kernel faster than that;
    // the null stream and a stream already capturing time back-to-back
    // launches instead.
    // kReps covers ~100 us (1 to 8 launches): enough to hide the timing's
    // own cost on small grids without making large grids' tuning, which a
    // serving job waits for, eight times longer.
    const int kReps = std::max(1, std::min(8, (int)std::ceil(100.0 / std::max(launch_us, 1.0))));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (st) PAT_HIP(hipStreamIsCapturing(st, &cap));
    const bool by_graph = st != nullptr && cap == hipStreamCaptureStatusNone;
    struct Execs {  // one graph per candidate, destroyed on every path
        std::vector<hipGraphExec_t> v;
        ~Execs()
        {
            for (hipGraphExec_t e : v)
                if (e) (void)hipGraphExecDestroy(e);
        }
    } ex;
    ex.v.assign(cands.size(), nullptr);
    auto build = [&](size_t c) -> int {
        hipGraph_t g = nullptr;
        PAT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = PAT_OK;
        for (int i = 0; i < kReps && rc == PAT_OK; i++)
            rc = heat3d_launch(in, out, nx, ny, nz, alpha, cands[c], PAT_STENCIL_LDS25D, st);
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc != PAT_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PAT_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&ex.v[c], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PAT_HIP(ei);
        PAT_HIP(hipGraphLaunch(ex.v[c], st));  // the first replay carries one-time costs
        return PAT_OK;
    };
    std::vector<float> tm(cands.size(), 1e30f);
    for (int r = 0; r < 3; r++)
        for (size_t c = 0; c < cands.size(); c++) {
            if (by_graph && !ex.v[c]) PAT_TRY(build(c));
            PAT_HIP(hipEventRecord(a, st));
            if (by_graph) {
                PAT_HIP(hipGraphLaunch(ex.v[c], st));
            } else {
                for (int i = 0; i < kReps; i++)
                    PAT_TRY(heat3d_launch(in, out, nx, ny, nz, alpha, cands[c], PAT_STENCIL_LDS25D, st));
            }
            PAT_HIP(hipEventRecord(b, st));
            PAT_HIP(hipEventSynchronize(b));
            float ms = 0;
            PAT_HIP(hipEventElapsedTime(&ms, a, b));
            ms /= kReps;
            tm[c] = std::min(tm[c], ms / heat3d_steps_per_launch(cands[c]));
        }
    size_t bi = 0;
    for (size_t c = 1; c < cands.size(); c++)
        if (tm[c] < tm[bi]) bi = c;
    if (pat::tune_log())
        for (size_t c = 0; c < cands.size(); c++)
            fprintf(stderr, "[tune heat3d %lldx%lldx%lld] fused %d bx %d by %d ry %d zc %d: %.2f us/step%s\n",
                    (long long)nx, (long long)ny, (long long)nz, cands[c].fused, cands[c].bx, cands[c].by,
                    cands[c].ry, cands[c].zc, 1000.0f * tm[c], c == bi ? "  <- chosen" : "");
    *t = cands[bi];
    std::lock_guard<std::mutex> g(g_tile_mutex);
    g_tile_cache[key] = *t;
    return PAT_OK;
}

int heat3d_tuned_tile(const float *in, float *out, int64_t nx, int64_t ny, int64_t nz, float alpha, hipStream_t st,
                      int *bx, int *by, int *ry, int *zc)
{
    HeatTile t;
    PAT_TRY(heat3d_tuned(in, out, nx, ny, nz, alpha, st, false, &t));
    *bx = t.bx;
    *by = t.by;
    *ry = t.ry;
    *zc = t.zc;
    return PAT_OK;
}

// Loads this file's code object (kernels are otherwise loaded at their first
// launch, which a serving job would wait for).
int stencil_preload() { return cube_prepare(); }

int heat3d_cube_tile(int shape, HeatTile *t)
{
    PAT_REQUIRE(shape >= 0 && shape < kNumCubes, PAT_ERR_UNSUPPORTED, "heat3d: cube shape %d not instantiated",
                shape);
    *t = HeatTile{2, shape, 0, kCubes[shape].t, 0};
    return PAT_OK;
}

int heat3d_shapes(const int **triples)
{
    static int flat[sizeof(kShapes) / sizeof(kShapes[0]) * 3];
    int n = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
    for (int i = 0; i < n; i++) {
        flat[3 * i] = kShapes[i].bx;
        flat[3 * i + 1] = kShapes[i].by;
        flat[3 * i + 2] = kShapes[i].ry;
    }
    *triples = flat;
    return n;
}

}  // namespace pat

extern "C" int pat_heat3d_run_dev(float *a, float *b, int64_t nx, int64_t ny, int64_t nz, int nsteps, float alpha,
                                  int variant, int bx, int by, int zc, int *result_in_b, void *stream)
{
    PAT_REQUIRE(a && b && a != b, PAT_ERR_INVALID, "pat_heat3d_run_dev: need two distinct device grids");
    PAT_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nsteps >= 0, PAT_ERR_INVALID, "pat_heat3d_run_dev: bad sizes");
    PAT_REQUIRE((nx + 2) * (ny + 2) * (nz + 2) < (1LL << 40), PAT_ERR_INVALID, "pat_heat3d_run_dev: grid too large");
    PAT_REQUIRE(variant >= PAT_STENCIL_AUTO && variant <= PAT_STENCIL_CUBES, PAT_ERR_UNSUPPORTED,
                "pat_heat3d_run_dev: unknown variant %d", variant);
    hipStream_t st = (hipStream_t)stream;
    pat::HeatTile t;
    if (variant == PAT_STENCIL_AUTO && bx == 0 && by == 0 && zc == 0) {
        PAT_TRY(pat::heat3d_tuned(a, b, nx, ny, nz, alpha, st, true, &t));
    } else if (variant == PAT_STENCIL_CUBES) {
        PAT_TRY(pat::heat3d_cube_tile(bx, &t));
    } else if (variant == PAT_STENCIL_TEMPORAL2) {
        PAT_REQUIRE(bx == 0 || bx == 64, PAT_ERR_UNSUPPORTED, "pat_heat3d_run_dev: fused windows are 64 wide");
        PAT_REQUIRE(by == 0 || by == 8 || by == 16 || by == 32, PAT_ERR_UNSUPPORTED,
                    "pat_heat3d_run_dev: fused windows are 8, 16 or 32 rows");
        t = pat::HeatTile{1, 64, 4, by == 8 ? 2 : (by == 16 ? 4 : 8), zc > 0 ? zc : (int)std::min<int64_t>(64, nz)};
    } else {
        t.fused = 0;
        default_tile(nx, ny, nz, &t.bx, &t.by, &t.ry, &t.zc);
        // 'by' at the C ABI is the tile height in rows; split it into thread
        // rows and rows per thread using the instantiated shapes.
        const int tbx = bx > 0 ? bx : t.bx;
        if (by > 0 || tbx != t.bx) {
            bool found = false;
            for (const auto &sh : kShapes)
                if (sh.bx == tbx && (by <= 0 || sh.by * sh.ry == by)) {
                    t.by = sh.by;
                    t.ry = sh.ry;
                    found = true;
                    break;
                }
            PAT_REQUIRE(found, PAT_ERR_UNSUPPORTED, "pat_heat3d_run_dev: tile %dx%d not instantiated", tbx, by);
        }
        t.bx = tbx;
        if (zc > 0) t.zc = zc;
    }
    const int v = variant == PAT_STENCIL_GENERIC ? PAT_STENCIL_GENERIC : PAT_STENCIL_LDS25D;
    float *const buf[2] = {a, b};
    int end = 0;
    PAT_TRY(pat::heat3d_steps(buf, 0, nx, ny, nz, alpha, t, v, nsteps, st, &end));
    if (result_in_b) *result_in_b = end;
    return PAT_OK;
}
