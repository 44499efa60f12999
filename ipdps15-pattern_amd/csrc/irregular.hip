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
#include <type_traits>
#include <mutex>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "pat_common.hpp"
#include "pat_device.hpp"

namespace {

using pat::kWave;

// ------------------------------------------------------------------ CSR build

// p[0..n) = v, and dst[0..n) = src[0..n): the CSR builds clear and copy
// their counters with kernels rather than hipMemsetAsync/hipMemcpyAsync, so
// a build recorded into a captured graph holds kernel nodes only.
__global__ void fill_i32(int *__restrict__ p, int64_t n, int v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void copy_i32(int *__restrict__ dst, const int *__restrict__ src, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256 + 1, 4096); }

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

// Value held by lane ^ M (M < 64), by the cheapest cross-lane form: DPP
// quad permutes and row mirrors for 1, 2, 3, 7, 15, a bit-mode swizzle
// inside 32 lanes for the other masks below 32, a permute otherwise.
template <int M>
__device__ __forceinline__ unsigned xor_lane(unsigned v)
{
    const int x = (int)v;
    if constexpr (M == 1) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // [1,0,3,2]
    else if constexpr (M == 2) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]
    else if constexpr (M == 3) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0x1B, 0xF, 0xF, false);  // [3,2,1,0]
    else if constexpr (M == 7) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);  // half mirror
    else if constexpr (M == 15) return (unsigned)__builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false);  // mirror
    else if constexpr (M < 32) return (unsigned)__builtin_amdgcn_ds_swizzle(x, 0x1f | (M << 10));
    else return (unsigned)__shfl_xor(x, M, kWave);
}

// Rows sorted per wavefront by csr_row_sort.
constexpr int kSortRows = 4;

// Compare-exchange of every row's value `first` (0: entry lane, 1: entry
// 64 + lane) with lane ^ M's; the lane whose bit LOW is clear keeps the
// smaller value.
template <int M, int LOW>
__device__ __forceinline__ void sort_cx(unsigned (&a)[2 * kSortRows], int first, int lane)
{
#pragma unroll
    for (int t = 0; t < kSortRows; t++) {
        const unsigned v = a[2 * t + first], o = xor_lane<M>(v);
        a[2 * t + first] = (lane & LOW) ? max(v, o) : min(v, o);
    }
}

template <int J>
__device__ __forceinline__ void sort_half_clean(unsigned (&a)[2 * kSortRows], int lane)
{
    if constexpr (J >= 1) {
        sort_cx<J, J>(a, 0, lane);
        sort_cx<J, J>(a, 1, lane);
        sort_half_clean<J / 2>(a, lane);
    }
}

// Merge step K of the bitonic network in its flip form (Batcher's network
// with the first comparison of each merge mirrored, so every comparison
// sorts ascending): element i against i ^ (K - 1), then half-cleaners at
// distances K/4 .. 1.  Element i < 64 is value 0 of lane i, element 64 + i
// value 1 of lane i.
template <int K>
__device__ __forceinline__ void sort_merge(unsigned (&a)[2 * kSortRows], int lane)
{
    if constexpr (K <= kWave) {
        sort_cx<K - 1, K / 2>(a, 0, lane);
        sort_cx<K - 1, K / 2>(a, 1, lane);
    } else {  // K = 128: element lane against element 127 - lane (value 1 of lane ^ 63)
#pragma unroll
        for (int t = 0; t < kSortRows; t++) {
            const unsigned o1 = xor_lane<63>(a[2 * t + 1]), o0 = xor_lane<63>(a[2 * t]);
            a[2 * t] = min(a[2 * t], o1);
            a[2 * t + 1] = max(a[2 * t + 1], o0);
        }
    }
    sort_half_clean<K / 4>(a, lane);
}

// kSortRows rows per wavefront, each row of at most 128 columns sorted by
// the bitonic network held two values per lane (padding sorts last); the
// rows' pointers and values are all loaded before the networks run, which
// interleave.  Longer rows take the rank sort afterwards.  Columns of a row
// are distinct, so any correct sort gives the rank sort's order.  The
// unsorted row is tmp[rowptr[row] ...] (slot == 0) or tmp[row*slot ...]
// (rows of at most slot entries; longer ones are skipped).
__global__ __launch_bounds__(256) void csr_row_sort(const int *__restrict__ rowptr, int64_t N,
                                                    const int *__restrict__ tmp, int *__restrict__ cols, int slot)
{
    const int64_t row0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave * kSortRows;
    const int lane = threadIdx.x % kWave;
    if (row0 >= N) return;  // whole wavefronts
    // lane t <= kSortRows holds rowptr[row0 + t] (clamped to rowptr[N])
    const int rpl = rowptr[min(row0 + min(lane, kSortRows), N)];
    int b[kSortRows], n[kSortRows];
    unsigned a[2 * kSortRows];
#pragma unroll
    for (int t = 0; t < kSortRows; t++) {
        b[t] = __shfl(rpl, t, kWave);
        n[t] = row0 + t < N ? __shfl(rpl, t + 1, kWave) - b[t] : 0;
        if (slot && n[t] > slot) n[t] = 0;  // never produced by a slot pass
        const int m = n[t] <= 2 * kWave ? n[t] : 0;  // longer rows: rank sort below
        const int *__restrict__ src = tmp + (slot ? (row0 + t) * slot : (int64_t)b[t]);
        a[2 * t] = lane < m ? (unsigned)src[lane] : 0xffffffffu;
        a[2 * t + 1] = kWave + lane < m ? (unsigned)src[kWave + lane] : 0xffffffffu;
    }
    sort_merge<2>(a, lane);
    sort_merge<4>(a, lane);
    sort_merge<8>(a, lane);
    sort_merge<16>(a, lane);
    sort_merge<32>(a, lane);
    sort_merge<64>(a, lane);
    sort_merge<128>(a, lane);
#pragma unroll
    for (int t = 0; t < kSortRows; t++) {
        if (n[t] > 2 * kWave) {  // uniform per wavefront
            const int *__restrict__ src = tmp + (slot ? (row0 + t) * slot : (int64_t)b[t]);
            for (int k = lane; k < n[t]; k += kWave) {
                const int v = src[k];
                int rank = 0;
                for (int q = 0; q < n[t]; q++) {
                    const int u = src[q];
                    rank += (u < v) || (u == v && q < k);
                }
                cols[b[t] + rank] = v;
            }
        } else {
            if (lane < n[t]) cols[b[t] + lane] = (int)a[2 * t];
            if (kWave + lane < n[t]) cols[b[t] + kWave + lane] = (int)a[2 * t + 1];
        }
    }
}

// Grid of csr_row_sort over N rows.
static inline unsigned row_sort_grid(int64_t N)
{
    return (unsigned)(((N + kSortRows - 1) / kSortRows * kWave + 255) / 256);
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

// One periodic cube of a (possibly batched) build: its atoms are
// [atom0, atom1) of the concatenated positions and its nb^3 bins are
// [cell0, cell0 + nb^3) of the concatenated bin arrays.
struct NbSys {
    float L, hL, inv;  // box, half box, nb / L
    int nb, cell0, atom0, atom1, pad;
};

__device__ __forceinline__ int bin_of(float x, float inv, int nb)
{
    int c = (int)floorf(x * inv);
    return c < 0 ? 0 : (c >= nb ? nb - 1 : c);
}

// Index of the system holding atom i (by atom0) or bin q (by cell0).
template <bool BY_CELL>
__device__ __forceinline__ int nb_sys_of(const NbSys *__restrict__ sys, int nsys, int v)
{
    int lo = 0, hi = nsys;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((BY_CELL ? sys[mid].cell0 : sys[mid].atom0) <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void nb_bin_atoms(const float4 *__restrict__ pos, int N, const NbSys *__restrict__ sys, int nsys,
                             int *__restrict__ cell, int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const NbSys S = sys[nb_sys_of<false>(sys, nsys, i)];
    const float4 p = pos[i];
    const int c = S.cell0 + bin_of(p.x, S.inv, S.nb) + S.nb * (bin_of(p.y, S.inv, S.nb) + S.nb * bin_of(p.z, S.inv, S.nb));
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

// The same test where every coordinate difference is known to lie in
// [-L/2, L/2], on which the minimum image returns its argument unchanged.
__device__ __forceinline__ bool nb_within_near(float4 a, float4 b, const BinGeom &g)
{
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float r2 = dx * dx + dy * dy;
    r2 = r2 + dz * dz;
    return r2 < g.cut2;
}

// Candidate slots per lane of nb_cells: a wavefront holds 64 * kNbSlots
// atoms of a bin's 27-bin neighbourhood in registers at a time (about 510
// at the FCC 0.8442 density and 2.8-wide bins; larger ones take more
// passes over the bin's atoms).
constexpr int kNbSlots = 12;

// Entries per atom slot of the one-pass neighbour build (rows up to this
// long take the register bitonic sort).
constexpr int kNbSlot = 128;

enum NbMode {
    kNbCount = 0,  // deg[i] = row length
    kNbFill = 1,   // row into tmp[rowptr[i] ...]
    kNbSlotFill = 2,  // deg[i] = row length and the row into tmp[i*kNbSlot ...] (at most kNbSlot
                      // entries; *overflow = 1 when a row is longer)
};

// One wavefront per bin q: the atoms of the 27 bins around q (periodic
// wrap; nb >= 3 keeps them distinct) are laid out as one flat candidate
// range, gathered into registers slice by slice, and every atom i of q is
// tested against them, all lanes at once.  Rows are written compacted by
// ballot (order within a row is left to the row sort), never past their
// end (row pointer or slot).
template <int MODE>
__global__ __launch_bounds__(256) void nb_cells(const float4 *__restrict__ pos, const NbSys *__restrict__ sys,
                                                int nsys, int ncell, float cut2, const int *__restrict__ cstart,
                                                const int *__restrict__ catoms, int *__restrict__ deg,
                                                const int *__restrict__ rowptr, int *__restrict__ tmp,
                                                int *__restrict__ overflow)
{
    constexpr bool FILL = MODE != kNbCount;
    const int q = blockIdx.x * 4 + pat::wave_in_block();
    const int lane = threadIdx.x % kWave;
    if (q >= ncell) return;  // whole wavefronts
    const NbSys S = sys[nb_sys_of<true>(sys, nsys, q)];
    const int nb = S.nb, local = q - S.cell0;
    const int cx = local % nb, cy = (local / nb) % nb, cz = local / (nb * nb);
    // lane k < 27: neighbour bin k, its first atom slot and atom count
    int ns = 0, nc = 0;
    if (lane < 27) {
        const int qx = (cx + lane % 3 - 1 + nb) % nb;
        const int qy = (cy + (lane / 3) % 3 - 1 + nb) % nb;
        const int qz = (cz + lane / 9 - 1 + nb) % nb;
        const int nq = S.cell0 + qx + nb * (qy + nb * qz);
        ns = cstart[nq];
        nc = cstart[nq + 1] - ns;
    }
    int incl = nc;  // inclusive scan over lanes 0..31
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const int t = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += t;
    }
    const int excl = incl - nc;
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 26, kWave));  // uniform
    BinGeom g;
    g.L = S.L;
    g.hL = S.hL;
    g.cut2 = cut2;
    g.inv = S.inv;
    g.nb = nb;
    const int a_beg = cstart[q], a_end = cstart[q + 1];
    for (int a0 = a_beg; a0 < a_end; a0 += kWave) {
        const int na = min(kWave, a_end - a0);
        int acc = 0;  // lane l: entries found so far for atom slot a0 + l
        // lane l: atom a0 + l, its position and row start / end, loaded
        // together and broadcast in the atom loop
        int my_i = 0, my_w = 0, my_wend = 0;
        float4 my_p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane < na) {
            my_i = catoms[a0 + lane];
            my_p = pos[my_i];
            if (MODE == kNbFill) {
                my_w = rowptr[my_i];
                my_wend = rowptr[my_i + 1];
            }
        }
        for (int base = 0; base < total; base += kWave * kNbSlots) {
            float4 pj[kNbSlots];
            int jj[kNbSlots];
#pragma unroll
            for (int r = 0; r < kNbSlots; r++) {
                jj[r] = -1;
                pj[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            // slots past the last candidate stay empty and are not visited
            // (uniform over the wavefront)
            const int slots = min(kNbSlots, (total - base + kWave - 1) / kWave);
#pragma unroll
            for (int r = 0; r < kNbSlots; r++) {
                if (r < slots) {
                const int f = base + r * kWave + lane;
                // largest k <= 26 with excl_k <= f (excl is nondecreasing,
                // so bin k is non-empty and holds f when f < total)
                int k = 0;
#pragma unroll
                for (int step = 16; step > 0; step >>= 1) {
                    const int e = __shfl(excl, min(k + step, 26), kWave);
                    k = (k + step <= 26 && e <= f) ? k + step : k;
                }
                const int ex = __shfl(excl, k, kWave), s0 = __shfl(ns, k, kWave);
                if (f < total) {
                    jj[r] = catoms[s0 + f - ex];
                    pj[r] = pos[jj[r]];
                }
                }
            }
            // When one pass holds every candidate (the bin's own atoms among
            // them) and, per axis, the rounded span of their coordinates is
            // below L/2, the exact span is too (rounding is monotone), so
            // every difference lies in [-L/2, L/2] and the minimum image is
            // the identity: skipping it gives the same bits.  Bins whose
            // neighbourhood wraps around the box keep it.
            bool near = false;
            if (total <= kWave * kNbSlots) {  // uniform per wavefront
                float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
#pragma unroll
                for (int r = 0; r < kNbSlots; r++) {
                    if (jj[r] >= 0) {
                        lx = fminf(lx, pj[r].x), hx = fmaxf(hx, pj[r].x);
                        ly = fminf(ly, pj[r].y), hy = fmaxf(hy, pj[r].y);
                        lz = fminf(lz, pj[r].z), hz = fmaxf(hz, pj[r].z);
                    }
                }
#pragma unroll
                for (int d = kWave / 2; d > 0; d >>= 1) {
                    lx = fminf(lx, __shfl_xor(lx, d, kWave)), hx = fmaxf(hx, __shfl_xor(hx, d, kWave));
                    ly = fminf(ly, __shfl_xor(ly, d, kWave)), hy = fmaxf(hy, __shfl_xor(hy, d, kWave));
                    lz = fminf(lz, __shfl_xor(lz, d, kWave)), hz = fmaxf(hz, __shfl_xor(hz, d, kWave));
                }
                near = hx - lx < g.hL && hy - ly < g.hL && hz - lz < g.hL;
            }
            // Atom a's values are broadcast from lane a - a0 (a uniform
            // index: scalar reads of one lane, no LDS exchange).
            auto from_lane = [&](auto v, int l) {
                if constexpr (std::is_same_v<decltype(v), float>)
                    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
                else
                    return __builtin_amdgcn_readlane(v, l);
            };
            auto test_atoms = [&](auto near_c) {
                constexpr bool NEAR = decltype(near_c)::value;
                for (int a = a0; a < a0 + na; a++) {
                    const int l = a - a0;
                    const int i = from_lane(my_i, l);
                    const float4 pi = make_float4(from_lane(my_p.x, l), from_lane(my_p.y, l), from_lane(my_p.z, l), 0.f);
                    int w = 0, wend = 0;
                    if (MODE == kNbFill) {
                        w = from_lane(my_w, l) + from_lane(acc, l);
                        wend = from_lane(my_wend, l);  // the counting pass sized the row; never write past it
                    } else if (MODE == kNbSlotFill) {
                        w = i * kNbSlot + from_lane(acc, l);
                        wend = (i + 1) * kNbSlot;
                    }
                    int found = 0;
#pragma unroll
                    for (int r = 0; r < kNbSlots; r++) {
                        if (r < slots) {
                            // all three conditions evaluated (no branch around the distance test)
                            const bool hit = (jj[r] >= 0) & (jj[r] != i) &
                                             (NEAR ? nb_within_near(pi, pj[r], g) : nb_within(pi, pj[r], g));
                            const unsigned long long m = __ballot(hit);
                            if (FILL && hit) {
                                // hits in lanes below this one
                                const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                                const int o = w + found + below;
                                if (o < wend) tmp[o] = jj[r];
                            }
                            found += __popcll(m);
                        }
                    }
                    if (lane == l) acc += found;
                }
            };
            if (near)
                test_atoms(std::true_type{});
            else
                test_atoms(std::false_type{});
        }
        if (MODE != kNbFill && lane < na) deg[my_i] = acc;
        if (MODE == kNbSlotFill && lane < na && acc > kNbSlot) *overflow = 1;
    }
}

// Small boxes (two bins per side, L >= 2 cut: every pair has at most one
// image within the cutoff): one wavefront per atom i tests every atom j in
// increasing index order with the same distance test as nb_cells, so each
// row comes out already sorted.  Count mode writes deg[i]; fill mode writes
// the row at rowptr[i], never past rowptr[i + 1].
constexpr int kNbBruteMax = 1 << 16;  // atoms: O(N^2) tests per build

template <int MODE>
__global__ __launch_bounds__(256) void nb_brute(const float4 *__restrict__ pos, int N, BinGeom g,
                                                int *__restrict__ deg, const int *__restrict__ rowptr,
                                                int *__restrict__ cols)
{
    const int i = blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (i >= N) return;  // whole wavefronts
    const float4 pi = pos[i];
    const int w = MODE == kNbFill ? rowptr[i] : 0;
    const int wend = MODE == kNbFill ? rowptr[i + 1] : 0;
    int found = 0;
    for (int j0 = 0; j0 < N; j0 += kWave) {
        const int j = j0 + lane;
        const bool hit = j < N && j != i && nb_within(pi, pos[min(j, N - 1)], g);
        const unsigned long long m = __ballot(hit);
        if (MODE == kNbFill && hit) {
            const int o = w + found + __popcll(m & ((1ull << lane) - 1ull));
            if (o < wend) cols[o] = j;
        }
        found += __popcll(m);
    }
    if (MODE == kNbCount && lane == 0) deg[i] = found;
}

// ------------------------------------------------------------------ force

struct LJParams {
    float L, hL;   // box side and half side (pbc)
    float rc2;     // cutoff squared
    float dt, hdt; // time step and half step (hdt = 0.5f*dt rounded once)
    int pbc;
    float W0;      // least float w with w*w >= rc2 (rounded product)
    int *inbox;    // 1 while every position lies in [0, L] (may be null)
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

// The same refined reciprocal without the range guard, for r^2 inside a
// cutoff (rc^2 < 2^31).  Below 2^-126 -- two nodes closer than 2^-63 -- the
// serial code's r^-4 already overflows and its force is NaN, and so is this
// one's (r^-1 is then >= 2^125 or inf, r^-4 inf or NaN).
__device__ __forceinline__ float rcp_in_cutoff(float b)
{
    const float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    return __builtin_fmaf(e, y, y);
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

// The same move when every position lies in [0, L]: a row with x_i < L/2
// can only see d = x_i - x_j < -L/2 (x_j > L is impossible), one with
// x_i >= L/2 only d > L/2, so the move is d + s with s = +L or -L fixed
// per row (the image sign), and d + (-L) is exactly d - L.
__device__ __forceinline__ float min_image_dir(float d, float s, float hL)
{
    return fabsf(d) > hL ? d + s : d;
}

struct ImageSign {
    float x, y, z;  // +L or -L, see min_image_dir
};

__device__ __forceinline__ ImageSign image_sign(float4 pi, const LJParams &p)
{
    return ImageSign{pi.x < p.hL ? p.L : -p.L, pi.y < p.hL ? p.L : -p.L, pi.z < p.hL ? p.L : -p.L};
}

// Force on i from j (both float4 positions), the minimum image taken on the
// axes set in WM (bit 0 x, 1 y, 2 z): in the general two-sided form when
// bit 3 is set, else in the per-row form of min_image_dir (signs sg).  Branch-free: beyond the cutoff the
// reciprocal is replaced by 0, so F = 0 and each component is +0 or -0;
// added to a running sum that started at +0 either leaves every bit of it
// unchanged (a sum started at +0 is never -0 under round-to-nearest), so the
// sums equal the serial code's, which skips such pairs.
// rrd7 - 0.5*rrd4 is one fma: 0.5*rrd4 is exact for rrd4 >= 2^-125, which
// holds inside any cutoff with rc^2 < 2^31.
template <int WM>
__device__ __forceinline__ void pair_force(float4 pi, float4 pj, const LJParams &p, const ImageSign &sg, float &fx,
                                           float &fy, float &fz)
{
    float dx = pi.x - pj.x;
    float dy = pi.y - pj.y;
    float dz = pi.z - pj.z;
    if (WM & 8) {
        if (WM & 1) dx = min_image_sel(dx, p.L, p.hL);
        if (WM & 2) dy = min_image_sel(dy, p.L, p.hL);
        if (WM & 4) dz = min_image_sel(dz, p.L, p.hL);
    } else {
        if (WM & 1) dx = min_image_dir(dx, sg.x, p.hL);
        if (WM & 2) dy = min_image_dir(dy, sg.y, p.hL);
        if (WM & 4) dz = min_image_dir(dz, sg.z, p.hL);
    }
    float rd = dx * dx + dy * dy;
    rd = rd + dz * dz;
    const float rrd = rd < p.rc2 ? rcp_in_cutoff(rd) : 0.0f;
    const float rrd2 = rrd * rrd;
    const float rrd4 = rrd2 * rrd2;
    const float rrd6 = rrd2 * rrd4;
    const float rrd7 = rrd6 * rrd;
    const float F = 48.0f * __builtin_fmaf(-0.5f, rrd4, rrd7);
    fx = dx * F;
    fy = dy * F;
    fz = dz * F;
}

__device__ __forceinline__ void pair_force(float4 pi, float4 pj, const LJParams &p, float &fx, float &fy, float &fz)
{
    const ImageSign none{0.0f, 0.0f, 0.0f};
    if (p.pbc) pair_force<15>(pi, pj, p, none, fx, fy, fz);
    else pair_force<0>(pi, pj, p, none, fx, fy, fz);
}

// Axes on which a wavefront's pairs may need the minimum image (the WM of
// pair_force).  When every position lies in [0, L] (the inbox flag) and all
// 64 rows' own coordinate on axis a lies in [W, L - W], W = W0 + L*2^-20,
// no pair of these rows needs the image on a, or needs it only where both
// forms lie beyond the cutoff: |d| > L/2 >= W0 without the move, and
// L - |d| >= L - max(x_i, L - x_i) >= W minus a few roundings of L (each
// under L*2^-24) >= W0 with it -- both give r^2 >= rc^2, a zero force.  The
// remaining axes take the per-row form.  Without the flag: all axes, two-
// sided form.  The sentinel node stays beyond the cutoff in every form.
template <bool SENT>
__device__ __forceinline__ int axes_to_wrap(float4 pi, const LJParams &p)
{
    if (!p.pbc) return 0;
    if (!SENT || !p.inbox || !*p.inbox) return 15;
    const float W = p.W0 + p.L * 0x1p-20f, hi = p.L - W;
    int wm = 0;
    if (!__all(pi.x >= W && pi.x <= hi)) wm |= 1;
    if (!__all(pi.y >= W && pi.y <= hi)) wm |= 2;
    if (!__all(pi.z >= W && pi.z <= hi)) wm |= 4;
    return wm;
}

// Class of a node for the slice order: bit a set when its coordinate on
// axis a lies within W of a face (W as in axes_to_wrap).  Slices filled
// with one class take the minimum image on exactly those axes.
__device__ __forceinline__ int row_class(float4 x, float L, float W0)
{
    const float W = W0 + L * 0x1p-20f, hi = L - W;
    return (x.x >= W && x.x <= hi ? 0 : 1) | (x.y >= W && x.y <= hi ? 0 : 2) | (x.z >= W && x.z <= hi ? 0 : 4);
}

// Periodic wrap of a drifted position; clears the inbox flag when a
// coordinate still lies outside [0, L] (a step longer than the box).
__device__ __forceinline__ void drift_wrap(float4 &x, const LJParams &p)
{
    if (!p.pbc) return;
    x.x = pat::wrap_coord(x.x, p.L);
    x.y = pat::wrap_coord(x.y, p.L);
    x.z = pat::wrap_coord(x.z, p.L);
    if (p.inbox && !(x.x >= 0.0f && x.x <= p.L && x.y >= 0.0f && x.y <= p.L && x.z >= 0.0f && x.z <= p.L))
        *p.inbox = 0;
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
    kKickDrift = 2,   // v += hdt*f; v += hdt*f; x_out = wrap(x + dt*v) (f itself not stored)
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
                drift_wrap(x, p);
                pos_out[row] = x;
            }
            vel[row] = v;
        }
    }
}

// Shared epilogue: the fused velocity-Verlet update and the force write,
// v being the node's velocity (loaded by the caller; unused for
// kForceOnly).  A kick-drift step's force only feeds its own two half
// kicks, so it is not stored; a force evaluation or the closing kKick step
// of a run stores it (the next run's first half kick reads it).
__device__ __forceinline__ void lj_finish(int row, float4 pi, float4 v, float fx, float fy, float fz,
                                          const LJParams &p, float4 *__restrict__ force, float4 *__restrict__ vel,
                                          float4 *__restrict__ pos_out, int mode)
{
    if (mode != kKickDrift) force[row] = make_float4(fx, fy, fz, 0.0f);
    if (mode == kForceOnly) return;
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
        drift_wrap(x, p);
        pos_out[row] = x;
    }
    vel[row] = v;
}

__device__ __forceinline__ void lj_epilogue(int row, float4 pi, float fx, float fy, float fz, const LJParams &p,
                                            float4 *__restrict__ force, float4 *__restrict__ vel,
                                            float4 *__restrict__ pos_out, int mode)
{
    const float4 v = mode != kForceOnly ? vel[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    lj_finish(row, pi, v, fx, fy, fz, p, force, vel, pos_out, mode);
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

// Slice s holds rows perm[64s .. 64s+63] (rows 64s.. when perm is null;
// entries >= N are empty lanes).
__device__ __forceinline__ int slice_row(const int *__restrict__ perm, int i)
{
    return perm ? perm[i] : i;
}

__global__ void sell_width(const int *__restrict__ rowptr, const int *__restrict__ perm, int N, int nsl,
                           int *__restrict__ width)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsl) return;
    int w = 0;
    for (int i = s * kWave; i < s * kWave + kWave; i++) {
        const int r = slice_row(perm, i);
        if (r < N) w = max(w, rowptr[r + 1] - rowptr[r]);
    }
    width[s] = w;
}

// Columns per LDS tile of sell_fill.
constexpr int kFillK = 32;

// Orders this wavefront's LDS writes before its later LDS reads (and the
// reads before later overwrites).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront per slice.  pad < 0: entries hold node indices and padding
// repeats the row's own index; otherwise entries hold byte offsets 16*j into
// the float4 position array and padding is 16*pad (a sentinel node placed
// beyond every cutoff).  The slice is transposed through LDS 32 columns at a
// time: half-wavefronts read 32 consecutive entries of one row (coalesced),
// then every lane writes its row's entry k (one 256-byte store per k).
__global__ __launch_bounds__(256) void sell_fill(const int *__restrict__ rowptr, const int *__restrict__ cols,
                                                 const int *__restrict__ perm, int N, int nsl,
                                                 const int *__restrict__ sptr, int *__restrict__ scol, int pad)
{
    __shared__ int tile[4][kWave][kFillK + 1];
    const int wv = threadIdx.x / kWave;
    const int s = blockIdx.x * 4 + wv;
    const int lane = threadIdx.x % kWave;
    if (s >= nsl) return;  // whole wavefronts
    const int row = slice_row(perm, s * kWave + lane);
    const int b = row < N ? rowptr[row] : 0, e = row < N ? rowptr[row + 1] : 0;
    const int fill = pad < 0 ? min(row, N - 1) : (int)((uint32_t)pad << 4);
    const int sh = pad < 0 ? 0 : 4;
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    const int half = lane >> 5, t = lane & 31;
    const int last = rowptr[N] - 1;  // last entry of the CSR (-1: none)
    for (int k0 = 0; k0 < w; k0 += kFillK) {
        if (last >= 0) {  // uniform per wavefront
            // every lane loads (at an index clamped into the CSR) and the
            // padding is selected afterwards, so the 32 loads are in flight
            // together instead of one branch and wait each
#pragma unroll
            for (int m = 0; m < kWave / 2; m++) {
                const int rr = 2 * m + half;  // lanes 0-31: row 2m, lanes 32-63: row 2m+1
                const int rb = __shfl(b, rr, kWave), re = __shfl(e, rr, kWave), rf = __shfl(fill, rr, kWave);
                const int k = rb + k0 + t;
                const int v = cols[min(k, last)];
                tile[wv][rr][t] = k < re ? (int)((uint32_t)v << sh) : rf;
            }
        } else {
            for (int m = 0; m < kWave / 2; m++) {
                const int rr = 2 * m + half;
                tile[wv][rr][t] = __shfl(fill, rr, kWave);
            }
        }
        wave_lds_sync();
        const int kn = min(kFillK, w - k0);
        for (int kk = 0; kk < kn; kk++) scol[base + (k0 + kk) * kWave + lane] = tile[wv][lane][kk];
        wave_lds_sync();
    }
}

// Slice order by row class (row_class) within windows of 256 rows: each
// block sorts its rows by class, stably (ballot ranks), so the window's
// interior rows fill whole slices of their own while every slice keeps
// rows from the same stretch of the node order (the same neighbourhood,
// and the same share of boundary rows per block of the grid).  Rows >= N
// sort last and become empty lanes (N) where they fall below total.
__global__ __launch_bounds__(256) void perm_block_sort(const float4 *__restrict__ pos, int N, int total, float L,
                                                       float W0, int pbc, int *__restrict__ perm)
{
    __shared__ int wc[4][9];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    int c = 8;
    if (i < N) {
        const float4 x = pos[i];
        c = row_class(x, pbc == 2 ? x.w : L, W0);
    }
    int rank = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const unsigned long long m = __ballot(c == k);
        if (lane == 0) wc[w][k] = __popcll(m);
        if (c == k) rank = __popcll(m & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    int off = 0;
    for (int k = 0; k < c; k++) off += wc[0][k] + wc[1][k] + wc[2][k] + wc[3][k];
    for (int v = 0; v < w; v++) off += wc[v][c];
    const int dst = blockIdx.x * 256 + off + rank;
    if (dst < total) perm[dst] = i < N ? i : N;
}

// Spatial slice order: nodes sorted by the Morton code of a bin about h
// wide (at most 1024 per axis), system by system, ties in node order, so a
// slice's 64 rows sit in a compact region and share one small gather
// neighbourhood.  Rows keep their own entries and summation order.
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// keys[i] = (system << 30) | morton(bin of node i); vals[i] = i.  The box
// side is box, or each node's w for pbc 2; sysoff (nsys + 1 offsets, null
// for one system) gives the systems' node ranges.
__global__ void morton_keys(const float4 *__restrict__ pos, int N, float box, int pbc, float h,
                            const int *__restrict__ sysoff, int nsys, uint64_t *__restrict__ keys,
                            int *__restrict__ vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 p = pos[i];
    const float L = pbc == 2 ? p.w : box;
    const int nb = (int)fminf(1024.0f, fmaxf(1.0f, floorf(L / h)));
    const float inv = (float)nb / L;
    const uint32_t bx = (uint32_t)min(max((int)floorf(p.x * inv), 0), nb - 1);
    const uint32_t by = (uint32_t)min(max((int)floorf(p.y * inv), 0), nb - 1);
    const uint32_t bz = (uint32_t)min(max((int)floorf(p.z * inv), 0), nb - 1);
    int s = 0;
    if (sysoff) {
        int lo = 0, hi = nsys;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sysoff[mid] <= i) lo = mid;
            else hi = mid;
        }
        s = lo;
    }
    keys[i] = ((uint64_t)s << 30) | (uint64_t)(spread3(bx) | (spread3(by) << 1) | (spread3(bz) << 2));
    vals[i] = i;
}

// perm[k] = sorted node k for k < N, the empty lane N up to total.
__global__ void perm_pad(const int *__restrict__ sorted, int N, int total, int *__restrict__ perm)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < total) perm[k] = k < N ? sorted[k] : N;
}

// Neighbour position.  SENT layouts store the byte offset j*16 of node j
// (the handle keeps fewer than 2^28 nodes), the others the index.
template <bool SENT>
__device__ __forceinline__ float4 gather_pos(const float4 *__restrict__ pos, int j)
{
    if (SENT) return *(const float4 *)((const char *)pos + (uint32_t)j);
    return pos[j];
}

// Row sums of one slice, lane = row: entries k = 0 .. w-1 of the lane's row
// at c[64k], U entries' column reads and position gathers issued ahead of
// their arithmetic.  Padding entries either repeat the row's own index and
// are masked (SENT = false: entries at or past deg), or point at a sentinel
// node N whose position is beyond every cutoff (SENT = true): its pair force
// is +0 or -0, and adding it to a sum that started at +0 leaves every bit
// unchanged, so no mask is needed.
template <int U, bool SENT, int WM>
__device__ __forceinline__ void sell_rows(const float4 *__restrict__ pos, const int *__restrict__ c, int w, int deg,
                                          float4 pi, const LJParams &p, const ImageSign &sg, float &fx, float &fy,
                                          float &fz)
{
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
            pair_force<WM>(pi, q[u], p, sg, ax, ay, az);
            if (SENT || k + u < deg) {
                fx = fx + ax;
                fy = fy + ay;
                fz = fz + az;
            }
        }
    }
    for (; k < w; k++) {
        float ax, ay, az;
        pair_force<WM>(pi, gather_pos<SENT>(pos, c[k * kWave]), p, sg, ax, ay, az);
        if (SENT || k < deg) {
            fx = fx + ax;
            fy = fy + ay;
            fz = fz + az;
        }
    }
}

// The same sums, software-pipelined: while a group of U pairs is evaluated,
// the positions of the next group and the column indices of the one after
// are already in flight (two register sets used alternately), so a lane
// always has a group's memory latency to hide behind arithmetic.  Past the
// slice width a lane reads a valid dummy column (its own row, masked, or
// the sentinel).
template <int U, bool SENT, int WM>
__device__ __forceinline__ void sell_rows_pipe(const float4 *__restrict__ pos, const int *__restrict__ c, int w,
                                               int deg, int dummy, float4 pi, const LJParams &p, const ImageSign &sg,
                                               float &fx, float &fy, float &fz)
{
    auto col_at = [&](int k) { return k < w ? c[k * kWave] : dummy; };
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
            pair_force<WM>(pi, qc[u], p, sg, ax, ay, az);
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
}

// In-order row sums over the sliced layout, one wavefront per slice and one
// lane per row (PIPE: the software-pipelined loop).  The loop is
// instantiated for the axis sets {} (no minimum image), {x} and {x, y, z}
// and each wavefront takes the least one covering axes_to_wrap.
// Bit-identical to lj_csr_seq.
template <int U, bool SENT, bool PIPE>
__global__ __launch_bounds__(256) void lj_sell(const float4 *__restrict__ pos, const int *__restrict__ rowptr,
                                               const int *__restrict__ sptr, const int *__restrict__ scol,
                                               const int *__restrict__ perm, int N, LJParams p,
                                               float4 *__restrict__ force, float4 *__restrict__ vel,
                                               float4 *__restrict__ pos_out, int mode)
{
    const int nsl = (N + kWave - 1) / kWave;
    const int s = pat::xcd_swizzle(blockIdx.x, gridDim.x) * 4 + pat::wave_in_block();
    if (s >= nsl) return;  // whole wavefronts
    const int lane = threadIdx.x % kWave;
    const int row = slice_row(perm, s * kWave + lane);
    const int self = min(row, N - 1);
    const float4 pi = pos[self];
    if (p.pbc == 2) {
        p.L = pi.w;
        p.hL = 0.5f * pi.w;
    }
    const int deg = SENT ? 0 : (row < N ? rowptr[row + 1] - rowptr[row] : 0);
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    const int *__restrict__ c = scol + base + lane;
    const int dummy = SENT ? (int)((uint32_t)N << 4) : self;
    // the velocity is read ahead of the loop, its latency hidden by the sums
    const float4 v0 = mode != kForceOnly && row < N ? vel[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int wm = axes_to_wrap<SENT>(pi, p);
    const ImageSign sg = image_sign(pi, p);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    auto rows = [&](auto tag) {
        constexpr int WM = decltype(tag)::value;
        if (PIPE) sell_rows_pipe<U, SENT, WM>(pos, c, w, deg, dummy, pi, p, sg, fx, fy, fz);
        else sell_rows<U, SENT, WM>(pos, c, w, deg, pi, p, sg, fx, fy, fz);
    };
    if (wm == 0) {
        rows(std::integral_constant<int, 0>());
    } else if (wm == 15 || !SENT) {
        rows(std::integral_constant<int, 15>());
    } else if constexpr (SENT) {
        switch (wm) {
        case 1: rows(std::integral_constant<int, 1>()); break;
        case 2: rows(std::integral_constant<int, 2>()); break;
        case 3: rows(std::integral_constant<int, 3>()); break;
        case 4: rows(std::integral_constant<int, 4>()); break;
        case 5: rows(std::integral_constant<int, 5>()); break;
        case 6: rows(std::integral_constant<int, 6>()); break;
        default: rows(std::integral_constant<int, 7>()); break;
        }
    }
    if (row < N) lj_finish(row, pi, v0, fx, fy, fz, p, force, vel, pos_out, mode);
}

// ------------------------------------------------------------ halo layout
//
// The sentinel layout's slices read every neighbour position from global
// memory, one 16-byte gather per pair, and the texture data path that
// returns those gathers, not the arithmetic, bounds the kernel.  The halo
// layout stages positions through LDS instead: for each block of 256 rows
// (4 slices) the distinct neighbour nodes are listed once (hidx, sorted by
// their offset from the block in the periodic index space), the block
// loads their positions into LDS at its start, and the slices' entries
// become 16-bit LDS byte offsets (local index * 16; the sentinel is the
// slot after the list).  Same pairs in the same order: the same bits.
//
// A block's neighbour set is found with a bitmap over a window of the node
// index space centred on the block (modular: columns across the periodic
// index wrap fall in it too), at most kHaloBits wide; a block whose window
// or list is larger fails the build and the handle keeps the global form.
constexpr int kHaloBits = 1 << 18;
constexpr int kHaloCap = 4095;  // list entries per block (16-bit byte offsets, sentinel slot included)
constexpr int kHaloWords = kHaloBits / 32;

// Signed offset of column col from center in the periodic index space
// [0, N), in (-N/2, N/2].
__device__ __forceinline__ int mod_delta(int col, int center, int N)
{
    int d = col - center;
    if (d > N / 2) d -= N;
    else if (d <= -((N + 1) / 2)) d += N;
    return d;
}

// Window of block b: base and width (in bits) from the block's largest
// column offset m; shared by the count and fill passes.
struct HaloWindow {
    int base, width;
};

__device__ HaloWindow halo_window(int m, int center, int N)
{
    if (2 * (int64_t)m + 1 >= N) return HaloWindow{0, N};
    int base = center - m;
    if (base < 0) base += N;
    return HaloWindow{base, 2 * m + 1};
}

// Builds the block's bitmap in LDS (words [0, ceil(width/32))) and returns
// its window; win.width = 0 when the window is too wide.
__device__ HaloWindow halo_bitmap(const int *__restrict__ rowptr, const int *__restrict__ cols,
                                  const int *__restrict__ perm, int N, int nsl, uint32_t *bits, int *red)
{
    const int b = blockIdx.x, t = threadIdx.x;
    const int s = b * 4 + t / kWave;
    const int row = s < nsl ? slice_row(perm, s * kWave + t % kWave) : N;
    const int center = min(b * 256 + 128, N - 1);
    int m = 0;
    const int e0 = row < N ? rowptr[row] : 0, e1 = row < N ? rowptr[row + 1] : 0;
    for (int k = e0; k < e1; k++) m = max(m, abs(mod_delta(cols[k], center, N)));
    if (t == 0) *red = 0;
    __syncthreads();
    atomicMax(red, m);
    __syncthreads();
    const HaloWindow win = halo_window(*red, center, N);
    if (win.width > kHaloBits) return HaloWindow{0, 0};
    const int nw = (win.width + 31) / 32;
    for (int i = t; i < nw; i += 256) bits[i] = 0u;
    __syncthreads();
    for (int k = e0; k < e1; k++) {
        int x = cols[k] - win.base;
        if (x < 0) x += N;
        atomicOr(&bits[x >> 5], 1u << (x & 31));
    }
    __syncthreads();
    return win;
}

// Pass 1: hcount[b] = the block's distinct neighbours, or -1 when the block
// does not fit the halo form.
__global__ __launch_bounds__(256) void halo_count(const int *__restrict__ rowptr, const int *__restrict__ cols,
                                                  const int *__restrict__ perm, int N, int nsl,
                                                  int *__restrict__ hcount)
{
    __shared__ uint32_t bits[kHaloWords];
    __shared__ int red, tot;
    const HaloWindow win = halo_bitmap(rowptr, cols, perm, N, nsl, bits, &red);
    if (win.width == 0) {
        if (threadIdx.x == 0) hcount[blockIdx.x] = -1;
        return;
    }
    int c = 0;
    for (int i = threadIdx.x; i < (win.width + 31) / 32; i += 256) c += __popc(bits[i]);
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    atomicAdd(&tot, c);
    __syncthreads();
    if (threadIdx.x == 0) hcount[blockIdx.x] = tot < kHaloCap ? tot : -1;
}

// Pass 2: the block's list (hidx from hptr[b], in window order) and its
// slices' entries as LDS byte offsets.
__global__ __launch_bounds__(256) void halo_fill(const int *__restrict__ rowptr, const int *__restrict__ cols,
                                                 const int *__restrict__ perm, int N, int nsl,
                                                 const int *__restrict__ sptr, const int *__restrict__ hptr,
                                                 int *__restrict__ hidx, uint16_t *__restrict__ scol16)
{
    __shared__ uint32_t bits[kHaloWords];
    __shared__ int wbase[kHaloWords];
    __shared__ int red, scan[256], carry;
    const HaloWindow win = halo_bitmap(rowptr, cols, perm, N, nsl, bits, &red);
    const int b = blockIdx.x, t = threadIdx.x;
    const int nw = (win.width + 31) / 32;
    // exclusive prefix of the words' bit counts, 256 words at a time
    if (t == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nw; c0 += 256) {
        const int i = c0 + t;
        const int v = i < nw ? __popc(bits[i]) : 0;
        scan[t] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int a = t >= o ? scan[t - o] : 0;
            __syncthreads();
            scan[t] += a;
            __syncthreads();
        }
        if (i < nw) wbase[i] = carry + scan[t] - v;
        __syncthreads();
        if (t == 255) carry += scan[255];
        __syncthreads();
    }
    const int h0 = hptr[b], nh = hptr[b + 1] - h0;
    for (int i = t; i < nw; i += 256) {
        uint32_t m = bits[i];
        int r = wbase[i];
        while (m) {
            const int bit = __ffs(m) - 1;
            m &= m - 1;
            int g = win.base + i * 32 + bit;
            if (g >= N) g -= N;
            hidx[h0 + r++] = g;
        }
    }
    // entries of this thread's row: slice s = 4b + wave, lane = row slot
    const int s = b * 4 + t / kWave, lane = t % kWave;
    if (s >= nsl) return;
    const int row = slice_row(perm, s * kWave + lane);
    const int e0 = row < N ? rowptr[row] : 0, e1 = row < N ? rowptr[row + 1] : 0;
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    for (int k = 0; k < w; k++) {
        int local = nh;  // sentinel slot
        if (e0 + k < e1) {
            int x = cols[e0 + k] - win.base;
            if (x < 0) x += N;
            const uint32_t word = bits[x >> 5];
            local = wbase[x >> 5] + __popc(word & ((1u << (x & 31)) - 1u));
        }
        scol16[base + k * kWave + lane] = (uint16_t)(local << 4);
    }
}

// Row sums of one slice from the block's LDS positions (see sell_rows).
template <int U, int WM>
__device__ __forceinline__ void sell_rows_lds(const float4 *hpos, const uint16_t *__restrict__ c, int w, float4 pi,
                                              const LJParams &p, const ImageSign &sg, float &fx, float &fy,
                                              float &fz)
{
    auto at = [&](uint32_t off) { return *(const float4 *)((const char *)hpos + off); };
    int k = 0;
    for (; k + U <= w; k += U) {
        uint32_t j[U];
        float4 q[U];
#pragma unroll
        for (int u = 0; u < U; u++) j[u] = c[(k + u) * kWave];
#pragma unroll
        for (int u = 0; u < U; u++) q[u] = at(j[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            float ax, ay, az;
            pair_force<WM>(pi, q[u], p, sg, ax, ay, az);
            fx = fx + ax;
            fy = fy + ay;
            fz = fz + az;
        }
    }
    for (; k < w; k++) {
        float ax, ay, az;
        pair_force<WM>(pi, at(c[k * kWave]), p, sg, ax, ay, az);
        fx = fx + ax;
        fy = fy + ay;
        fz = fz + az;
    }
}

// The sliced kernel over the halo layout: the block's neighbour positions
// are loaded into LDS (dynamic, nh+1 float4) first, then each wavefront
// sums its slice as lj_sell does.  Bit-identical to lj_csr_seq.
template <int U>
__global__ __launch_bounds__(256) void lj_sell_halo(const float4 *__restrict__ pos, const int *__restrict__ sptr,
                                                    const uint16_t *__restrict__ scol16,
                                                    const int *__restrict__ hptr, const int *__restrict__ hidx,
                                                    const int *__restrict__ perm, int N, LJParams p,
                                                    float4 *__restrict__ force, float4 *__restrict__ vel,
                                                    float4 *__restrict__ pos_out, int mode)
{
    extern __shared__ float4 hpos[];
    const int nsl = (N + kWave - 1) / kWave;
    const int b = pat::xcd_swizzle(blockIdx.x, gridDim.x);
    const int h0 = hptr[b], nh = hptr[b + 1] - h0;
    // list entries (and the sentinel, node N, at entry nh) four per thread
    // at a time: the four index reads, then the four position reads, so
    // the loads overlap instead of queueing
    for (int i0 = threadIdx.x; i0 <= nh; i0 += 4 * 256) {
        int g[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = i0 + 256 * r;
            g[r] = i < nh ? hidx[h0 + i] : N;
        }
        float4 v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = pos[g[r]];
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (i0 + 256 * r <= nh) hpos[i0 + 256 * r] = v[r];
    }
    const int s = b * 4 + pat::wave_in_block();
    const int lane = threadIdx.x % kWave;
    const int row = s < nsl ? slice_row(perm, s * kWave + lane) : N;
    const int self = min(row, N - 1);
    const float4 pi = pos[self];
    const float4 v0 = mode != kForceOnly && row < N ? vel[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    if (s >= nsl) return;  // whole wavefronts, after the barrier
    if (p.pbc == 2) {
        p.L = pi.w;
        p.hL = 0.5f * pi.w;
    }
    const int base = sptr[s], w = (sptr[s + 1] - base) / kWave;
    const uint16_t *__restrict__ c = scol16 + base + lane;
    const int wm = axes_to_wrap<true>(pi, p);
    const ImageSign sg = image_sign(pi, p);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    auto rows = [&](auto tag) {
        constexpr int WM = decltype(tag)::value;
        sell_rows_lds<U, WM>(hpos, c, w, pi, p, sg, fx, fy, fz);
    };
    switch (wm) {
    case 0: rows(std::integral_constant<int, 0>()); break;
    case 1: rows(std::integral_constant<int, 1>()); break;
    case 2: rows(std::integral_constant<int, 2>()); break;
    case 3: rows(std::integral_constant<int, 3>()); break;
    case 4: rows(std::integral_constant<int, 4>()); break;
    case 5: rows(std::integral_constant<int, 5>()); break;
    case 6: rows(std::integral_constant<int, 6>()); break;
    case 7: rows(std::integral_constant<int, 7>()); break;
    default: rows(std::integral_constant<int, 15>()); break;
    }
    if (row < N) lj_finish(row, pi, v0, fx, fy, fz, p, force, vel, pos_out, mode);
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
    drift_wrap(x, p);
    vel[i] = v;
    pos[i] = x;
}

__global__ void box_flag_set(int *flag)
{
    if (threadIdx.x == 0) *flag = 1;
}

__global__ void box_check(const float4 *__restrict__ pos, int N, float L, int pbc, int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 x = pos[i];
    const float b = pbc == 2 ? x.w : L;
    if (!(x.x >= 0.0f && x.x <= b && x.y >= 0.0f && x.y <= b && x.z >= 0.0f && x.z <= b)) *flag = 0;
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

static LJParams make_params(float box, int pbc, float rc2, float dt, int *inbox = nullptr)
{
    LJParams p;
    p.L = box;
    p.hL = 0.5f * box;
    p.rc2 = rc2;
    p.dt = dt;
    p.hdt = 0.5f * dt;
    p.pbc = pbc;
    // least w with fl(w*w) >= rc2, in the device's (IEEE) float product
    volatile float w = std::sqrt(std::max(rc2, 0.0f));
    while (w * w < rc2) w = std::nextafter((float)w, INFINITY);
    p.W0 = w;
    p.inbox = inbox;
    return p;
}

// The pair functor's reciprocal and single-fma force factor hold for
// cutoffs with rc^2 below 2^31.
static int check_cutoff(float rc2, const char *who)
{
    PAT_REQUIRE(rc2 >= 0.0f && rc2 < 0x1p31f, PAT_ERR_UNSUPPORTED, "%s: cutoff^2 %g outside [0, 2^31)", who,
                (double)rc2);
    return PAT_OK;
}

int64_t csr_workspace_bytes(int64_t N, int64_t E)
{
    const int64_t nb = (N + 1 + kScanBlock - 1) / kScanBlock;
    const int64_t nb2 = (nb + kScanBlock - 1) / kScanBlock;
    // deg/cursor (N+1), block sums (nb), their scan (nb), second level (nb2 x 2), tmp cols and edge ids (2 x 2E)
    return (int64_t)sizeof(int) * ((N + 1) + 2 * nb + 2 * nb2 + 2 + 4 * E) + 256;
}

// Sum of n non-negative ints in 64 bits (into *out, which starts at 0):
// the neighbour build checks the total entry count with it before the int32
// row pointers are used, since an int32 scan of more than 2^31 entries wraps.
// *out += sum of n non-negative ints, in 64 bits (a 16-byte aligned; 256
// threads per workgroup).  Four 16-byte loads in flight per thread, and one
// atomic per workgroup: atomics from every wavefront to one word serialise.
__global__ __launch_bounds__(256) void sum_nonneg_i64(const int *__restrict__ a, int64_t n,
                                                      unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long part[4];
    auto add4 = [](const int4 v) {
        return (unsigned long long)(unsigned)v.x + (unsigned)v.y + (unsigned long long)(unsigned)v.z + (unsigned)v.w;
    };
    const int4 *__restrict__ a4 = (const int4 *)a;
    const int64_t n4 = n / 4, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long s = 0;
    int64_t i = gid;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const int4 v0 = a4[i], v1 = a4[i + stride], v2 = a4[i + 2 * stride], v3 = a4[i + 3 * stride];
        s += add4(v0) + add4(v1) + add4(v2) + add4(v3);
    }
    for (; i < n4; i += stride) s += add4(a4[i]);
    if (gid < n - 4 * n4) s += (unsigned)a[4 * n4 + gid];  // the last n % 4 ints
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long b = part[0] + part[1] + part[2] + part[3];
        if (b) atomicAdd(out, b);
    }
}

// Launches sum_nonneg_i64 over n ints at a (a 16-byte aligned device array).
static int sum_nonneg(const int *a, int64_t n, unsigned long long *out, hipStream_t st)
{
    PAT_REQUIRE(((uintptr_t)a & 15) == 0, PAT_ERR_INVALID, "sum_nonneg: array not 16-byte aligned");
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n / 4 + 255) / 256, 256));
    hipLaunchKernelGGL(sum_nonneg_i64, dim3((unsigned)blocks), dim3(256), 0, st, a, n, out);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
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

// The edge ids of the CSR by destination (row r: every edge with r as an
// endpoint, in increasing edge id -- the serial scatter's order for r);
// rowptr as csr_build writes it.  Same workspace as csr_build.
int csr_edge_ids(const int32_t *edges, int64_t E, int64_t N, int32_t *rowptr, int32_t *eids, void *work,
                 int64_t work_bytes, hipStream_t st)
{
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - 1 && E >= 0 && 2 * E < (1LL << 31), PAT_ERR_INVALID,
                "csr_edge_ids: sizes out of range (N=%lld, E=%lld)", (long long)N, (long long)E);
    PAT_REQUIRE(work_bytes >= csr_workspace_bytes(N, E), PAT_ERR_INVALID, "csr_edge_ids: workspace too small");
    const int64_t nb = (N + 1 + kScanBlock - 1) / kScanBlock;
    int *deg = (int *)work;
    int *ws = deg + (N + 1);
    int *tmp = ws + 2 * nb + 2 * ((nb + kScanBlock - 1) / kScanBlock) + 2;
    hipLaunchKernelGGL(fill_i32, dim3(grid_for(N + 1)), dim3(256), 0, st, deg, N + 1, 0);
    PAT_LAUNCH_CHECK();
    const unsigned grid = (unsigned)std::min<int64_t>((E + 255) / 256 + 1, 65536);
    if (E > 0) {
        hipLaunchKernelGGL(csr_degree, dim3(grid), dim3(256), 0, st, (const int2 *)edges, E, deg);
        PAT_LAUNCH_CHECK();
    }
    PAT_TRY(scan_exclusive(deg, rowptr, N + 1, ws, st));
    hipLaunchKernelGGL(copy_i32, dim3(grid_for(N + 1)), dim3(256), 0, st, deg, (const int *)rowptr, N + 1);
    PAT_LAUNCH_CHECK();
    if (E > 0) {
        int *tmpe = tmp + 2 * E;
        hipLaunchKernelGGL(csr_scatter, dim3(grid), dim3(256), 0, st, (const int2 *)edges, E, deg, tmp, tmpe);
        PAT_LAUNCH_CHECK();
        const unsigned g2 = (unsigned)((N * kWave + 255) / 256);
        hipLaunchKernelGGL(csr_row_sort_by_edge, dim3(g2), dim3(256), 0, st, rowptr, N, tmpe, tmpe, eids);
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
    hipLaunchKernelGGL(fill_i32, dim3(grid_for(N + 1)), dim3(256), 0, st, deg, N + 1, 0);
    PAT_LAUNCH_CHECK();
    const unsigned grid = (unsigned)std::min<int64_t>((E + 255) / 256 + 1, 65536);
    if (E > 0) {
        hipLaunchKernelGGL(csr_degree, dim3(grid), dim3(256), 0, st, (const int2 *)edges, E, deg);
        PAT_LAUNCH_CHECK();
    }
    PAT_TRY(scan_exclusive(deg, rowptr, N + 1, ws, st));
    // cursor = rowptr (reuse deg)
    hipLaunchKernelGGL(copy_i32, dim3(grid_for(N + 1)), dim3(256), 0, st, deg, (const int *)rowptr, N + 1);
    PAT_LAUNCH_CHECK();
    if (E > 0) {
        int *tmpe = tmp + 2 * E;
        hipLaunchKernelGGL(csr_scatter, dim3(grid), dim3(256), 0, st, (const int2 *)edges, E, deg, tmp, tmpe);
        PAT_LAUNCH_CHECK();
        const unsigned g2 = (unsigned)((N * kWave + 255) / 256);
        if (order == 0)
            hipLaunchKernelGGL(csr_row_sort_by_edge, dim3(g2), dim3(256), 0, st, rowptr, N, tmp, tmpe, cols);
        else
            hipLaunchKernelGGL(csr_row_sort, dim3(row_sort_grid(N)), dim3(256), 0, st, rowptr, N, tmp, cols, 0);
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
                 float rc2, float dt, int lpr, int mode, float4 *force, float4 *vel, float4 *pos_out, int *inbox,
                 hipStream_t st)
{
    PAT_TRY(check_cutoff(rc2, "lj_force_csr"));
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
                       make_params(box, pbc, rc2, dt, inbox), force, vel, pos_out, mode);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// Sliced layout of a CSR: widths on the device, slice offsets scanned on the
// host in 64-bit (the padded total must stay below 2^31), then the fill.
// width_h / sptr_h: host scratch of nsl and nsl+1 ints; *scol grows to the
// padded total (capacity *cap).  Refuses (PAT_ERR_UNSUPPORTED, nothing
// changed) when padding would exceed max_pad times the entry count plus one
// slice of 64 entries per slice.
int sell_build(const int32_t *rowptr, const int32_t *cols, const int32_t *perm, int64_t N, double max_pad,
               int32_t *width_d, int32_t *sptr, int32_t **scol, int64_t *cap, int64_t *total, int pad, hipStream_t st)
{
    const int64_t nsl = (N + kWave - 1) / kWave;
    hipLaunchKernelGGL(sell_width, dim3((unsigned)((nsl + 255) / 256)), dim3(256), 0, st, (const int *)rowptr,
                       (const int *)perm, (int)N, (int)nsl, (int *)width_d);
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
                       (const int *)cols, (const int *)perm, (int)N, (int)nsl, (const int *)sptr, (int *)*scol, pad);
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
    hipLaunchKernelGGL(sell_width, dim3((unsigned)((nsl + 255) / 256)), dim3(256), 0, st, (const int *)rowptr,
                       (const int *)nullptr, (int)N, (int)nsl, (int *)sptr);
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
                       (const int *)cols, (const int *)nullptr, (int)N, (int)nsl, (const int *)sptr, (int *)scol, -1);
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipStreamSynchronize(st));  // sp leaves scope
    return PAT_OK;
}

// Slice order by row class from positions pos (see perm_block_sort): perm
// gets nsl*64 entries.
int sell_perm_build(const float4 *pos, int64_t N, float box, int pbc, float rc2, int32_t *perm, hipStream_t st)
{
    const int64_t nsl = (N + kWave - 1) / kWave, total = nsl * kWave;
    const LJParams p = make_params(box, pbc, rc2, 0.f);
    hipLaunchKernelGGL(perm_block_sort, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, pos, (int)N,
                       (int)total, box, p.W0, pbc, (int *)perm);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// Buffers of the spatial slice order, grown as needed.
struct SortWork {
    int64_t n = 0;
    size_t tmp_bytes = 0;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    int *vals = nullptr, *vals2 = nullptr;
    void *tmp = nullptr;
};

SortWork *sort_create() { return new SortWork(); }
void sort_destroy(SortWork *w)
{
    if (!w) return;
    for (void *p : {(void *)w->keys, (void *)w->keys2, (void *)w->vals, (void *)w->vals2, w->tmp})
        if (p) (void)pat::dev_free(p);
    delete w;
}

// Spatial slice order of N nodes (see morton_keys): perm gets nsl*64
// entries.  sysoff: nsys + 1 node offsets on the device (null: one system).
int sell_perm_spatial(const float4 *pos, int64_t N, float box, int pbc, float h, const int32_t *sysoff, int nsys,
                      int32_t *perm, SortWork *w, hipStream_t st)
{
    PAT_REQUIRE(pbc != 0 && h > 0 && nsys >= 1 && nsys < (1 << 20), PAT_ERR_INVALID,
                "spatial slice order: needs a periodic box, h > 0 and 1 <= systems < 2^20");
    const int64_t nsl = (N + kWave - 1) / kWave, total = nsl * kWave;
    int end_bit = 30;
    while ((1LL << (end_bit - 30)) < nsys) end_bit++;
    size_t need = 0;
    PAT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                               (const int *)nullptr, (int *)nullptr, (int)N, 0, end_bit, st));
    if (w->n < N || w->tmp_bytes < need) {
        for (void **p : {(void **)&w->keys, (void **)&w->keys2, (void **)&w->vals, (void **)&w->vals2, &w->tmp})
            if (*p) {
                (void)pat::dev_free(*p);
                *p = nullptr;
            }
        const int64_t n = std::max(w->n, N);
        const size_t tb = std::max(w->tmp_bytes, need);
        if (pat::dev_malloc((void **)&w->keys, sizeof(uint64_t) * (size_t)n) != hipSuccess ||
            pat::dev_malloc((void **)&w->keys2, sizeof(uint64_t) * (size_t)n) != hipSuccess ||
            pat::dev_malloc((void **)&w->vals, sizeof(int) * (size_t)n) != hipSuccess ||
            pat::dev_malloc((void **)&w->vals2, sizeof(int) * (size_t)n) != hipSuccess ||
            pat::dev_malloc(&w->tmp, tb > 0 ? tb : 1) != hipSuccess) {
            (void)hipGetLastError();
            w->n = 0;
            w->tmp_bytes = 0;
            pat::set_error("spatial slice order: hipMalloc for %lld nodes failed", (long long)N);
            return PAT_ERR_NOMEM;
        }
        w->n = n;
        w->tmp_bytes = tb;
    }
    const unsigned g = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(morton_keys, dim3(g), dim3(256), 0, st, pos, (int)N, box, pbc, h, (const int *)sysoff, nsys,
                       w->keys, w->vals);
    PAT_LAUNCH_CHECK();
    size_t tb = w->tmp_bytes;
    PAT_HIP(hipcub::DeviceRadixSort::SortPairs(w->tmp, tb, (const uint64_t *)w->keys, w->keys2,
                                               (const int *)w->vals, w->vals2, (int)N, 0, end_bit, st));
    hipLaunchKernelGGL(perm_pad, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, (const int *)w->vals2,
                       (int)N, (int)total, (int *)perm);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// Halo layout of the sliced layout (sptr, perm as built for the sentinel
// kinds): hptr gets nblk+1 offsets (nblk = ceil(nsl/4)); hidx and scol16
// are (re)allocated as needed (capacities in *hcap, *scap16).  Returns
// PAT_ERR_UNSUPPORTED when some block does not fit the halo form; *max_nh
// = the largest list.
int halo_build(const int32_t *rowptr, const int32_t *cols, const int32_t *perm, const int32_t *sptr, int64_t N,
               int64_t stotal, int32_t *hptr, int32_t **hidx, int64_t *hcap, uint16_t **scol16, int64_t *scap16,
               int *max_nh, hipStream_t st)
{
    const int64_t nsl = (N + kWave - 1) / kWave, nblk = (nsl + 3) / 4;
    hipLaunchKernelGGL(halo_count, dim3((unsigned)nblk), dim3(256), 0, st, (const int *)rowptr, (const int *)cols,
                       (const int *)perm, (int)N, (int)nsl, (int *)hptr);
    PAT_LAUNCH_CHECK();
    std::vector<int32_t> h((size_t)nblk + 1);
    PAT_HIP(hipMemcpyAsync(h.data(), hptr, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    int64_t acc = 0;
    int mx = 0;
    for (int64_t b = 0; b < nblk; b++) {
        const int c = h[b];
        if (c < 0) {
            pat::set_error("halo_build: block %lld does not fit the halo form", (long long)b);
            return PAT_ERR_UNSUPPORTED;
        }
        h[b] = (int32_t)acc;
        acc += c;
        mx = std::max(mx, c);
    }
    h[nblk] = (int32_t)acc;
    PAT_REQUIRE(acc < (1LL << 31), PAT_ERR_UNSUPPORTED, "halo_build: %lld list entries", (long long)acc);
    auto grow = [](void **ptr, int64_t *cap, int64_t need, size_t elem) -> int {
        if (need <= *cap) return PAT_OK;
        const int64_t want = need + need / 4 + 4096;
        void *n = nullptr;
        if (pat::dev_malloc(&n, elem * (size_t)want) != hipSuccess) {
            (void)hipGetLastError();
            pat::set_error("halo_build: hipMalloc of %lld entries failed", (long long)want);
            return PAT_ERR_NOMEM;
        }
        if (*ptr) (void)pat::dev_free(*ptr);
        *ptr = n;
        *cap = want;
        return PAT_OK;
    };
    PAT_TRY(grow((void **)hidx, hcap, acc, sizeof(int32_t)));
    PAT_TRY(grow((void **)scol16, scap16, stotal, sizeof(uint16_t)));
    PAT_HIP(hipMemcpyAsync(hptr, h.data(), sizeof(int32_t) * (nblk + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(halo_fill, dim3((unsigned)nblk), dim3(256), 0, st, (const int *)rowptr, (const int *)cols,
                       (const int *)perm, (int)N, (int)nsl, (const int *)sptr, (const int *)hptr, (int *)*hidx,
                       *scol16);
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipStreamSynchronize(st));  // h leaves scope
    *max_nh = mx;
    return PAT_OK;
}

// Loads this file's code object now (kernels are otherwise loaded at their
// first launch, which a serving job would wait for: ~55 ms for this file).
int irregular_preload()
{
    hipFuncAttributes a;
    PAT_HIP(hipFuncGetAttributes(&a, (const void *)lj_sell<4, true, false>));
    return PAT_OK;
}

// Forces over the halo layout; U = 4 or 8 entries' loads ahead.
int lj_force_halo(const float4 *pos, const int32_t *sptr, const uint16_t *scol16, const int32_t *hptr,
                  const int32_t *hidx, const int32_t *perm, int64_t N, int max_nh, float box, int pbc, float rc2,
                  float dt, int mode, float4 *force, float4 *vel, float4 *pos_out, int U, int *inbox, hipStream_t st)
{
    PAT_REQUIRE(N > 0 && N < (1LL << 28) - 1, PAT_ERR_INVALID, "lj_force_halo: bad node count");
    PAT_REQUIRE(max_nh >= 0 && max_nh < kHaloCap, PAT_ERR_INVALID, "lj_force_halo: list of %d entries", max_nh);
    PAT_TRY(check_cutoff(rc2, "lj_force_halo"));
    const int64_t nsl = (N + kWave - 1) / kWave, nblk = (nsl + 3) / 4;
    const size_t lds = sizeof(float4) * (size_t)(max_nh + 1);
    auto k = U == 8 ? lj_sell_halo<8> : lj_sell_halo<4>;
    hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(256), lds, st, pos, (const int *)sptr, scol16,
                       (const int *)hptr, (const int *)hidx, (const int *)perm, (int)N,
                       make_params(box, pbc, rc2, dt, inbox), force, vel, pos_out, mode);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// kind: 0 = lj_sell<4>, 1 = pipelined <4>, 2 = pipelined <8> over an
// index layout whose padding repeats the row; 3, 4, 5 = the same over a
// byte-offset layout padded with the sentinel node N (pos then holds N+1
// entries, the last beyond every cutoff, and N < 2^28), which may skip the
// minimum image per wavefront while *inbox is set.  Same sums for all.
int lj_force_sell(const float4 *pos, const int32_t *rowptr, const int32_t *sptr, const int32_t *scol,
                  const int32_t *perm, int64_t N, float box, int pbc, float rc2, float dt, int mode, float4 *force,
                  float4 *vel, float4 *pos_out, int kind, int *inbox, hipStream_t st)
{
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - kWave, PAT_ERR_INVALID, "lj_force_sell: bad node count");
    PAT_TRY(check_cutoff(rc2, "lj_force_sell"));
    using K = void (*)(const float4 *, const int *, const int *, const int *, const int *, int, LJParams, float4 *,
                       float4 *, float4 *, int);
    const K ks[] = {lj_sell<4, false, false>, lj_sell<4, false, true>, lj_sell<8, false, true>,
                    lj_sell<4, true, false>,  lj_sell<4, true, true>,  lj_sell<8, true, true>};
    PAT_REQUIRE(kind >= 0 && kind < 6, PAT_ERR_INVALID, "lj_force_sell: kind %d", kind);
    PAT_REQUIRE(kind < 3 || N < (1LL << 28) - 1, PAT_ERR_INVALID, "lj_force_sell: sentinel kinds need N < 2^28 - 1");
    const int64_t nsl = (N + kWave - 1) / kWave;
    hipLaunchKernelGGL(ks[kind], dim3((unsigned)((nsl + 3) / 4)), dim3(256), 0, st, pos, (const int *)rowptr,
                       (const int *)sptr, (const int *)scol, (const int *)perm, (int)N,
                       make_params(box, pbc, rc2, dt, inbox), force, vel, pos_out, mode);
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
    PAT_TRY(check_cutoff(rc2, "lj_force_atomic"));
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
                      int *inbox, hipStream_t st)
{
    hipLaunchKernelGGL(vv_kick_drift, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, pos, vel, force, (int)N,
                       make_params(box, pbc, 0.f, dt, inbox));
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

// *flag = 1 when every position of nodes [0, N) lies in [0, L] on all
// three axes (L = box, or each node's w for pbc 2), else 0.
int box_check_dev(const float4 *pos, int64_t N, float box, int pbc, int *flag, hipStream_t st)
{
    hipLaunchKernelGGL(box_flag_set, dim3(1), dim3(64), 0, st, flag);
    PAT_LAUNCH_CHECK();
    if (pbc && N > 0) {
        hipLaunchKernelGGL(box_check, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, pos, (int)N, box, pbc,
                           flag);
        PAT_LAUNCH_CHECK();
    }
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
    int64_t N = 0, ncell = 0, cap = 0;  // N, ncell: capacities of the per-atom / per-bin buffers; cap: coltmp
    int64_t nsys_cap = 0;
    int *cell = nullptr, *cstart = nullptr, *cursor = nullptr, *atmp = nullptr, *catoms = nullptr, *deg = nullptr,
        *ws = nullptr, *coltmp = nullptr;
    NbSys *sys = nullptr;
    int *slot = nullptr, *flag = nullptr;  // one-pass rows: kNbSlot entries per atom; overflow flag
    unsigned long long *tot64 = nullptr;   // 64-bit sum of the row lengths
    int64_t small_n = 0;                   // deg / ws capacity in atoms of the two-bin (all pairs) build
    int64_t slot_cap = 0;
    bool slot_fail = false;  // a row outgrew its slot: two passes from then on
    std::vector<NbSys> sys_h;  // source of the last upload (read before the build returns)
};

static void nb_free(NbWork &w)
{
    for (int *p : {w.cell, w.cstart, w.cursor, w.atmp, w.catoms, w.deg, w.ws, w.coltmp, w.slot, w.flag})
        if (p) (void)pat::dev_free(p);
    if (w.sys) (void)pat::dev_free(w.sys);
    if (w.tot64) (void)pat::dev_free(w.tot64);
    w = NbWork{};
}

// The neighbour build's entry count check: the 64-bit sum of the row
// lengths must fit an int32 row pointer and equal the int32 scan's total.
extern "C" int pat_internal_nb_check_total(unsigned long long tot64, int tot32);
int nb_check_total(unsigned long long tot64, int tot32) { return pat_internal_nb_check_total(tot64, tot32); }
}  // namespace pat

// Exported (outside the public header) so the CPU tests can check it with
// synthetic totals, including sums that wrap an int32 scan.
extern "C" int pat_internal_nb_check_total(unsigned long long tot64, int tot32)
{
    PAT_REQUIRE(tot64 <= (unsigned long long)INT32_MAX, PAT_ERR_UNSUPPORTED,
                "neighbour build: %llu list entries exceed the int32 row pointers (at most 2^31-1)", tot64);
    PAT_REQUIRE((long long)tot32 == (long long)tot64, PAT_ERR_INVALID,
                "neighbour build: scanned total %d differs from the 64-bit sum %llu", tot32, tot64);
    return PAT_OK;
}

namespace pat {

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

// The neighbour CSR of one periodic cube whose side holds two bins of the
// cutoff (2 cut <= L < 3 cut, at most kNbBruteMax atoms): every pair tested
// (nb_brute), rows sorted by construction.  Same outputs and growth rules as
// neighbor_build_multi.
static int neighbor_build_small(const float4 *pos, int64_t N, double L, double cut, NbWork *w, int32_t *rowptr,
                                int32_t **cols, int64_t *cap, int64_t *total, hipStream_t st)
{
    if (!w->deg || !w->ws || std::max(w->N, w->small_n) < N) {
        if (w->deg) (void)pat::dev_free(w->deg);
        if (w->ws) (void)pat::dev_free(w->ws);
        w->deg = w->ws = nullptr;
        w->N = 0;  // a binned build on this workspace reallocates all of its buffers
        const int64_t scan_ws = 2 * ((N + 1 + kScanBlock - 1) / kScanBlock) + 2 * kScanBlock + 4;
        PAT_TRY(nb_alloc(&w->deg, N + 1));
        PAT_TRY(nb_alloc(&w->ws, scan_ws));
        w->small_n = N;
    }
    if (!w->tot64 && pat::dev_malloc((void **)&w->tot64, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        w->tot64 = nullptr;
        pat::set_error("neighbour build: hipMalloc of 8 bytes failed");
        return PAT_ERR_NOMEM;
    }
    BinGeom g;
    g.L = (float)L;
    g.hL = 0.5f * g.L;
    g.cut2 = (float)(cut * cut);
    g.inv = 0.f;
    g.nb = 2;
    const unsigned grid = (unsigned)((N + 3) / 4);
    hipLaunchKernelGGL(nb_brute<kNbCount>, dim3(grid), dim3(256), 0, st, pos, (int)N, g, w->deg, (const int *)nullptr,
                       (int *)nullptr);
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipMemsetAsync(w->deg + N, 0, sizeof(int), st));
    PAT_HIP(hipMemsetAsync(w->tot64, 0, sizeof(unsigned long long), st));
    PAT_TRY(sum_nonneg((const int *)w->deg, N, w->tot64, st));
    PAT_TRY(scan_exclusive(w->deg, rowptr, N + 1, w->ws, st));
    int tot = 0;
    unsigned long long tot64 = 0;
    PAT_HIP(hipMemcpyAsync(&tot, rowptr + N, sizeof(int), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipMemcpyAsync(&tot64, w->tot64, sizeof(tot64), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    PAT_TRY(nb_check_total(tot64, tot));
    if (tot > *cap) {
        const int64_t want = (int64_t)(tot * 1.25) + 1024;
        if (*cols) (void)pat::dev_free(*cols);
        *cols = nullptr;
        PAT_TRY(nb_alloc(cols, want));
        *cap = want;
    }
    hipLaunchKernelGGL(nb_brute<kNbFill>, dim3(grid), dim3(256), 0, st, pos, (int)N, g, (int *)nullptr,
                       (const int *)rowptr, *cols);
    PAT_LAUNCH_CHECK();
    *total = tot;
    return PAT_OK;
}

// Rebuilds the full, row-sorted neighbour CSR of nsys independent periodic
// cubes at once: system s holds atoms [off[s], off[s+1]) of the float4
// positions in a cube of side Ls[s]; pairs closer than cut (within one
// system) land in both rows, with global atom indices.  *cols (capacity
// *cap entries) is reallocated when too small.  Synchronises the stream
// once (to size the column array).
int neighbor_build_multi(const float4 *pos, int nsys, const double *Ls, const int64_t *off, double cut, NbWork *w,
                         int32_t *rowptr, int32_t **cols, int64_t *cap, int64_t *total, hipStream_t st)
{
    PAT_REQUIRE(nsys >= 1 && Ls && off && off[0] == 0 && cut > 0, PAT_ERR_INVALID, "neighbour build: bad arguments");
    const int64_t N = off[nsys];
    PAT_REQUIRE(N > 0 && N < (1LL << 31) - 1, PAT_ERR_INVALID, "neighbour build: bad atom count");
    w->sys_h.resize((size_t)nsys);
    int64_t ncell = 0;
    for (int s = 0; s < nsys; s++) {
        const double L = Ls[s];
        PAT_REQUIRE(L > 0 && off[s + 1] > off[s], PAT_ERR_INVALID, "neighbour build: system %d is empty", s);
        const int nb = (int)std::floor(L / (cut * 1.0001));
        if (nb == 2 && nsys == 1 && off[1] <= kNbBruteMax) {
            const int64_t n = off[1];
            return neighbor_build_small(pos, n, L, cut, w, rowptr, cols, cap, total, st);
        }
        PAT_REQUIRE(nb >= 3, PAT_ERR_UNSUPPORTED,
                    "neighbour build: box %.3f holds fewer than 3 bins of %.3f (a single system of at most %d "
                    "atoms may hold 2)", L, cut, kNbBruteMax);
        PAT_REQUIRE(nb < 1024, PAT_ERR_INVALID, "neighbour build: too many bins");
        NbSys &S = w->sys_h[(size_t)s];
        S.L = (float)L;
        S.hL = 0.5f * S.L;
        S.inv = (float)nb / S.L;
        S.nb = nb;
        S.cell0 = (int)ncell;
        S.atom0 = (int)off[s];
        S.atom1 = (int)off[s + 1];
        S.pad = 0;
        ncell += (int64_t)nb * nb * nb;
        PAT_REQUIRE(ncell < (1LL << 30), PAT_ERR_INVALID, "neighbour build: too many bins");
    }
    // the buffers only grow: rebuilds of other sizes reuse them (a free
    // here would wait for the whole device)
    if (w->N < N || w->ncell < ncell || w->nsys_cap < nsys) {
        const int64_t cap_keep = w->cap;
        int *coltmp = w->coltmp;
        w->coltmp = nullptr;
        const int64_t Nn = std::max(w->N, N), nc = std::max(w->ncell, ncell), ns = std::max<int64_t>(w->nsys_cap, nsys);
        std::vector<NbSys> keep = std::move(w->sys_h);
        nb_free(*w);
        w->sys_h = std::move(keep);
        w->coltmp = coltmp;
        w->cap = cap_keep;
        const int64_t m = std::max<int64_t>(Nn, nc) + 1;
        const int64_t scan_ws = 2 * ((m + kScanBlock - 1) / kScanBlock) + 2 * kScanBlock + 4;
        PAT_TRY(nb_alloc(&w->cell, Nn));
        PAT_TRY(nb_alloc(&w->cstart, nc + 1));
        PAT_TRY(nb_alloc(&w->cursor, nc + 1));
        PAT_TRY(nb_alloc(&w->atmp, Nn));
        PAT_TRY(nb_alloc(&w->catoms, Nn));
        PAT_TRY(nb_alloc(&w->deg, Nn + 1));
        PAT_TRY(nb_alloc(&w->ws, scan_ws));
        if (pat::dev_malloc((void **)&w->sys, sizeof(NbSys) * (size_t)ns) != hipSuccess) {
            (void)hipGetLastError();
            pat::set_error("neighbour build: hipMalloc of %lld systems failed", (long long)ns);
            return PAT_ERR_NOMEM;
        }
        w->N = Nn;
        w->ncell = nc;
        w->nsys_cap = ns;
    }
    const float cut2 = (float)(cut * cut);
    PAT_HIP(hipMemcpyAsync(w->sys, w->sys_h.data(), sizeof(NbSys) * (size_t)nsys, hipMemcpyHostToDevice, st));
    const unsigned gN = (unsigned)((N + 255) / 256);
    const unsigned gC = (unsigned)((ncell + 3) / 4);
    // bin the atoms and sort each bin's list by atom index
    PAT_HIP(hipMemsetAsync(w->cursor, 0, sizeof(int) * (ncell + 1), st));
    hipLaunchKernelGGL(nb_bin_atoms, dim3(gN), dim3(256), 0, st, pos, (int)N, (const NbSys *)w->sys, nsys, w->cell,
                       w->cursor);
    PAT_LAUNCH_CHECK();
    PAT_TRY(scan_exclusive(w->cursor, w->cstart, ncell + 1, w->ws, st));
    PAT_HIP(hipMemcpyAsync(w->cursor, w->cstart, sizeof(int) * (ncell + 1), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(nb_scatter_atoms, dim3(gN), dim3(256), 0, st, w->cell, (int)N, w->cursor, w->atmp);
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(csr_row_sort, dim3(row_sort_grid(ncell)), dim3(256), 0, st, w->cstart, ncell,
                       w->atmp, w->catoms, 0);
    PAT_LAUNCH_CHECK();
    // One pass: row lengths, and each row into a fixed slot of kNbSlot
    // entries per atom; the rows are then sorted straight out of the slots.
    // A row longer than its slot sets a flag, and that build (and every
    // later one on this workspace) fills the rows again at their final
    // offsets instead.  Without slots (or for more than 2^31 slot entries):
    // a counting pass first.
    const bool slots = !w->slot_fail && N * kNbSlot < (1LL << 31);
    if (slots) {
        if (w->slot_cap < N * kNbSlot) {
            if (w->slot) (void)pat::dev_free(w->slot);
            w->slot = nullptr;
            w->slot_cap = 0;
            PAT_TRY(nb_alloc(&w->slot, N * kNbSlot));
            w->slot_cap = N * kNbSlot;
        }
        if (!w->flag) PAT_TRY(nb_alloc(&w->flag, 1));
        PAT_HIP(hipMemsetAsync(w->flag, 0, sizeof(int), st));
        hipLaunchKernelGGL(nb_cells<kNbSlotFill>, dim3(gC), dim3(256), 0, st, pos, (const NbSys *)w->sys, nsys, (int)ncell,
                           cut2, (const int *)w->cstart, (const int *)w->catoms, w->deg, (const int *)nullptr, w->slot,
                           w->flag);
    } else {
        hipLaunchKernelGGL(nb_cells<kNbCount>, dim3(gC), dim3(256), 0, st, pos, (const NbSys *)w->sys, nsys,
                           (int)ncell, cut2, (const int *)w->cstart, (const int *)w->catoms, w->deg,
                           (const int *)nullptr, (int *)nullptr, (int *)nullptr);
    }
    PAT_LAUNCH_CHECK();
    PAT_HIP(hipMemsetAsync(w->deg + N, 0, sizeof(int), st));
    if (!w->tot64 && pat::dev_malloc((void **)&w->tot64, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        w->tot64 = nullptr;
        pat::set_error("neighbour build: hipMalloc of 8 bytes failed");
        return PAT_ERR_NOMEM;
    }
    PAT_HIP(hipMemsetAsync(w->tot64, 0, sizeof(unsigned long long), st));
    PAT_TRY(sum_nonneg((const int *)w->deg, N, w->tot64, st));
    PAT_TRY(scan_exclusive(w->deg, rowptr, N + 1, w->ws, st));
    int tot = 0, over = 0;
    unsigned long long tot64 = 0;
    PAT_HIP(hipMemcpyAsync(&tot, rowptr + N, sizeof(int), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipMemcpyAsync(&tot64, w->tot64, sizeof(tot64), hipMemcpyDeviceToHost, st));
    if (slots) PAT_HIP(hipMemcpyAsync(&over, w->flag, sizeof(int), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    // Nothing has been written at the scanned offsets yet: refuse before any
    // fill when the entries do not fit int32 row pointers.
    PAT_TRY(nb_check_total(tot64, tot));
    if (over) w->slot_fail = true;
    const bool from_slots = slots && !over;
    if (tot > *cap) {
        const int64_t want = (int64_t)(tot * 1.25) + 1024;
        if (*cols) (void)pat::dev_free(*cols);
        *cols = nullptr;
        PAT_TRY(nb_alloc(cols, want));
        *cap = want;
    }
    if (!from_slots && tot > w->cap) {
        const int64_t want = (int64_t)(tot * 1.25) + 1024;
        if (w->coltmp) (void)pat::dev_free(w->coltmp);
        w->coltmp = nullptr;
        w->cap = 0;
        PAT_TRY(nb_alloc(&w->coltmp, want));
        w->cap = want;
    }
    if (!from_slots) {
        hipLaunchKernelGGL(nb_cells<kNbFill>, dim3(gC), dim3(256), 0, st, pos, (const NbSys *)w->sys, nsys,
                           (int)ncell, cut2, (const int *)w->cstart, (const int *)w->catoms, w->deg,
                           (const int *)rowptr, w->coltmp, (int *)nullptr);
        PAT_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(csr_row_sort, dim3(row_sort_grid(N)), dim3(256), 0, st, (const int *)rowptr,
                       N, from_slots ? (const int *)w->slot : (const int *)w->coltmp, *cols, from_slots ? kNbSlot : 0);
    PAT_LAUNCH_CHECK();
    *total = tot;
    return PAT_OK;
}

// The same for one cube of N atoms and side L.
int neighbor_build(const float4 *pos, int64_t N, double L, double cut, NbWork *w, int32_t *rowptr, int32_t **cols,
                   int64_t *cap, int64_t *total, hipStream_t st)
{
    const int64_t off[2] = {0, N};
    return neighbor_build_multi(pos, 1, &L, off, cut, w, rowptr, cols, cap, total, st);
}

}  // namespace pat

// ------------------------------------------------------------------
// CSR build without allocation: the work space of pat_ir_csr_build_dev comes
// from a static device array, for callers that cannot allocate (code under
// graph capture, serving loops, harnesses that forbid hipMalloc).  Every GPU
// the library runs on holds its own copy until the process ends; calls that
// overlap on one GPU share it and race.

namespace {

constexpr int64_t kCsrScratchBytes = 128LL << 20;  // the csr_build work space of ~8.3M edges
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kCsrScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices] = {};

// The current device's address of 'scratch', looked up once per device.
int scratch_address(void **out)
{
    int dev = 0;
    PAT_HIP(hipGetDevice(&dev));
    PAT_REQUIRE(dev >= 0 && dev < kScratchDevices, PAT_ERR_UNSUPPORTED, "scratch: device %d beyond %d", dev,
                kScratchDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        PAT_REQUIRE(e == hipSuccess && p, PAT_ERR_HIP, "scratch: hipGetSymbolAddress failed on device %d: %s", dev,
                    hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PAT_OK;
}

}  // namespace

extern "C" int pat_ir_csr_build_dev_noalloc(const int32_t *edges, int64_t num_edges, int64_t num_nodes, int order,
                                            int32_t *rowptr, int32_t *cols, void *stream)
{
    PAT_REQUIRE(rowptr && (num_edges == 0 || (edges && cols)), PAT_ERR_INVALID,
                "pat_ir_csr_build_dev_noalloc: null argument");
    PAT_REQUIRE(order == PAT_IR_ORDER_EDGE || order == PAT_IR_ORDER_NODE, PAT_ERR_INVALID,
                "pat_ir_csr_build_dev_noalloc: row order %d not in {PAT_IR_ORDER_EDGE, PAT_IR_ORDER_NODE}", order);
    PAT_REQUIRE(num_nodes > 0 && num_nodes < (1LL << 31) - 1 && num_edges >= 0 && 2 * num_edges < (1LL << 31),
                PAT_ERR_INVALID, "pat_ir_csr_build_dev_noalloc: sizes out of range (N=%lld, E=%lld)",
                (long long)num_nodes, (long long)num_edges);
    const int64_t need = pat::csr_workspace_bytes(num_nodes, num_edges);
    PAT_REQUIRE(need <= kCsrScratchBytes, PAT_ERR_UNSUPPORTED,
                "pat_ir_csr_build_dev_noalloc: %lld nodes and %lld edges need %lld bytes of work space; the static "
                "scratch holds %lld (use pat_ir_csr_build_dev with a work space of your own)",
                (long long)num_nodes, (long long)num_edges, (long long)need, (long long)kCsrScratchBytes);
    void *ws = nullptr;
    PAT_TRY(scratch_address(&ws));
    return pat::csr_build(edges, num_edges, num_nodes, order, rowptr, cols, ws, kCsrScratchBytes, (hipStream_t)stream);
}
