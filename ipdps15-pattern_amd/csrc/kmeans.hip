// Generalized-reduction engine for gfx950: kmeans assignment + centroid
// accumulation (the reference's map1/reduce1 functors, kmeans/regular.h:
// 75-136, run by the regular compute_gpu kernel, kmeans/regular/
// gpu_kernel.cu:39-183, into an LDS hash reduction object, ros.cu:15-170).
//
// Reduction object.  map1 emits {x_0..x_{D-1}, 1, min_dist} under key =
// nearest centre and reduce1 adds values per key.  The reference adds FP32
// values in lock-acquisition order, which is not reproducible.  Here each
// coordinate (and min_dist) is rounded once to an integer multiple of 2^-s
// (s chosen so every rounded value fits in 31 bits and the largest possible
// sum in 62) and summed as int64:
// integer addition is associative, so every summation order -- LDS atomics
// within a workgroup, global atomics across workgroups, or the CPU oracle's
// loop -- gives identical bits, and so identical centroids every iteration.
// The dense per-workgroup object (K x D sums, K counts, K distance sums)
// lives in LDS and is flushed once per workgroup with global atomics.
//
// Assignment, plain path: the functor's arithmetic exactly -- for each centre
// in order, dist = sum_d (c_d - x_d)^2 in FP32 in dimension order, sqrt,
// strict < against min_dist starting at 65536*65 (lowest index wins ties).
// Compiled with -ffp-contract=off so no multiply-add is fused.
//
// Assignment, MFMA path: a wavefront takes 32 points and computes
// ||c||^2 - 2 x.c for every centre with v_mfma_f32_32x32x16_bf16 (FP32
// accumulation, ||c||^2 split into three BF16 terms on an extra K step).
// With NT = 3 terms, x and -2c are each split into BF16 high and low parts
// and the product is xh.ch + xh.cl + xl.ch (the dropped terms are O(2^-16)
// relative); with NT = 1 both are rounded to BF16 once (O(2^-8)).  A
// per-point bound E on the error of that value (operand rounding, the norm
// split, FP32 accumulation) plus the FP32 rounding of the reference's own
// difference form gives a margin:
// every centre whose approximate value is within the margin of the smallest
// is rescored with the plain path's exact FP32 arithmetic, and the exact
// results decide (same tie rule).  A centre outside the margin provably has a
// larger reference distance than the best one, so the assignment equals the
// plain path's bit for bit.
#include <algorithm>
#include <cmath>

#include "pat_common.hpp"
#include "pat_device.hpp"
#include "kmeans_dev.hpp"

namespace {

using pat::kWave;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2;
typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4;

// (lo, hi) rounded to BF16 (nearest even, as a scalar conversion rounds)
// in one packed conversion: lo in bits 0-15, hi in bits 16-31.
__device__ __forceinline__ unsigned bf16_pair(float lo, float hi)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr float kInitMin = 65536 * 65;  // map1's initial min_dist (kmeans/regular.h:86)

// Position of element e of centre row k in the BF16 MFMA operand array,
// stored in fragment order: for 32-row block b and K-step s, lane
// l = (k % 32) + 32 * ((e % 16) / 8) holds elements 8*(l/32) .. +8 of row
// b*32 + l%32 at 16-byte slot (b*NS + s)*64 + l, so a wavefront's operand
// read is 1 KB of consecutive LDS bytes (no bank conflicts).
__host__ __device__ __forceinline__ int64_t frag_index(int k, int e, int ns)
{
    const int b = k / 32, col = k % 32, s = e / 16, h = (e % 16) / 8, j = e % 8;
    return ((((int64_t)b * ns + s) * 2 + h) * 32 + col) * 8 + j;
}

// rint(v * 2^s) as a 64-bit two's-complement integer, with scale = 2^s an
// FP32 power of two chosen so that |v| 2^s <= 2^30 for every value summed:
// v * scale is exact, rounding it to an integer is exact in FP32, and the
// integer converts exactly to int32.  The same integer as
// rint((double)v * 2^s) in FP64.
__device__ __forceinline__ unsigned long long to_fixed(float v, float scale)
{
    return (unsigned long long)(long long)(int)__builtin_rintf(v * scale);
}

// ------------------------------------------------------------ shared parts

struct KmArgs {
    const float *x;         // N x D points
    const float *c;         // K x D centres (FP32)
    int64_t N;
    int D, K;
    float scale, dscale;    // fixed-point scales 2^s, 2^ds (FP32 powers of two)
    int32_t *assign;        // N
    float *mindist;         // N (may be null)
    unsigned long long *gsum;   // K x D
    unsigned long long *gcnt;   // K
    unsigned long long *gdsum;  // K
    int use_lds_acc;        // 1: LDS reduction object, 0: global atomics directly
    int defer_acc;          // 1: assignments only; the bucketed reduction (km_bucket_*) sums afterwards
    int rs;                 // row stride (8-byte words) of the coordinate sums: D in global memory,
                            // odd in LDS so that rows of different keys start in different banks
};

// Adds the wavefront's points [base, base+np) with keys key_of(j) into the
// reduction object.  Lanes cover (point, dimension) pairs so consecutive
// lanes touch consecutive 8-byte words.
template <typename KeyFn, typename DistFn>
__device__ __forceinline__ void accumulate_points(const KmArgs &a, int64_t base, int np, KeyFn key_of, DistFn dist_of,
                                                  unsigned long long *acc, unsigned long long *cnt,
                                                  unsigned long long *dsum)
{
    // Trip counts are uniform over the wavefront so that every lane is active
    // at each cross-lane read of the keys.
    const int lane = threadIdx.x % kWave;
    const int D = a.D;
    const int total = np * D;
    // lane covers flat index t = j*D + d, advanced by kWave per trip without
    // a division inside the loop
    const int step_j = kWave / D, step_d = kWave % D;
    int j = lane / D, d = lane % D;
    for (int t0 = 0; t0 < total; t0 += kWave) {
        const bool in = t0 + lane < total;
        const int k = key_of(in ? j : np - 1);
        if (in) {
            const float v = a.x[(base + j) * D + d];
            atomicAdd(&acc[(int64_t)k * a.rs + d], to_fixed(v, a.scale));
        }
        j += step_j;
        d += step_d;
        if (d >= D) {
            d -= D;
            j++;
        }
    }
    const int jj = lane < np ? lane : np - 1;
    const int k = key_of(jj);
    const float dist = dist_of(jj);
    if (lane < np) {
        atomicAdd(&cnt[k], 1ULL);
        atomicAdd(&dsum[k], to_fixed(dist, a.dscale));
    }
}

__device__ void flush_lds_object(const KmArgs &a, unsigned long long *acc, unsigned long long *cnt,
                                 unsigned long long *dsum)
{
    const int KD = a.K * a.D;
    for (int i = threadIdx.x; i < KD; i += blockDim.x) {
        const int k = i / a.D, d = i - k * a.D;
        const unsigned long long v = acc[(int64_t)k * a.rs + d];
        if (v) atomicAdd(&a.gsum[i], v);
    }
    for (int i = threadIdx.x; i < a.K; i += blockDim.x) {
        if (cnt[i]) atomicAdd(&a.gcnt[i], cnt[i]);
        if (dsum[i]) atomicAdd(&a.gdsum[i], dsum[i]);
    }
}

__device__ void zero_lds_object(const KmArgs &a, unsigned long long *acc, unsigned long long *cnt,
                                unsigned long long *dsum)
{
    for (int i = threadIdx.x; i < a.K * a.rs; i += blockDim.x) acc[i] = 0;
    for (int i = threadIdx.x; i < a.K; i += blockDim.x) {
        cnt[i] = 0;
        dsum[i] = 0;
    }
}

// Row stride of the LDS reduction object: D rounded up to odd, so that the
// rows of different keys start in different banks (with an even stride every
// row would start in the same bank, and the wavefront's atomics on one
// dimension of 32 points would serialise).
__host__ __device__ __forceinline__ int object_stride(int D) { return D | 1; }

// Words of the LDS reduction object, rounded up to a 16-byte boundary.
__host__ __device__ __forceinline__ int64_t object_words(int K, int D)
{
    return ((int64_t)K * object_stride(D) + 2 * K + 1) & ~1LL;
}

// LDS layout of the reduction object (when it fits).
__device__ __forceinline__ void lds_object(const KmArgs &a, unsigned long long *base, unsigned long long *&acc,
                                           unsigned long long *&cnt, unsigned long long *&dsum)
{
    if (a.use_lds_acc) {
        acc = base;
        cnt = base + (int64_t)a.K * a.rs;
        dsum = cnt + a.K;
    } else {
        acc = a.gsum;
        cnt = a.gcnt;
        dsum = a.gdsum;
    }
}

// ------------------------------------------------------------ plain path

// One thread per point; centres read from LDS when they fit (cs != null).
// BS threads per workgroup: the LDS object allows one workgroup per CU, so
// BS sets the waves per SIMD that hide LDS latency (4 at 1024 threads).
template <int DMAX, int BS>
__global__ __launch_bounds__(BS) void km_plain(KmArgs a, int centres_in_lds)
{
    extern __shared__ unsigned long long smem[];
    unsigned long long *acc, *cnt, *dsum;
    lds_object(a, smem, acc, cnt, dsum);
    // centres in LDS: rows of D4 = D rounded up to 4 floats, zero-padded
    float *cs = (float *)(smem + (a.use_lds_acc ? object_words(a.K, a.D) : 0));
    const int D = a.D, K = a.K, D4 = (D + 3) & ~3;
    if (a.use_lds_acc) zero_lds_object(a, acc, cnt, dsum);
    if (centres_in_lds)
        for (int i = threadIdx.x; i < K * D4; i += blockDim.x) {
            const int k = i / D4, d = i - k * D4;
            cs[i] = d < D ? a.c[(int64_t)k * D + d] : 0.0f;
        }
    __syncthreads();

    const int lane = threadIdx.x % kWave;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    const int64_t w0 = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    for (int64_t base = w0 * kWave; base < a.N; base += nwaves * kWave) {
        const int64_t p = base + lane;
        const bool live = p < a.N;
        float x[DMAX];
#pragma unroll
        for (int d = 0; d < DMAX; d++) x[d] = (live && d < D) ? a.x[p * D + d] : 0.0f;
        unsigned int key = 0;
        float min_dist = kInitMin;
        // dist = sum over d < D of (c_d - x_d)^2 in dimension order.  From
        // the LDS rows: groups of four dimensions (one 16-byte read, the same
        // address in every lane); past D both the row and x hold zeros, and
        // adding (0 - 0)^2 = +0 to the non-negative sum changes no bit.
        // From global memory: one guarded dimension at a time.
        auto dist_of = [&](int k) {
            float dist = 0;
            if (centres_in_lds) {
                const float *ci = cs + k * D4;
#pragma unroll
                for (int g = 0; g < DMAX / 4; g++) {
                    if (4 * g < D) {
                        const float4 c4 = *(const float4 *)(ci + 4 * g);
                        float t = c4.x - x[4 * g];
                        dist = dist + t * t;
                        t = c4.y - x[4 * g + 1];
                        dist = dist + t * t;
                        t = c4.z - x[4 * g + 2];
                        dist = dist + t * t;
                        t = c4.w - x[4 * g + 3];
                        dist = dist + t * t;
                    }
                }
            } else {
                const float *ci = a.c + (int64_t)k * D;
#pragma unroll
                for (int d = 0; d < DMAX; d++) {
                    if (d < D) {
                        const float t = ci[d] - x[d];
                        dist = dist + t * t;
                    }
                }
            }
            return sqrtf(dist);
        };
        int i = 0;
        for (; i + 4 <= K; i += 4) {
            // four independent chains for ILP; compared in index order
            float dv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) dv[q] = dist_of(i + q);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (dv[q] < min_dist) {
                    min_dist = dv[q];
                    key = i + q;
                }
        }
        for (; i < K; i++) {
            const float dist = dist_of(i);
            if (dist < min_dist) {
                min_dist = dist;
                key = i;
            }
        }
        if (live) {
            a.assign[p] = (int32_t)key;
            if (a.mindist) a.mindist[p] = min_dist;
        }
        const int np = (int)min<int64_t>(kWave, a.N - base);
        if (!a.defer_acc) {
            auto key_of = [&](int j) { return (int)__shfl((int)key, j, kWave); };
            auto dist_at = [&](int j) { return __shfl(min_dist, j, kWave); };
            // the LDS object addressed from the shared array itself, so that
            // its adds compile to LDS atomics
            if (a.use_lds_acc)
                accumulate_points(a, base, np, key_of, dist_at, smem, smem + K * a.rs, smem + K * a.rs + K);
            else
                accumulate_points(a, base, np, key_of, dist_at, a.gsum, a.gcnt, a.gdsum);
        }
    }
    if (a.use_lds_acc) {
        __syncthreads();
        flush_lds_object(a, acc, cnt, dsum);
    }
}

// ------------------------------------------------------------ MFMA path

// Per-point bound inputs: u = 2^-8 (BF16 unit roundoff), g = FP32
// accumulation over the MFMA's terms with a factor 2 of slack for any
// intermediate rounding mode.
struct MfmaBound {
    float c_dot;   // multiplies ||x|| * max||c||
    float c_norm;  // multiplies max||c||^2
    float c_ref;   // relative error of the reference's FP32 difference form
};

// Exact plain-path distance (sqrt of the FP32 difference form, dimensions in
// order) of point row xp to centre row ck, both read from memory (the point
// row was just loaded by the wavefront, so it is in cache).
__device__ __forceinline__ float exact_dist(const float *__restrict__ xp, const float *__restrict__ ck, int D, bool vec)
{
    float dist = 0;
    if (vec) {  // D % 4 == 0 and 16-byte aligned rows
        for (int d = 0; d < D; d += 4) {
            const float4 cv = *(const float4 *)(ck + d), xv = *(const float4 *)(xp + d);
            float t = cv.x - xv.x;
            dist = dist + t * t;
            t = cv.y - xv.y;
            dist = dist + t * t;
            t = cv.z - xv.z;
            dist = dist + t * t;
            t = cv.w - xv.w;
            dist = dist + t * t;
        }
    } else {
        for (int d = 0; d < D; d++) {
            const float t = ck[d] - xp[d];
            dist = dist + t * t;
        }
    }
    return sqrtf(dist);
}

// The same distance with the point row held by the wavefront's two halves
// (lane l and l+32 hold the point's dimensions 16s + 8*half + e in xo):
// v_permlane32_swap gives both lanes dimensions 16s + e in its first result
// and 16s + 8 + e in its second, so the row is read in order from registers
// instead of memory.  Called by every lane (the swap is cross-lane); the
// row is only formed where D == 16*KS.
template <int KS>
__device__ __forceinline__ void point_row(const float (&xo)[8 * KS], float (&row)[16 * KS])
{
#pragma unroll
    for (int s = 0; s < KS; s++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const unsigned v = __float_as_uint(xo[8 * s + e]);
            const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
            row[16 * s + e] = __uint_as_float(r[0]);
            row[16 * s + 8 + e] = __uint_as_float(r[1]);
        }
}

template <int KS>
__device__ __forceinline__ float exact_dist_row(const float (&row)[16 * KS], const float *__restrict__ ck)
{
    float dist = 0;
#pragma unroll
    for (int d = 0; d < 16 * KS; d += 4) {
        const float4 cv = *(const float4 *)(ck + d);
        float t = cv.x - row[d];
        dist = dist + t * t;
        t = cv.y - row[d + 1];
        dist = dist + t * t;
        t = cv.z - row[d + 2];
        dist = dist + t * t;
        t = cv.w - row[d + 3];
        dist = dist + t * t;
    }
    return sqrtf(dist);
}

// Approximate value v with its position idx (0..127) in the low 7 bits of
// the significand: a key that orders like v (up to 127 ulps) and carries its
// index through a plain float minimum.  |key - v| <= 2^-16 |v| + 2^-126.
// keep: ~127 held in a vector register (key_mask), so that with idx in a
// scalar register the packing is one v_and_or_b32 (an instruction reads at
// most one scalar operand, and the mask as a literal would be a second).
__device__ __forceinline__ float pack_key(float v, int idx, int keep)
{
    return __int_as_float((__float_as_int(v) & keep) | idx);
}

// ~127 in a vector register, opaque to constant folding.
__device__ __forceinline__ int key_mask()
{
    int m;
    asm volatile("v_mov_b32 %0, 0xffffff80" : "=v"(m));
    return m;
}

// base | i for a wavefront-uniform base, in one scalar register (a scalar
// ALU operation, opaque to the compiler, which would otherwise fold the OR
// into the packing as a second vector instruction).
__device__ __forceinline__ int scalar_or(int base, int i)
{
    int r;
    asm("s_or_b32 %0, %1, %2" : "=s"(r) : "s"(base), "s"(i));
    return r;
}

// Stand-in for a value past the last centre row: finite (so its key is not
// a NaN) and above any threshold.
constexpr float kFar = 3.0e38f;

// KS: K-steps of 16 dimensions holding the data (D <= 16*KS); NT: terms
// (1 or 3), each KS steps; one more K-step holds the norm terms.  KB:
// 32-centre blocks per chunk.  cb: K_pad x (16*(NT*KS+1)) BF16 centre rows
// [-2c | n1 n2 n3 0...] (NT = 1) or [hi(-2c) | lo(-2c) | hi(-2c) | n1 n2 n3
// 0...] (NT = 3), data zero-padded to 16*KS, stored in fragment order
// (frag_index); padded rows have n1 = 1e30 (far, but finite so that their
// keys are numbers).  cnorm: ||c_k|| (K entries).
//
// Per 32-point tile a wavefront (lanes l and l+32 hold halves of point l):
// computes the approximate values of every centre with MFMAs, packs each
// with its index (pack_key) and keeps the three smallest keys of each half
// (the two smallest per chunk when the index bits cannot tell the chunks
// apart).  The threshold is the minimum plus the error bound plus the
// packing error; when only the smallest key is under it (the usual case)
// that one centre is rescored, when a half's third key lies above it its
// kept keys under it are, and otherwise every centre under it is (the
// values recomputed chunk by chunk).  The exact FP32 results
// decide with map1's rule, and the point is added to the LDS reduction object
// from the registers holding its half rows.
// TAIL = false: D == 16*KS and the rows are 16-byte aligned, so every
// K-step is full and the loads need no per-element guard.
// GC: the form for K x D too large for LDS.  The workgroup's wavefronts
// step through the centre chunks together: before each chunk the
// workgroup copies its KB blocks of BF16 rows from global memory into one
// LDS chunk buffer (between two barriers) and every wavefront computes its
// own tile against it, so each copy of the rows serves 8 tiles.  Tiles are
// taken block-uniformly (a wavefront past the last tile computes with every
// lane dead) and the recompute path is a workgroup vote, so every wavefront
// reaches every barrier.
// BS: threads per workgroup (the LDS-resident centre rows and reduction
// object are shared by all its wavefronts, so larger workgroups raise the
// occupancy the LDS allows when the registers permit).
template <int KS, int KB, int NT, bool TAIL, bool GC, int BS = 512>
__global__ __launch_bounds__(BS) void km_mfma(KmArgs a, const __bf16 *__restrict__ cb, const float *__restrict__ cnorm,
                                               int K_pad, MfmaBound bd, unsigned long long *__restrict__ rescored)
{
    constexpr int NS = NT * KS + 1;  // K-steps per centre row
    constexpr int KW = 16 * NS;      // BF16 row width of cb
    static_assert(KB * 16 <= 128, "pack_key indexes at most 128 values per chunk");
    extern __shared__ unsigned long long smem[];
    unsigned long long *acc_o, *cnt_o, *dsum_o;
    lds_object(a, smem, acc_o, cnt_o, dsum_o);
    __bf16 *cbs = (__bf16 *)(smem + (a.use_lds_acc ? object_words(a.K, a.D) : 0));
    __shared__ float red[16];
    if (a.use_lds_acc) zero_lds_object(a, acc_o, cnt_o, dsum_o);
    // stage the BF16 centre matrix in LDS (16-byte copies); the GC form
    // stages one chunk at a time (stage_chunk)
    if (!GC) {
        const int n16 = K_pad * KW / 8;
        const bf16x8 *src = (const bf16x8 *)cb;
        bf16x8 *dst = (bf16x8 *)cbs;
        for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
    }
    const __bf16 *__restrict__ cfr = cbs;
    // max ||c|| over the centres (for the bound)
    float cm = 0.0f;
    for (int i = threadIdx.x; i < a.K; i += blockDim.x) cm = fmaxf(cm, cnorm[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cm = fmaxf(cm, __shfl_xor(cm, o, kWave));
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = cm;
    __syncthreads();
    float cmax = 0.0f;
    for (int w = 0; w < (int)(blockDim.x / kWave); w++) cmax = fmaxf(cmax, red[w]);
    const float cc_max = cmax * cmax;

    const int D = a.D;
    const bool vec_x = (D % 4) == 0 && ((uintptr_t)a.x % 16) == 0;  // rows start 16-byte aligned
    const bool vec_c = (D % 4) == 0 && ((uintptr_t)a.c % 16) == 0;
    const int lane = threadIdx.x % kWave;
    const int col = lane % 32;   // point within the tile
    const int half = lane / 32;  // which 8 of each 16 K elements this lane holds
    const int nchunks = (K_pad / 32 + KB - 1) / KB;
    unsigned my_rescored = 0;  // this lane's rescored points (under 2^32 per lane)

    // approximate values of chunk ch for this lane's point: value i of block
    // b is centre row 8*(i/4) + 4*half + i%4 of the block (the 32x32 MFMA
    // output layout).  The A fragments are read a few K-steps ahead of the
    // MFMA that uses them.
    // Blocks past the last centre row (when K_pad is not a multiple of
    // KB*32) are computed from block 0's rows -- the block index is clamped
    // on the scalar unit -- and skipped by the scans.
    const int kblocks = K_pad / 32;
    const int keep = key_mask();
    auto chunk_values = [&](int ch, const bf16x8 (&bx)[NS], f32x16 (&acc)[KB]) {
        constexpr int T = KB * NS, P = 3;
        const bf16x8 *ab = (const bf16x8 *)cfr + lane;
        auto frag = [&](int t) {
            if (GC) return ab[t * kWave];  // the staged chunk: block t / NS, K-step t % NS
            const int blk = ch * KB + t / NS;
            return ab[((int64_t)(blk < kblocks ? blk : 0) * NS + t % NS) * kWave];
        };
        bf16x8 fr[P];
#pragma unroll
        for (int t = 0; t < P; t++) fr[t] = frag(t);
#pragma unroll
        for (int t = 0; t < T; t++) {
            const int b = t / NS, st = t % NS;
            const bf16x8 av = fr[t % P];
            if (t + P < T) fr[t % P] = frag(t + P);
            if (st == 0) acc[b] = f32x16{};
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bx[st], acc[b], 0, 0, 0);
        }
    };
    auto centre_of = [&](int ch, int idx) {
        const int b = idx >> 4, i = idx & 15;
        return (ch * KB + b) * 32 + 8 * (i >> 2) + 4 * half + (i & 3);
    };
    // GC: chunk ch's blocks into the LDS chunk buffer (blocks past the last
    // centre row are not copied; their values are never scanned).  Called
    // by every thread of the workgroup.
    auto stage_chunk = [&](int ch) {
        __syncthreads();  // every wavefront is done with the previous chunk
        const int nblk = min(KB, kblocks - ch * KB);
        const int n16 = nblk * NS * kWave;
        const bf16x8 *src = (const bf16x8 *)cb + (int64_t)ch * KB * NS * kWave;
        bf16x8 *dst = (bf16x8 *)cbs;
        for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    };

    const int64_t ntiles = (a.N + 31) / 32;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / kWave);
    // this lane's half of tile t's point row (dimensions 16s + 8*half + e);
    // dead lanes read a valid row and keep zeros
    auto load_half_row = [&](int64_t t, float (&xo)[8 * KS]) {
        const int64_t p = t * 32 + col;
        const bool live = p < a.N;
        const float *xp = a.x + (live ? p : a.N - 1) * D;
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int d0 = 16 * s + 8 * half;
            if (!TAIL) {
                const float4 lo = *(const float4 *)(xp + d0), hi = *(const float4 *)(xp + d0 + 4);
                xo[8 * s] = lo.x; xo[8 * s + 1] = lo.y; xo[8 * s + 2] = lo.z; xo[8 * s + 3] = lo.w;
                xo[8 * s + 4] = hi.x; xo[8 * s + 5] = hi.y; xo[8 * s + 6] = hi.z; xo[8 * s + 7] = hi.w;
            } else if (vec_x) {
                float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
                if (live && d0 < D) lo = *(const float4 *)(xp + d0);
                if (live && d0 + 4 < D) hi = *(const float4 *)(xp + d0 + 4);
                xo[8 * s] = lo.x; xo[8 * s + 1] = lo.y; xo[8 * s + 2] = lo.z; xo[8 * s + 3] = lo.w;
                xo[8 * s + 4] = hi.x; xo[8 * s + 5] = hi.y; xo[8 * s + 6] = hi.z; xo[8 * s + 7] = hi.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) xo[8 * s + e] = (live && d0 + e < D) ? xp[d0 + e] : 0.0f;
            }
        }
    };
    // With the centre fragments in LDS the next tile's rows are loaded
    // while this tile computes (its loads return during the MFMAs); the
    // global-fragment form loads them in place, since vector loads complete
    // in order and the fragment loads would wait behind them, and so do rows
    // of more than 64 dimensions or in workgroups over 512 threads (the
    // extra registers would spill).
    constexpr bool PF = !GC && KS <= 4 && BS <= 512;
    // software-pipelined chunk scan (one more accumulator of KB x 16 values)
    constexpr bool PIPE = !GC && KB == 1;
    const int64_t tile0 = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    float xnext[8 * KS];
    if (PF && tile0 < ntiles) load_half_row(tile0, xnext);
    // GC: the workgroup's loop ends when its first wavefront's tile does
    const int64_t block_tile0 = (int64_t)blockIdx.x * (blockDim.x / kWave);
    for (int64_t tile = tile0; (GC ? block_tile0 + (tile - tile0) : tile) < ntiles; tile += nw) {
        const int64_t p = tile * 32 + col;
        const bool live = p < a.N;
        const float *xp = a.x + (live ? p : a.N - 1) * D;  // dead lanes read a valid row

        float xo[8 * KS];
        if (!PF) {
            load_half_row(tile, xo);
        } else {
#pragma unroll
            for (int i = 0; i < 8 * KS; i++) xo[i] = xnext[i];
            if (tile + nw < ntiles) load_half_row(tile + nw, xnext);
        }
        // B operand: point p, K elements [16s + 8*half, +8) of
        // [x, 1, 1, 1, 0...] or [xh, xh, xl, 1, 1, 1, 0...]
        bf16x8 bx[NS];
        float xx = 0.0f;
        // pairs of elements, one packed conversion each, straight into the
        // operand's 32-bit words (xh = bf16(x), xl = bf16(x - xh))
#pragma unroll
        for (int s = 0; s < KS; s++) {
            u32x4 wh, wl;
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                const float v0 = xo[8 * s + e], v1 = xo[8 * s + e + 1];
                xx = __builtin_fmaf(v0, v0, xx);  // a bound input only (used with slack)
                xx = __builtin_fmaf(v1, v1, xx);
                const unsigned h = bf16_pair(v0, v1);
                wh[e / 2] = h;
                if (NT == 3)
                    wl[e / 2] = bf16_pair(v0 - __uint_as_float(h << 16), v1 - __uint_as_float(h & 0xffff0000u));
            }
            bx[s] = __builtin_bit_cast(bf16x8, wh);
            if (NT == 3) {
                bx[KS + s] = bx[s];
                bx[2 * KS + s] = __builtin_bit_cast(bf16x8, wl);
            }
        }
#pragma unroll
        for (int e = 0; e < 8; e++) bx[NT * KS][e] = (__bf16)((half == 0 && e < 3) ? 1.0f : 0.0f);
        xx = xx + __shfl_xor(xx, 32, kWave);  // ||x||^2 of point p (both halves)
        // an upper bound of ||x|| only: the hardware square root (within an
        // ulp or two for normal inputs, 0 for flushed subnormals, whose
        // root lies below 2^-63) with relative and absolute slack, in place
        // of the correctly rounded sequence
        const float xn = __builtin_amdgcn_sqrtf(xx) * 1.0001f + 0x1p-40f;
        const float E = bd.c_dot * xn * cmax + bd.c_norm * (cc_max + xx);
        // packing error bound: |values| <= ||c||^2 + 2 ||x|| ||c||
        const float dp = 0x1p-16f * (cc_max + 2.0f * xn * cmax) * 1.01f + 0x1p-126f;

        f32x16 acc[KB];
        // The three smallest keys of this half.  When the lane's values of all
        // chunks fit the 7 index bits (gidx: at most 128 per lane), every key
        // carries the lane's global value index (value i of block blk ->
        // 16 blk + i) and one sorted triple runs over all chunks; otherwise
        // each chunk's two smallest are merged with their chunk recorded and
        // k3 stays unknown (kFar, never used).
        const bool gidx = nchunks * KB <= 8;  // uniform
        float k1 = kFar, k2 = kFar, k3 = kFar;
        int ch1 = 0;
        auto scan3 = [&](int ch, const f32x16 (&v)[KB]) {
#pragma unroll
            for (int b = 0; b < KB; b++) {
                if (b > 0 && ch * KB + b >= kblocks) break;  // uniform; block 0 of a chunk always exists
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const float key = pack_key(v[b][i], scalar_or((ch * KB + b) * 16, i), keep);
                    k3 = __builtin_amdgcn_fmed3f(k2, k3, key);
                    k2 = __builtin_amdgcn_fmed3f(k1, k2, key);
                    k1 = __builtin_amdgcn_fmed3f(k1, key, -kFar);  // min (keys are finite)
                }
            }
        };
        if (PIPE && gidx) {
            // the next chunk's MFMAs are issued before this chunk's keys are
            // scanned, so one wavefront keeps the matrix cores and the vector
            // ALU busy together; two accumulators take the roles in turn
            // (no branch between a chunk's MFMAs and the other's scan, so the
            // scheduler can interleave them)
            f32x16 acc2[KB];
            chunk_values(0, bx, acc);
            int ch = 0;
            for (; ch + 2 < nchunks; ch += 2) {
                chunk_values(ch + 1, bx, acc2);
                scan3(ch, acc);
                chunk_values(ch + 2, bx, acc);
                scan3(ch + 1, acc2);
            }
            if (ch + 1 < nchunks) {
                chunk_values(ch + 1, bx, acc2);
                scan3(ch, acc);
                scan3(ch + 1, acc2);
            } else {
                scan3(ch, acc);
            }
        }
        for (int ch = 0; ch < (PIPE && gidx ? 0 : nchunks); ch++) {
            if (GC) stage_chunk(ch);
            chunk_values(ch, bx, acc);
            if (gidx) {
                scan3(ch, acc);
                continue;
            }
            float c1 = kFar, c2 = kFar;
#pragma unroll
            for (int b = 0; b < KB; b++) {
                if (ch * KB + b >= kblocks) break;  // uniform
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const float key = pack_key(acc[b][i], b * 16 + i, keep);
                    c2 = __builtin_amdgcn_fmed3f(c1, c2, key);
                    c1 = __builtin_amdgcn_fmed3f(c1, key, -kFar);  // min (keys are finite)
                }
            }
            if (c1 < k1) {
                k2 = fminf(k1, c2);
                k1 = c1;
                ch1 = ch;
            } else {
                k2 = fminf(k2, c1);
            }
        }
        // both halves: overall smallest and second smallest
        const float o1 = __shfl_xor(k1, 32, kWave), o2 = __shfl_xor(k2, 32, kWave);
        const float m1 = fminf(k1, o1);
        const float m2 = fminf(fmaxf(k1, o1), fminf(k2, o2));
        // candidates: keys within the margin of the minimum
        const float dhi = (xx + m1 + 4.0f * E + 4.0f * dp) * 1.001f + 1e-30f;
        const float thr = m1 + 2.0f * E + 2.0f * dp + 2.0f * bd.c_ref * dhi;
        const bool multi = !(m2 > thr);
        // lanes whose candidates may lie beyond the keys they kept: every
        // value under the threshold is then visited again (full)
        const bool full = multi && (!gidx || !(k3 > thr));
        // the whole point row in registers (full K-steps) where the form has
        // the registers to spare (the others would spill: they reload it)
        constexpr bool ROW = !TAIL && !GC && KS == 2 && KB == 1;
        float row[16 * KS];
        if (ROW) point_row<KS>(xo, row);

        float best_d = kInitMin;
        int best_k = 0x7fffffff;
        // exact distance to value idx of chunk ch (the pack_key index)
        auto rescore = [&](int ch, int idx) {
            const int k = centre_of(ch, idx);
            if (k < a.K) {
                const float *ck = a.c + (int64_t)k * D;
                const float dk = ROW && vec_c ? exact_dist_row<KS>(row, ck) : exact_dist(xp, ck, D, vec_c && vec_x);
                if (dk < best_d || (dk == best_d && k < best_k)) {
                    best_d = dk;
                    best_k = k;
                }
            }
        };
        if (live && !full) {
            // one candidate: rescored by the half that holds it; several
            // (gidx only): each half rescores its kept keys under the
            // threshold, which are all of its candidates since its third
            // smallest lies above it
            if (multi ? !(k1 > thr) : k1 < o1) rescore(gidx ? 0 : ch1, __float_as_int(k1) & 127);
            if (multi && !(k2 > thr)) rescore(0, __float_as_int(k2) & 127);
        }
        // Every value under the threshold for the full lanes.  Taken by the
        // whole wavefront (a vote), because recomputing a chunk's values
        // needs every lane's operands.
        if (GC ? __syncthreads_or(full && live) != 0 : __any(full)) {
            for (int ch = 0; ch < nchunks; ch++) {
                if (GC) {
                    stage_chunk(ch);
                    chunk_values(ch, bx, acc);
                } else if (nchunks > 1) {
                    chunk_values(ch, bx, acc);
                }
                if (!full) continue;
                uint64_t w0 = 0, w1 = 0;
#pragma unroll
                for (int b = 0; b < KB; b++) {
                    if (ch * KB + b >= kblocks) break;  // uniform
#pragma unroll
                    for (int i = 0; i < 16; i++)
                        if (pack_key(acc[b][i], b * 16 + i, keep) <= thr) {
                            if (b < 4) w0 |= 1ULL << (b * 16 + i);
                            else w1 |= 1ULL << ((b - 4) * 16 + i);
                        }
                }
                for (int w = 0; w < 2; w++) {
                    uint64_t mw = w ? w1 : w0;
                    while (mw) {
                        const int bit = __builtin_ctzll(mw);
                        mw &= mw - 1;
                        if (live) rescore(ch, 64 * w + bit);
                    }
                }
            }
        }
        // merge the two half-waves' exact winners (same tie rule)
        const float od = __shfl_xor(best_d, 32, kWave);
        const int ok = __shfl_xor(best_k, 32, kWave);
        if (od < best_d || (od == best_d && ok < best_k)) {
            best_d = od;
            best_k = ok;
        }
        if (best_k == 0x7fffffff) best_k = 0;  // unreachable for live points
        if (live && a.defer_acc) {
            if (half == 0) {
                a.assign[p] = best_k;
                a.mindist[p] = best_d;
                if (multi) my_rescored++;
            }
        } else if (live) {
            if (half == 0) {
                a.assign[p] = best_k;
                if (a.mindist) a.mindist[p] = best_d;
                if (multi) my_rescored++;
            }
            // the point (count, distance, this half-wave's dimensions) into
            // the reduction object; the LDS form addresses the shared array
            // itself, so that its adds compile to LDS atomics
            auto add_point = [&](unsigned long long *acc, unsigned long long *cnt, unsigned long long *dsum) {
                if (half == 0) {
                    atomicAdd(&cnt[best_k], 1ULL);
                    atomicAdd(&dsum[best_k], to_fixed(best_d, a.dscale));
                }
                unsigned long long *row = acc + best_k * a.rs;
#pragma unroll
                for (int s = 0; s < KS; s++)
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const int d = 16 * s + 8 * half + e;
                        if (!TAIL || d < D) atomicAdd(&row[d], to_fixed(xo[8 * s + e], a.scale));
                    }
            };
            if (a.use_lds_acc) add_point(smem, smem + a.K * a.rs, smem + a.K * a.rs + a.K);
            else add_point(a.gsum, a.gcnt, a.gdsum);
        }
    }
    // rescored-point counter
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) my_rescored += __shfl_xor(my_rescored, o, kWave);
    if (lane == 0 && my_rescored) atomicAdd(rescored, (unsigned long long)my_rescored);
    if (a.use_lds_acc) {
        __syncthreads();
        flush_lds_object(a, acc_o, cnt_o, dsum_o);
    }
}

// ------------------------------------------------------------ update

// One wavefront per centre: applies the Lloyd update from the reduction
// object (c = (sum * 2^-s) / count in FP64, rounded to FP32; empty clusters
// keep their centre), writes the BF16 MFMA row and the norm, and clears the
// object for the next pass.  With update == 0 only the rows are prepared.
template <int KS, int NT>
__global__ __launch_bounds__(256) void km_finalize(float *__restrict__ c, int K, int K_pad, int D,
                                                   unsigned long long *__restrict__ gsum,
                                                   unsigned long long *__restrict__ gcnt,
                                                   unsigned long long *__restrict__ gdsum, double inv_scale,
                                                   int update, __bf16 *__restrict__ cb, float *__restrict__ cnorm,
                                                   unsigned long long *__restrict__ csnap)
{
    constexpr int KW = 16 * (NT * KS + 1);
    constexpr int NB = 16 * NT * KS;  // first norm column
    const int k = blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (k >= K_pad) return;
    if (k >= K) {  // padding row: never the minimum
        for (int e = lane; e < KW; e += kWave)
            cb[frag_index(k, e, KW / 16)] = (__bf16)((e == NB) ? 1.0e30f : 0.0f);  // far, finite
        return;
    }
    const long long n = (long long)gcnt[k];
    double ss = 0.0;
    for (int d = lane; d < D; d += kWave) {
        float v = c[(int64_t)k * D + d];
        if (update && n > 0) v = (float)(((double)(long long)gsum[(int64_t)k * D + d] * inv_scale) / (double)n);
        if (update) c[(int64_t)k * D + d] = v;
        ss += (double)v * (double)v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, kWave);
    for (int d = lane; d < 16 * KS; d += kWave) {
        const float v = d < D ? -2.0f * c[(int64_t)k * D + d] : 0.0f;  // exact
        const __bf16 h = (__bf16)v;
        cb[frag_index(k, d, KW / 16)] = h;
        if (NT == 3) {
            cb[frag_index(k, 16 * KS + d, KW / 16)] = (__bf16)(v - (float)h);
            cb[frag_index(k, 32 * KS + d, KW / 16)] = h;
        }
    }
    if (lane < 16) {
        const __bf16 n1 = (__bf16)(float)ss;
        const double r1 = ss - (double)(float)n1;
        const __bf16 n2 = (__bf16)(float)r1;
        const double r2 = r1 - (double)(float)n2;
        const __bf16 n3 = (__bf16)(float)r2;
        __bf16 v = (__bf16)0.0f;
        if (lane == 0) v = n1;
        if (lane == 1) v = n2;
        if (lane == 2) v = n3;
        cb[frag_index(k, NB + lane, KW / 16)] = v;
    }
    if (lane == 0) {
        cnorm[k] = (float)sqrt(ss) * 1.0001f;
        if (csnap) csnap[k] = (unsigned long long)n;
    }
    // clear the reduction object for the next pass (a fill, not a memset, so
    // the sequence can be captured in a graph)
    for (int d = lane; d < D; d += kWave) gsum[(int64_t)k * D + d] = 0;
    if (lane == 0) {
        gcnt[k] = 0;
        gdsum[k] = 0;
    }
}

// 256 threads per workgroup.  Sixteen-byte loads, four in flight per thread,
// when x is 16-byte aligned (else one float at a time); the workgroup's
// maximum reaches *out by one atomic (atomics from every wavefront to one
// word serialise).  The maximum does not depend on the order.
__global__ __launch_bounds__(256) void km_maxabs(const float *__restrict__ x, int64_t n,
                                                 unsigned int *__restrict__ out)
{
    __shared__ float part[4];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (((uintptr_t)x & 15) == 0) {
        auto mx = [](const float4 v) { return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))); };
        const float4 *__restrict__ x4 = (const float4 *)x;
        const int64_t n4 = n / 4;
        int64_t i = gid;
        for (; i + 3 * stride < n4; i += 4 * stride) {
            const float4 v0 = x4[i], v1 = x4[i + stride], v2 = x4[i + 2 * stride], v3 = x4[i + 3 * stride];
            m = fmaxf(m, fmaxf(fmaxf(mx(v0), mx(v1)), fmaxf(mx(v2), mx(v3))));
        }
        for (; i < n4; i += stride) m = fmaxf(m, mx(x4[i]));
        if (gid < n - 4 * n4) m = fmaxf(m, fabsf(x[4 * n4 + gid]));  // the last n % 4 values
    } else {
        for (int64_t i = gid; i < n; i += stride) m = fmaxf(m, fabsf(x[i]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, kWave));
    if (threadIdx.x % kWave == 0) part[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
        if (b > 0.0f) atomicMax(out, __float_as_uint(b));
    }
}

// ------------------------------------------------------------ bucketed reduction
//
// For a reduction object too large for LDS the assignment kernels only
// write assign/mindist, and the object is summed afterwards bucket by
// bucket: points are counted per centre, the counts scanned, the point
// indices scattered into their centre's bucket, and wavefronts walk the
// buckets in runs of kRun positions, summing each centre's rows in
// registers (the same fixed-point integers as the direct path, so the sums
// are the same bit for bit) and adding them to the object once per centre
// run -- about (N / kRun + K) * D atomics instead of N * (D + 2).

constexpr int kRun = 256;

// cnt[k] += points assigned to k (LDS histogram per workgroup when K fits).
__global__ __launch_bounds__(256) void km_bucket_count(const int32_t *__restrict__ assign, int64_t N, int K,
                                                       int *__restrict__ cnt, int lds_hist)
{
    extern __shared__ int hist[];
    if (lds_hist) {
        for (int k = threadIdx.x; k < K; k += blockDim.x) hist[k] = 0;
        __syncthreads();
    }
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        const int k = assign[p];
        if (lds_hist) atomicAdd(&hist[k], 1);
        else atomicAdd(&cnt[k], 1);
    }
    if (lds_hist) {
        __syncthreads();
        for (int k = threadIdx.x; k < K; k += blockDim.x)
            if (hist[k]) atomicAdd(&cnt[k], hist[k]);
    }
}

// One workgroup of 1024 threads: off[k] = exclusive prefix sum of cnt
// (K <= 65536 entries; off has K + 1), and cur = off (scatter cursors).
__global__ __launch_bounds__(1024) void km_bucket_scan(const int *__restrict__ cnt, int K, int *__restrict__ off,
                                                       int *__restrict__ cur)
{
    __shared__ int part[1024];
    const int per = (K + 1023) / 1024;
    const int k0 = threadIdx.x * per, k1 = min(K, k0 + per);
    int sum = 0;
    for (int k = k0; k < k1; k++) sum += cnt[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
        const int v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - sum;
    for (int k = k0; k < k1; k++) {
        off[k] = run;
        cur[k] = run;
        run += cnt[k];
    }
    if (threadIdx.x == 1023) off[K] = part[1023];
}

// idx[slot] = p, each point into its centre's bucket (order within a bucket
// is arbitrary: the sums are integers).
__global__ __launch_bounds__(256) void km_bucket_scatter(const int32_t *__restrict__ assign, int64_t N,
                                                         int *__restrict__ cur, int *__restrict__ idx)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x)
        idx[atomicAdd(&cur[assign[p]], 1)] = (int)p;
}

// One wavefront per run of kRun bucket positions: lanes hold dimensions
// lane + 64t; each centre's points in the run are summed in registers and
// added to the object when the centre changes (and at the run's end).
template <int DT>
__global__ __launch_bounds__(256) void km_bucket_sum(const float *__restrict__ x, const int *__restrict__ idx,
                                                     const int32_t *__restrict__ assign,
                                                     const float *__restrict__ mindist, int64_t N, int D, float scale,
                                                     float dscale, unsigned long long *__restrict__ gsum,
                                                     unsigned long long *__restrict__ gcnt,
                                                     unsigned long long *__restrict__ gdsum)
{
    const int64_t run = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    const int64_t j0 = run * kRun;
    if (j0 >= N) return;  // whole wavefronts
    const int64_t j1 = min<int64_t>(N, j0 + kRun);
    unsigned long long acc[DT], cnt = 0, dsum = 0;
#pragma unroll
    for (int t = 0; t < DT; t++) acc[t] = 0;
    int kc = assign[idx[j0]];
    auto flush = [&](int k) {
#pragma unroll
        for (int t = 0; t < DT; t++) {
            const int d = lane + kWave * t;
            if (d < D && acc[t]) atomicAdd(&gsum[(int64_t)k * D + d], acc[t]);
            acc[t] = 0;
        }
        if (lane == 0) {
            atomicAdd(&gcnt[k], cnt);
            atomicAdd(&gdsum[k], dsum);
        }
        cnt = 0;
        dsum = 0;
    };
    for (int64_t j = j0; j < j1; j++) {
        const int p = idx[j];
        const int k = assign[p];
        if (k != kc) {  // uniform: every lane reads the same p
            flush(kc);
            kc = k;
        }
        const float *xp = x + (int64_t)p * D;
#pragma unroll
        for (int t = 0; t < DT; t++) {
            const int d = lane + kWave * t;
            if (d < D) acc[t] += to_fixed(xp[d], scale);
        }
        cnt++;
        if (lane == 0) dsum += to_fixed(mindist[p], dscale);
    }
    flush(kc);
}

__global__ void km_zero_u64(unsigned long long *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void km_zero_i32(int *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

namespace pat {

// Fixed-point shift for n values of magnitude < maxabs: each value times
// 2^s stays within 2^30 (so it converts to int32) and the largest possible
// sum n * maxabs * 2^s below 2^62.  Same formula as the oracle's.
int kmeans_fixed_shift(int64_t n, float maxabs)
{
    int eN = 0, eM = 0;
    while ((1LL << eN) < n) eN++;
    double m = maxabs > 0 ? maxabs : 1.0f;
    frexp(m, &eM);
    int s = std::min(62 - eN - eM, 30 - eM);
    if (s > 60) s = 60;
    return s;
}

int kmeans_ks_for(int D)
{
    if (D <= 16) return 1;
    if (D <= 32) return 2;
    if (D <= 64) return 4;
    if (D <= 128) return 8;
    return 0;
}

// 32-centre blocks per chunk (accumulator registers: 16 per block).
int kmeans_kb_for(int ks) { return ks <= 2 ? 8 : (ks == 4 ? 4 : 2); }

// LDS bytes of the reduction object and the staged centres.
static int64_t object_bytes(int K, int D) { return 8LL * object_words(K, D); }

int kmeans_maxabs(const float *x, int64_t n, unsigned int *out_bits, hipStream_t st)
{
    hipLaunchKernelGGL(km_zero_u64, dim3(1), dim3(1), 0, st, (unsigned long long *)out_bits, 1);
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(km_maxabs, dim3(1024), dim3(256), 0, st, x, n, out_bits);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}


// Largest dynamic LDS a workgroup may request on this device (queried once).
static int64_t g_lds_limit = 64 * 1024;

static KmArgs make_args(const KmDev &d)
{
    KmArgs a;
    a.x = d.x;
    a.c = d.c;
    a.N = d.N;
    a.D = d.D;
    a.K = d.K;
    a.scale = ldexpf(1.0f, d.shift);
    a.dscale = ldexpf(1.0f, d.dshift);
    a.assign = d.assign;
    a.mindist = d.mindist;
    a.gsum = d.gsum;
    a.gcnt = d.gcnt;
    a.gdsum = d.gdsum;
    a.use_lds_acc = kmeans_object_in_lds(d.K, d.D);
    a.defer_acc = d.bucketed;
    a.rs = a.use_lds_acc ? object_stride(d.D) : d.D;
    return a;
}


// Raises every kmeans kernel's dynamic-LDS limit to what the device allows
// per workgroup (160 KB on gfx950).  Best effort: the per-block limit the
// device reports is what the launch checks use.  Called once before any
// launch (outside graph capture).
int kmeans_prepare_kernels()
{
    static int done = 0;
    if (done) return PAT_OK;
    int dev = 0, lim = 0;
    PAT_HIP(hipGetDevice(&dev));
    PAT_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    g_lds_limit = lim - 256;  // room for the kernels' static LDS
    const void *fns[] = {
        (const void *)km_plain<8, 1024>, (const void *)km_plain<16, 1024>, (const void *)km_plain<32, 1024>,
        (const void *)km_plain<64, 1024>, (const void *)km_plain<128, 512>,
        // one term
        (const void *)km_mfma<1, 8, 1, true, false>, (const void *)km_mfma<2, 8, 1, true, false>,
        (const void *)km_mfma<4, 4, 1, true, false>, (const void *)km_mfma<8, 2, 1, true, false>,
        (const void *)km_mfma<1, 8, 1, false, false>, (const void *)km_mfma<2, 8, 1, false, false>,
        (const void *)km_mfma<4, 4, 1, false, false>, (const void *)km_mfma<8, 2, 1, false, false>,
        // three terms, rows in LDS
        (const void *)km_mfma<1, 4, 3, true, false, 768>, (const void *)km_mfma<1, 4, 3, false, false, 768>,
        (const void *)km_mfma<2, 1, 3, true, false, 768>, (const void *)km_mfma<2, 1, 3, false, false, 768>,
        (const void *)km_mfma<4, 2, 3, true, false, 768>, (const void *)km_mfma<4, 2, 3, false, false, 768>,
        (const void *)km_mfma<8, 2, 3, true, false>, (const void *)km_mfma<8, 2, 3, false, false>,
        // three terms, rows staged chunk by chunk
        (const void *)km_mfma<1, 8, 3, true, true>, (const void *)km_mfma<2, 8, 3, true, true>,
        (const void *)km_mfma<4, 4, 3, true, true>, (const void *)km_mfma<8, 2, 3, true, true>,
        (const void *)km_mfma<1, 8, 3, false, true>, (const void *)km_mfma<2, 8, 3, false, true>,
        (const void *)km_mfma<4, 4, 3, false, true>, (const void *)km_mfma<8, 2, 3, false, true>};
    for (const void *f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g_lds_limit) != hipSuccess)
            (void)hipGetLastError();
    done = 1;
    return PAT_OK;
}

// Whether the reduction object (K x D sums, counts, distance sums) is kept
// in LDS by the assignment kernels; otherwise a handle with fewer than 2^31
// points sums it with the bucketed reduction.
bool kmeans_object_in_lds(int K, int D)
{
    return object_bytes(K, D) <= std::min<int64_t>(96 * 1024, g_lds_limit - 32 * 1024);
}

// The bucketed reduction after an assignment pass (see km_bucket_sum).
static int bucket_reduce(const KmDev &d, hipStream_t st)
{
    const int K = d.K;
    hipLaunchKernelGGL(km_zero_i32, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st, d.bcnt, (int64_t)K);
    PAT_LAUNCH_CHECK();
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((d.N + 255) / 256, 8LL * d.cus));
    const int lds_hist = K <= 8192;
    hipLaunchKernelGGL(km_bucket_count, dim3(g), dim3(256), lds_hist ? (size_t)4 * K : 0, st, (const int32_t *)d.assign,
                       d.N, K, d.bcnt, lds_hist);
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(km_bucket_scan, dim3(1), dim3(1024), 0, st, (const int *)d.bcnt, K, d.boff, d.bcur);
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(km_bucket_scatter, dim3(g), dim3(256), 0, st, (const int32_t *)d.assign, d.N, d.bcur, d.bidx);
    PAT_LAUNCH_CHECK();
    const int64_t runs = (d.N + kRun - 1) / kRun;
    const unsigned gs = (unsigned)((runs + 3) / 4);
    const float scale = ldexpf(1.0f, d.shift), dscale = ldexpf(1.0f, d.dshift);
    if (d.D <= 64)
        hipLaunchKernelGGL(km_bucket_sum<1>, dim3(gs), dim3(256), 0, st, d.x, (const int *)d.bidx,
                           (const int32_t *)d.assign, (const float *)d.mindist, d.N, d.D, scale, dscale, d.gsum, d.gcnt,
                           d.gdsum);
    else
        hipLaunchKernelGGL(km_bucket_sum<2>, dim3(gs), dim3(256), 0, st, d.x, (const int *)d.bidx,
                           (const int32_t *)d.assign, (const float *)d.mindist, d.N, d.D, scale, dscale, d.gsum, d.gcnt,
                           d.gdsum);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

static int allow_lds(const void *, int64_t bytes)
{
    PAT_REQUIRE(bytes <= g_lds_limit, PAT_ERR_UNSUPPORTED, "kmeans: %lld bytes of LDS requested, device allows %lld",
                (long long)bytes, (long long)g_lds_limit);
    return PAT_OK;
}

int kmeans_finalize(const KmDev &d, int update, hipStream_t st)
{
    const double inv = ldexp(1.0, -d.shift);
    const dim3 g((unsigned)((d.K_pad + 3) / 4)), b(256);
#define PAT_KF(KS_, NT_)                                                                                               \
    hipLaunchKernelGGL((km_finalize<KS_, NT_>), g, b, 0, st, d.c, d.K, d.K_pad, d.D, d.gsum, d.gcnt, d.gdsum, inv,    \
                       update, (__bf16 *)d.cb, d.cnorm, d.csnap)
    const int code = d.KS * 4 + d.NT;
    switch (code) {
    case 1 * 4 + 1: PAT_KF(1, 1); break;
    case 2 * 4 + 1: PAT_KF(2, 1); break;
    case 4 * 4 + 1: PAT_KF(4, 1); break;
    case 8 * 4 + 1: PAT_KF(8, 1); break;
    case 1 * 4 + 3: PAT_KF(1, 3); break;
    case 2 * 4 + 3: PAT_KF(2, 3); break;
    case 4 * 4 + 3: PAT_KF(4, 3); break;
    case 8 * 4 + 3: PAT_KF(8, 3); break;
    default: pat::set_error("kmeans: D = %d or terms %d not supported", d.D, d.NT); return PAT_ERR_UNSUPPORTED;
    }
#undef PAT_KF
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int kmeans_assign_plain(const KmDev &d, hipStream_t st)
{
    KmArgs a = make_args(d);
    const int64_t cbytes = 4LL * d.K * ((d.D + 3) & ~3);  // rows padded to 4 floats
    int64_t lds = a.use_lds_acc ? object_bytes(d.K, d.D) : 0;
    int in_lds = lds + cbytes <= g_lds_limit;
    if (in_lds) lds += cbytes;
    const int64_t want = (d.N + 1023) / 1024;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)d.cus));
#define PAT_KP(DM)                                                                                                     \
    do {                                                                                                               \
        constexpr int bs = DM <= 64 ? 1024 : 512;                                                                      \
        PAT_TRY(allow_lds((const void *)km_plain<DM, bs>, lds));                                                       \
        hipLaunchKernelGGL((km_plain<DM, bs>), dim3(grid), dim3(bs), (size_t)lds, st, a, in_lds);                      \
    } while (0)
    if (d.D <= 8) PAT_KP(8);
    else if (d.D <= 16) PAT_KP(16);
    else if (d.D <= 32) PAT_KP(32);
    else if (d.D <= 64) PAT_KP(64);
    else if (d.D <= 128) PAT_KP(128);
    else {
        pat::set_error("kmeans: D = %d above 128", d.D);
        return PAT_ERR_UNSUPPORTED;
    }
#undef PAT_KP
    PAT_LAUNCH_CHECK();
    return d.bucketed ? bucket_reduce(d, st) : PAT_OK;
}

// LDS bytes of the MFMA kernel with nt terms (staged BF16 rows plus the
// reduction object when it is kept in LDS).
static int64_t mfma_lds_bytes(int K, int D, int nt, bool gc = false)
{
    const int ks = kmeans_ks_for(D);
    const int K_pad = (K + 31) / 32 * 32;
    // all rows, or (gc) one chunk: KB blocks of 32 rows (KB = 8, 8, 4, 2 for
    // KS = 1, 2, 4, 8)
    const int kb = kmeans_kb_for(ks);
    const int64_t cbb = gc ? 2LL * std::min(K_pad, 32 * kb) * 16 * (nt * ks + 1) : 2LL * K_pad * 16 * (nt * ks + 1);
    const int64_t obj = object_bytes(K, D) <= std::min<int64_t>(96 * 1024, g_lds_limit - 32 * 1024)
                            ? object_bytes(K, D) : 0;
    return cbb + obj;
}

// Terms of the MFMA path for this shape: 3 (split BF16) when its rows fit in
// LDS, else 1, 0 when neither fits.
int kmeans_nt_for(int K, int D)
{
    if (!kmeans_ks_for(D)) return 0;
    if (mfma_lds_bytes(K, D, 3) <= g_lds_limit) return 3;
    if (mfma_lds_bytes(K, D, 1) <= g_lds_limit) return 1;
    return 0;
}

// Whether the MFMA path reads its rows from global memory: when neither
// LDS form fits (it then uses three terms).
bool kmeans_gc_for(int K, int D) { return kmeans_ks_for(D) != 0 && kmeans_nt_for(K, D) == 0; }

bool kmeans_mfma_fits(int K, int D) { return kmeans_ks_for(D) != 0 && K <= 65536; }

int kmeans_assign_mfma(const KmDev &d, hipStream_t st)
{
    KmArgs a = make_args(d);
    PAT_REQUIRE(d.NT == 1 || d.NT == 3, PAT_ERR_UNSUPPORTED, "kmeans: K=%d, D=%d too large for the MFMA path", d.K,
                d.D);
    PAT_REQUIRE(!d.gc || d.NT == 3, PAT_ERR_STATE, "kmeans: global-row MFMA form needs three terms");
    const int64_t lds = mfma_lds_bytes(d.K, d.D, d.NT, d.gc != 0);
    PAT_REQUIRE(lds <= g_lds_limit, PAT_ERR_UNSUPPORTED, "kmeans: K=%d, D=%d too large for the MFMA path", d.K, d.D);
    // Error-bound constants.  u = 2^-8 (BF16 unit roundoff); g = FP32
    // accumulation over every term of the MFMA chain, with a factor 2 of
    // slack for any internal rounding mode.
    //  NT = 1: |x.c - bf(x).bf(c)| <= (2u + u^2) sum |x_d c_d|.
    //  NT = 3: x = xh + xl + rx with |rx| <= u^2 |x|, |xl| <= u(1+u)|x| (same
    //          for c); the dropped xl.cl + (terms with rx, rc) are bounded by
    //          (3u^2 + 4u^3)(1+u)^2 sum |x_d c_d|.
    //  Both scaled by 2 for the -2 factor; the norm split leaves u^3 ||c||^2.
    const double u = ldexp(1.0, -8), g = 2.0 * (16.0 * (d.NT * d.KS + 1)) * ldexp(1.0, -24);
    MfmaBound bd;
    if (d.NT == 3)
        bd.c_dot = (float)((2 * (3 * u * u + 4 * u * u * u) * (1 + u) * (1 + u) +
                            2 * g * (1 + u) * (1 + u) * (1 + 2 * u)) * 1.01);
    else
        bd.c_dot = (float)((4 * u + 2 * u * u + 4 * g * (1 + u) * (1 + u)) * 1.01);
    bd.c_norm = (float)((u * u * u + 2 * g * (1 + 2 * u)) * 1.01);
    bd.c_ref = (float)((3.0 + d.D + 16.0) * ldexp(1.0, -24) * 1.01);
    // one workgroup per CU (the LDS holds one), BS_ / 64 wavefronts each
    auto grid_for = [&](int bs) {
        const int64_t want = (d.N + 31) / 32 / (bs / 64);
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)d.cus));
    };
    // no tail: every K-step full and 16-byte aligned rows
    const bool tail = d.D != 16 * d.KS || ((uintptr_t)d.x % 16) != 0;
#define PAT_KM_BS(KS_, KB_, NT_, GC_, BS_)                                                                             \
    do {                                                                                                               \
        if (tail) {                                                                                                    \
            PAT_TRY(allow_lds((const void *)km_mfma<KS_, KB_, NT_, true, GC_, BS_>, lds));                             \
            hipLaunchKernelGGL((km_mfma<KS_, KB_, NT_, true, GC_, BS_>), dim3(grid_for(BS_)), dim3(BS_), (size_t)lds,  \
                               st, a, (const __bf16 *)d.cb, d.cnorm, d.K_pad, bd, d.rescored);                         \
        } else {                                                                                                       \
            PAT_TRY(allow_lds((const void *)km_mfma<KS_, KB_, NT_, false, GC_, BS_>, lds));                            \
            hipLaunchKernelGGL((km_mfma<KS_, KB_, NT_, false, GC_, BS_>), dim3(grid_for(BS_)), dim3(BS_), (size_t)lds, \
                               st, a, (const __bf16 *)d.cb, d.cnorm, d.K_pad, bd, d.rescored);                         \
        }                                                                                                              \
    } while (0)
#define PAT_KM_GC(KS_, KB_, NT_, GC_) PAT_KM_BS(KS_, KB_, NT_, GC_, 512)
#define PAT_KM(KS_, KB_, NT_) PAT_KM_GC(KS_, KB_, NT_, false)
    if (d.gc) {
        switch (d.KS) {
        case 1: PAT_KM_GC(1, 8, 3, true); break;
        case 2: PAT_KM_GC(2, 8, 3, true); break;
        case 4: PAT_KM_GC(4, 4, 3, true); break;
        case 8: PAT_KM_GC(8, 2, 3, true); break;
        default: pat::set_error("kmeans: D too large"); return PAT_ERR_UNSUPPORTED;
        }
        PAT_LAUNCH_CHECK();
        return d.bucketed ? bucket_reduce(d, st) : PAT_OK;
    }
    switch (d.KS * 4 + d.NT) {
    case 1 * 4 + 1: PAT_KM(1, 8, 1); break;
    case 2 * 4 + 1: PAT_KM(2, 8, 1); break;
    case 4 * 4 + 1: PAT_KM(4, 4, 1); break;
    case 8 * 4 + 1: PAT_KM(8, 2, 1); break;
    case 1 * 4 + 3: PAT_KM_BS(1, 4, 3, false, 768); break;  // 3 wavefronts per SIMD
    case 2 * 4 + 3: PAT_KM_BS(2, 1, 3, false, 768); break;  // 3 wavefronts per SIMD
    case 4 * 4 + 3: PAT_KM_BS(4, 2, 3, false, 768); break;  // 3 wavefronts per SIMD
    case 8 * 4 + 3: PAT_KM(8, 2, 3); break;
    default: pat::set_error("kmeans: D too large"); return PAT_ERR_UNSUPPORTED;
    }
#undef PAT_KM
#undef PAT_KM_GC
#undef PAT_KM_BS
    PAT_LAUNCH_CHECK();
    return d.bucketed ? bucket_reduce(d, st) : PAT_OK;
}

int kmeans_zero(unsigned long long *p, int64_t n, hipStream_t st)
{
    if (n <= 0) return PAT_OK;
    hipLaunchKernelGGL(km_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

}  // namespace pat
