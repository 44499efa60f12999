// Generalized reduction with additive functors: the reference's map/reduce
// pairs whose reduce adds values (kmeans/regular.h:30-73 map/reduce, the
// velocity statistics moldyn runs through RegularRuntime, kmeans/moldyn.cu:
// 440-470; miniMD/regular.h:139-181 map2/reduce2, the pair energy), run by
// the reference's regular compute_gpu kernel (kmeans/regular/gpu_kernel.cu:
// 39-183) into an LDS hash reduction object (ros.cu:15-170).
//
// Reduction object.  The reference's emits go to ANY_KEY buckets (threadIdx
// % 41, kmeans/regular/gpu_kernel_tools.cu:68-71) that the caller adds up,
// and its FP32 values are added in lock-acquisition order.  Here a functor
// emits (key, value[V]) with a key in [0, K) -- K = 1 for both functors, the
// ANY_KEY buckets folded -- and the object is a dense K x V array of FP64
// sums plus a count of emits per key, built without atomics so that every
// run gives the same bits: each wavefront takes 64 units at a time, sums
// each key's values across its lanes with a fixed butterfly, and lane 0 adds
// the result to the wavefront's own LDS copy; a workgroup adds its four
// copies in order into its slot of a per-workgroup array, and a final
// kernel adds the workgroups' slots in order.
#include <algorithm>
#include <cmath>

#include "pat_common.hpp"
#include "pat_device.hpp"

namespace {

using pat::kWave;

// kmeans/regular.h:30-73 (map/reduce): per velocity (double[3]) value[0] =
// (float)sqrt(vx^2 + vy^2 + vz^2) computed in FP64, value[1] = value[0] >
// vaverh (FP32 compare); reduce adds both.
struct VStats {
    static constexpr int kK = 1, kV = 2;
    float vaverh;
    __device__ bool map(const char *units, int64_t i, const double *, int &key, double (&v)[kV]) const
    {
        const double *u = (const double *)units + 3 * i;
        const double d1 = u[0], d2 = u[1], d3 = u[2];
        const float value0 = (float)sqrt(d1 * d1 + d2 * d2 + d3 * d3);
        key = 0;
        v[0] = (double)value0;
        v[1] = value0 > vaverh ? 1.0 : 0.0;
        return true;
    }
};

// miniMD/regular.h:139-181 (map2/reduce2): per edge (int[2] i, j) with the
// positions x (FP64, 3 per node) from the parameter: rsq = dx^2 + dy^2 + dz^2
// (no minimum image); within cutsq emit phi = sr6 * (sr6 - 1), sr2 = 1/rsq,
// sr6 = sr2^3; reduce adds.
struct LJEnergy {
    static constexpr int kK = 1, kV = 1;
    double cutsq;
    __device__ bool map(const char *units, int64_t k, const double *x, int &key, double (&v)[kV]) const
    {
        const int *e = (const int *)units + 2 * k;
        const int64_t i = e[0], j = e[1];
        const double delx = x[3 * i] - x[3 * j];
        const double dely = x[3 * i + 1] - x[3 * j + 1];
        const double delz = x[3 * i + 2] - x[3 * j + 2];
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (!(rsq < cutsq)) return false;
        const double sr2 = 1.0 / rsq;
        const double sr6 = sr2 * sr2 * sr2;
        key = 0;
        v[0] = sr6 * (sr6 - 1.0);
        return true;
    }
};

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) x = x + __shfl_xor(x, o, kWave);
    return x;
}

// Workgroup g reduces units [g*per, min((g+1)*per, n)) into part[g] (K*V
// sums) and cnt[g] (K counts).
template <class F>
__global__ __launch_bounds__(256) void gr_map_reduce(const char *__restrict__ units, int64_t n, int64_t per, F f,
                                                     const double *__restrict__ aux, double *__restrict__ part,
                                                     unsigned long long *__restrict__ cnt)
{
    constexpr int K = F::kK, V = F::kV;
    __shared__ double acc[4][K * V];
    __shared__ unsigned long long accn[4][K];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) {
        for (int t = 0; t < K * V; t++) acc[w][t] = 0.0;
        for (int t = 0; t < K; t++) accn[w][t] = 0;
    }
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min(b0 + per, n);
    for (int64_t c0 = b0 + (int64_t)w * kWave; c0 < b1; c0 += 4 * kWave) {
        const int64_t i = c0 + lane;
        int key = -1;
        double v[V];
#pragma unroll
        for (int t = 0; t < V; t++) v[t] = 0.0;
        if (i < b1 && !f.map(units, i, aux, key, v)) key = -1;
        for (int k = 0; k < K; k++) {
            const unsigned long long m = __ballot(key == k);
            if (!m) continue;
#pragma unroll
            for (int t = 0; t < V; t++) {
                const double s = wave_sum(key == k ? v[t] : 0.0);
                if (lane == 0) acc[w][k * V + t] += s;
            }
            if (lane == 0) accn[w][k] += (unsigned long long)__popcll(m);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < K * V; t += blockDim.x)
        part[(int64_t)blockIdx.x * K * V + t] = ((acc[0][t] + acc[1][t]) + acc[2][t]) + acc[3][t];
    for (int t = threadIdx.x; t < K; t += blockDim.x)
        cnt[(int64_t)blockIdx.x * K + t] = accn[0][t] + accn[1][t] + accn[2][t] + accn[3][t];
}

// Sums the workgroups' slots in order: out[t] = sum_g part[g][t].
__global__ void gr_final(const double *__restrict__ part, const unsigned long long *__restrict__ cnt, int nblk,
                         int kv, int k, double *__restrict__ out, unsigned long long *__restrict__ outn)
{
    const int t = threadIdx.x;
    if (t < kv) {
        double s = 0.0;
        for (int g = 0; g < nblk; g++) s += part[(int64_t)g * kv + t];
        out[t] = s;
    }
    if (t < k) {
        unsigned long long c = 0;
        for (int g = 0; g < nblk; g++) c += cnt[(int64_t)g * k + t];
        outn[t] = c;
    }
}

}  // namespace

namespace pat {

// Grid: enough workgroups to fill the chip, each a contiguous run of units
// (a function of n only, so the summation order is fixed).
static void gr_grid(int64_t n, int *nblk, int64_t *per)
{
    int64_t g = (n + 2047) / 2048;
    g = std::max<int64_t>(1, std::min<int64_t>(g, 4096));
    *per = std::max<int64_t>(1, (n + g - 1) / g);
    *nblk = (int)std::max<int64_t>(1, (n + *per - 1) / *per);
}

// functor: 0 = velocity statistics, 2 = pair energy.  part/cnt: scratch of
// 4096*K*V doubles and 4096*K counts; out/outn: K*V sums and K counts.
int gr_run(int functor, const void *units, int64_t n, const void *param, const double *aux, double *part,
           unsigned long long *cnt, double *out, unsigned long long *outn, hipStream_t st)
{
    int nblk;
    int64_t per;
    gr_grid(n, &nblk, &per);
    int kv, k;
    if (functor == 0) {
        VStats f;
        f.vaverh = *(const float *)param;
        kv = VStats::kK * VStats::kV;
        k = VStats::kK;
        hipLaunchKernelGGL(gr_map_reduce<VStats>, dim3(nblk), dim3(256), 0, st, (const char *)units, n, per, f, aux,
                           part, cnt);
    } else if (functor == 2) {
        LJEnergy f;
        f.cutsq = *(const double *)param;
        kv = LJEnergy::kK * LJEnergy::kV;
        k = LJEnergy::kK;
        hipLaunchKernelGGL(gr_map_reduce<LJEnergy>, dim3(nblk), dim3(256), 0, st, (const char *)units, n, per, f,
                           aux, part, cnt);
    } else {
        pat::set_error("gr_run: functor %d has no engine", functor);
        return PAT_ERR_UNSUPPORTED;
    }
    PAT_LAUNCH_CHECK();
    hipLaunchKernelGGL(gr_final, dim3(1), dim3(64), 0, st, (const double *)part, (const unsigned long long *)cnt, nblk,
                       kv, k, out, outn);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

}  // namespace pat
