// Device forms of the seeded workload generators (gen.cpp): the same
// counter-based SplitMix64 values computed where the data will live, so a
// job's input never crosses PCIe.  The heat3D grid and the FCC lattice are
// the host generators' values bit for bit (integer hashing and the same
// FP64 operations); the Gaussian-mixture points use the device's log, cos
// and sqrt, which may differ from the host library's in the last FP64 bit.
#include "pat_common.hpp"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rng(uint64_t seed, uint64_t s, uint64_t i)
{
    return mix64(mix64(seed ^ (s * 0xd1b54a32d192ed03ULL)) + i);
}

__device__ __forceinline__ float unif_f(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }
__device__ __forceinline__ double unif_d(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ double gauss(uint64_t seed, uint64_t s, uint64_t i)
{
    const double u1 = 1.0 - unif_d(rng(seed, s, 2 * i));
    const double u2 = unif_d(rng(seed, s, 2 * i + 1));
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ void gen_heat3d_k(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, float *__restrict__ grid)
{
    const int64_t sx = nx + 2, sy = ny + 2, total = sx * sy * (nz + 2);
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = g % sx, y = (g / sx) % sy, z = g / (sx * sy);
        float v = 0.0f;
        if (x >= 1 && x <= nx && y >= 1 && y <= ny && z >= 1 && z <= nz)
            v = unif_f(rng(seed, 0, (uint64_t)((x - 1) + nx * ((y - 1) + ny * (z - 1)))));
        grid[g] = v;
    }
}

__global__ void gen_fcc_k(int64_t ncx, int64_t ncy, int64_t natoms, double a, double L0, double L1, double L2,
                          double jitter, uint64_t seed, float4 *__restrict__ pos)
{
    for (int64_t atom = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; atom < natoms;
         atom += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = atom / ncx, m = atom % ncx;
        const int64_t k = r / (2 * ncy), j = r % (2 * ncy);
        const int64_t ijk[3] = {2 * m + ((j + k) & 1), j, k};
        const double L[3] = {L0, L1, L2};
        float out[3];
        for (int c = 0; c < 3; c++) {
            double v = 0.5 * a * (double)ijk[c];
            if (jitter > 0) v += jitter * (2.0 * unif_d(rng(seed, 10 + c, (uint64_t)atom)) - 1.0);
            float fv = (float)v;
            const float fL = (float)L[c];
            if (fv < 0.0f) fv += fL;
            else if (fv >= fL) fv -= fL;
            out[c] = fv;
        }
        pos[atom] = make_float4(out[0], out[1], out[2], 0.0f);
    }
}

__global__ void gen_gmm_k(int64_t n, int d, int ncomp, double sigma, double lo, double hi, uint64_t seed,
                          float *__restrict__ x)
{
    const int64_t total = n * d;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = g / d;
        const int k = (int)(g % d);
        const uint64_t r = rng(seed, 41, (uint64_t)p);
        const int comp = (int)(((r >> 32) * (uint64_t)ncomp) >> 32);
        const double c = lo + (hi - lo) * unif_d(rng(seed, 40, (uint64_t)comp * d + k));
        x[g] = (float)(c + sigma * gauss(seed, 42, (uint64_t)g));
    }
}

unsigned grid_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

}  // namespace

extern "C" {

int pat_gen_heat3d_dev(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, float *grid, void *stream)
{
    PAT_REQUIRE(nx > 0 && ny > 0 && nz > 0 && grid, PAT_ERR_INVALID, "pat_gen_heat3d_dev: bad arguments");
    const int64_t total = (nx + 2) * (ny + 2) * (nz + 2);
    hipLaunchKernelGGL(gen_heat3d_k, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, seed, grid);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int pat_gen_fcc_dev(int64_t ncx, int64_t ncy, int64_t ncz, double rho, double jitter, uint64_t seed, float *pos4,
                    double *box_out, void *stream)
{
    PAT_REQUIRE(ncx > 0 && ncy > 0 && ncz > 0 && rho > 0 && pos4, PAT_ERR_INVALID, "pat_gen_fcc_dev: bad arguments");
    const double a = std::pow(4.0 / rho, 1.0 / 3.0);
    const double L[3] = {ncx * a, ncy * a, ncz * a};
    if (box_out)
        for (int c = 0; c < 3; c++) box_out[c] = L[c];
    const int64_t natoms = 4 * ncx * ncy * ncz;
    hipLaunchKernelGGL(gen_fcc_k, dim3(grid_for(natoms)), dim3(256), 0, (hipStream_t)stream, ncx, ncy, natoms, a, L[0],
                       L[1], L[2], jitter, seed, (float4 *)pos4);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

int pat_gen_gmm_dev(int64_t n, int d, int ncomp, double sigma, double lo, double hi, uint64_t seed, float *x,
                    void *stream)
{
    PAT_REQUIRE(n > 0 && d > 0 && ncomp > 0 && x, PAT_ERR_INVALID, "pat_gen_gmm_dev: bad arguments");
    hipLaunchKernelGGL(gen_gmm_k, dim3(grid_for(n * d)), dim3(256), 0, (hipStream_t)stream, n, d, ncomp, sigma, lo, hi,
                       seed, x);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

}  // extern "C"

namespace {

constexpr int kSumBlocks = 256;

// Per-block FP64 sums of p[0..n) in a fixed order (grid-stride per thread,
// then a fixed tree over the block), so the total the host forms from the
// block sums is the same on every run.
__global__ __launch_bounds__(256) void sum_f32_k(const float *__restrict__ p, int64_t n, double *__restrict__ part)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += (double)p[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// The same over the x, y, z of n float4 (w left out), in a fixed order.
__global__ __launch_bounds__(256) void sum_xyz_k(const float4 *__restrict__ p, int64_t n, double *__restrict__ part)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float4 v = p[i];
        acc += (double)v.x;
        acc += (double)v.y;
        acc += (double)v.z;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void set_w_k(float4 *__restrict__ p, int64_t n, float w)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i].w = w;
}

int sum_parts(const double *part, double *out, hipStream_t st)
{
    double h[kSumBlocks];
    PAT_HIP(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, st));
    PAT_HIP(hipStreamSynchronize(st));
    double t = 0.0;
    for (double v : h) t += v;
    *out = t;
    return PAT_OK;
}

}  // namespace

namespace pat {

int sum_blocks() { return kSumBlocks; }

// Loads this file's code object now rather than at a first launch.
int gen_preload()
{
    hipFuncAttributes a;
    PAT_HIP(hipFuncGetAttributes(&a, (const void *)sum_f32_k));
    return PAT_OK;
}

// Sum of n floats at p (device) into *out, using part (kSumBlocks doubles of
// device memory) on stream st; waits for the stream.
int sum_f32_dev(const float *p, int64_t n, double *part, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(sum_f32_k, dim3(kSumBlocks), dim3(256), 0, st, p, n, part);
    PAT_LAUNCH_CHECK();
    return sum_parts(part, out, st);
}

// Sum of the x, y, z of n float4 at p (device), as sum_f32_dev.
int sum_xyz_dev(const float *p, int64_t n, double *part, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(sum_xyz_k, dim3(kSumBlocks), dim3(256), 0, st, (const float4 *)p, n, part);
    PAT_LAUNCH_CHECK();
    return sum_parts(part, out, st);
}

// w = value for n float4 at p (device), queued on st.
int set_w_dev(float *p, int64_t n, float w, hipStream_t st)
{
    hipLaunchKernelGGL(set_w_k, dim3(grid_for(n)), dim3(256), 0, st, (float4 *)p, n, w);
    PAT_LAUNCH_CHECK();
    return PAT_OK;
}

}  // namespace pat
