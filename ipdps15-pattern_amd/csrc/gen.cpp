// Seeded workload generators: the app drivers' data set-up (heat3D grid
// init, moldyn/miniMD FCC lattice and velocities, neighbour pairs, kmeans
// Gaussian-mixture points).  Every value is a pure function of (seed, index)
// through a counter-based SplitMix64 hash, so the output does not depend on
// the host thread count.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pat_common.hpp"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Independent 64-bit value for element i of stream s under seed.
inline uint64_t rng(uint64_t seed, uint64_t s, uint64_t i)
{
    return mix64(mix64(seed ^ (s * 0xd1b54a32d192ed03ULL)) + i);
}

// Uniform float in [0, 1) with 24 random bits (exact in FP32).
inline float unif_f(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }
// Uniform double in [0, 1) with 53 random bits.
inline double unif_d(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Standard normal from two hashed uniforms (Box-Muller, cosine branch).
inline double gauss(uint64_t seed, uint64_t s, uint64_t i)
{
    double u1 = 1.0 - unif_d(rng(seed, s, 2 * i));  // (0, 1]
    double u2 = unif_d(rng(seed, s, 2 * i + 1));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

extern "C" {

int pat_gen_heat3d(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, float *grid)
{
    PAT_REQUIRE(nx > 0 && ny > 0 && nz > 0 && grid, PAT_ERR_INVALID, "pat_gen_heat3d: bad arguments");
    const int64_t sx = nx + 2, sy = ny + 2, sz = nz + 2;
    pat::parallel_for(sz, [&](int64_t z0, int64_t z1) {
        for (int64_t z = z0; z < z1; z++)
            for (int64_t y = 0; y < sy; y++) {
                float *row = grid + (z * sy + y) * sx;
                if (z == 0 || z == sz - 1 || y == 0 || y == sy - 1) {
                    memset(row, 0, sizeof(float) * sx);
                    continue;
                }
                row[0] = 0.0f;
                row[sx - 1] = 0.0f;
                for (int64_t x = 1; x <= nx; x++) {
                    uint64_t lin = (x - 1) + nx * ((y - 1) + ny * (z - 1));
                    row[x] = unif_f(rng(seed, 0, lin));
                }
            }
    });
    return PAT_OK;
}

int pat_gen_fcc(int64_t ncx, int64_t ncy, int64_t ncz, double rho, double jitter, uint64_t seed, float *pos,
                double *box_out)
{
    PAT_REQUIRE(ncx > 0 && ncy > 0 && ncz > 0 && rho > 0 && pos, PAT_ERR_INVALID, "pat_gen_fcc: bad arguments");
    const double a = std::pow(4.0 / rho, 1.0 / 3.0);
    const double L[3] = {ncx * a, ncy * a, ncz * a};
    if (box_out) {
        box_out[0] = L[0];
        box_out[1] = L[1];
        box_out[2] = L[2];
    }
    // Lattice sites (i, j, k) on the half-spacing grid with i+j+k even, visited
    // z-outer, x-inner as miniMD create_atoms does; each (k, j) row holds
    // exactly ncx sites, so the atom index has a closed form.
    const int64_t rows = (2 * ncz) * (2 * ncy);
    pat::parallel_for(rows, [&](int64_t r0, int64_t r1) {
        for (int64_t r = r0; r < r1; r++) {
            int64_t k = r / (2 * ncy), j = r % (2 * ncy);
            for (int64_t m = 0; m < ncx; m++) {
                int64_t i = 2 * m + ((j + k) & 1);
                int64_t atom = r * ncx + m;
                const int64_t ijk[3] = {i, j, k};
                for (int c = 0; c < 3; c++) {
                    double v = 0.5 * a * (double)ijk[c];
                    if (jitter > 0) v += jitter * (2.0 * unif_d(rng(seed, 10 + c, (uint64_t)atom)) - 1.0);
                    float fv = (float)v;
                    float fL = (float)L[c];
                    if (fv < 0.0f) fv += fL;
                    else if (fv >= fL) fv -= fL;
                    pos[3 * atom + c] = fv;
                }
            }
        }
    });
    return PAT_OK;
}

int pat_gen_velocities(int64_t n, double t, uint64_t seed, float *vel)
{
    PAT_REQUIRE(n > 1 && vel && t >= 0, PAT_ERR_INVALID, "pat_gen_velocities: bad arguments");
    std::vector<double> v(3 * n);
    pat::parallel_for(3 * n, [&](int64_t b, int64_t e) {
        for (int64_t m = b; m < e; m++) v[m] = gauss(seed, 20, (uint64_t)m);
    });
    double tot[3] = {0, 0, 0};
    for (int64_t p = 0; p < n; p++)
        for (int c = 0; c < 3; c++) tot[c] += v[3 * p + c];
    for (int c = 0; c < 3; c++) tot[c] /= (double)n;
    double ke = 0;
    for (int64_t p = 0; p < n; p++)
        for (int c = 0; c < 3; c++) {
            v[3 * p + c] -= tot[c];
            ke += v[3 * p + c] * v[3 * p + c];
        }
    double tcur = ke / (3.0 * (double)(n - 1));
    double f = tcur > 0 ? std::sqrt(t / tcur) : 0.0;
    for (int64_t m = 0; m < 3 * n; m++) vel[m] = (float)(v[m] * f);
    return PAT_OK;
}

int pat_gen_pairs(int64_t n, const float *pos, double box, double cut, int shuffle, uint64_t shuffle_seed,
                  int32_t *edges, int64_t *num_edges)
{
    PAT_REQUIRE(n > 0 && n < (1LL << 31) && pos && num_edges && box > 0 && cut > 0, PAT_ERR_INVALID,
                "pat_gen_pairs: bad arguments");
    const float L = (float)box, hL = 0.5f * L, cut2 = (float)(cut * cut);
    // Minimum-image distance test, the same float expression as the
    // irregular engine's pair functor.
    auto within = [&](int64_t i, int64_t j) {
        float dx = pos[3 * i] - pos[3 * j];
        float dy = pos[3 * i + 1] - pos[3 * j + 1];
        float dz = pos[3 * i + 2] - pos[3 * j + 2];
        if (dx < -hL) dx += L;
        if (dy < -hL) dy += L;
        if (dz < -hL) dz += L;
        if (dx > hL) dx -= L;
        if (dy > hL) dy -= L;
        if (dz > hL) dz -= L;
        float r2 = dx * dx + dy * dy;
        r2 = r2 + dz * dz;
        return r2 < cut2;
    };
    int nb = (int)std::floor(box / (cut * (1.0 + 1e-6)));
    std::vector<std::vector<int32_t>> lists(n);
    if (nb < 3) {
        pat::parallel_for(n, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; i++)
                for (int64_t j = i + 1; j < n; j++)
                    if (within(i, j)) lists[i].push_back((int32_t)j);
        });
    } else {
        const int64_t ncell = (int64_t)nb * nb * nb;
        std::vector<int32_t> cell(n), head(ncell + 1, 0), order(n);
        auto bin = [&](float x) {
            int c = (int)std::floor((double)x * nb / box);
            return c < 0 ? 0 : (c >= nb ? nb - 1 : c);
        };
        for (int64_t i = 0; i < n; i++) {
            int cx = bin(pos[3 * i]), cy = bin(pos[3 * i + 1]), cz = bin(pos[3 * i + 2]);
            cell[i] = (int32_t)((int64_t)cx + nb * ((int64_t)cy + nb * cz));
            head[cell[i] + 1]++;
        }
        for (int64_t c = 0; c < ncell; c++) head[c + 1] += head[c];
        {
            std::vector<int32_t> cur(head.begin(), head.end() - 1);
            for (int64_t i = 0; i < n; i++) order[cur[cell[i]]++] = (int32_t)i;  // ascending i per cell
        }
        pat::parallel_for(n, [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; i++) {
                int c = cell[i];
                int cx = c % nb, cy = (c / nb) % nb, cz = c / (nb * nb);
                auto &li = lists[i];
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            int qx = (cx + dx + nb) % nb, qy = (cy + dy + nb) % nb, qz = (cz + dz + nb) % nb;
                            int64_t q = (int64_t)qx + nb * ((int64_t)qy + nb * qz);
                            for (int32_t s = head[q]; s < head[q + 1]; s++) {
                                int32_t j = order[s];
                                if (j > i && within(i, j)) li.push_back(j);
                            }
                        }
                std::sort(li.begin(), li.end());
            }
        });
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) total += (int64_t)lists[i].size();
    if (!edges) {
        *num_edges = total;
        return PAT_OK;
    }
    PAT_REQUIRE(*num_edges >= total, PAT_ERR_INVALID, "pat_gen_pairs: capacity %lld < %lld pairs",
                (long long)*num_edges, (long long)total);
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++)
        for (int32_t j : lists[i]) {
            edges[2 * w] = (int32_t)i;
            edges[2 * w + 1] = j;
            w++;
        }
    if (shuffle) {
        // Fisher-Yates with hashed draws: position m swaps with a draw in [0, m].
        for (int64_t m = total - 1; m > 0; m--) {
            uint64_t r = rng(shuffle_seed, 30, (uint64_t)m);
            int64_t q = (int64_t)((unsigned __int128)r * (uint64_t)(m + 1) >> 64);
            std::swap(edges[2 * m], edges[2 * q]);
            std::swap(edges[2 * m + 1], edges[2 * q + 1]);
        }
    }
    *num_edges = total;
    return PAT_OK;
}

int pat_gen_gmm(int64_t n, int d, int ncomp, double sigma, double lo, double hi, uint64_t seed, float *x)
{
    PAT_REQUIRE(n > 0 && d > 0 && ncomp > 0 && x, PAT_ERR_INVALID, "pat_gen_gmm: bad arguments");
    std::vector<double> centre((size_t)ncomp * d);
    for (int64_t m = 0; m < (int64_t)ncomp * d; m++) centre[m] = lo + (hi - lo) * unif_d(rng(seed, 40, (uint64_t)m));
    pat::parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t p = b; p < e; p++) {
            uint64_t r = rng(seed, 41, (uint64_t)p);
            int comp = (int)(((r >> 32) * (uint64_t)ncomp) >> 32);
            const double *c = centre.data() + (size_t)comp * d;
            for (int k = 0; k < d; k++) x[p * d + k] = (float)(c[k] + sigma * gauss(seed, 42, (uint64_t)(p * d + k)));
        }
    });
    return PAT_OK;
}

}  // extern "C"
