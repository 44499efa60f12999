// Generalized-reduction driver over RegularRuntime with the application's
// own functor tables compiled for the device (gpu_table.hip): the calls of
// kmeans/km.cu:20-87 and kmeans/moldyn.cu:440-470 (input + Offset array,
// RegularRuntime, RegularInit, set_map_idx, set_reduce_idx, RegularStart,
// RegularGetOutput) for the four map/reduce pairs of regular.h, each checked
// against the same functors run on the host into a CO object:
//   0 velocity statistics (41 ANY_KEY buckets on the device, one on the host),
//   1 kmeans nearest centre at D = 3, K = 40,
//   2 miniMD pair energy,
//   3 a 5,003-key histogram, its units placed by an Offset array with gaps,
//     once with the LDS objects forced to merge after 4 keys.
// Integer values (counts) must match exactly; FP32 sums to 1e-5 (kmeans:
// 1e-4 of the sum, or of count x 16 for coordinate sums that may cancel,
// since the host adds thousands of values one by one) and FP64 sums to 1e-10
// relative (the device folds in another order).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#define CPU_KERNEL
#include "include/runtime.h"
#include "regular.h"
#include "regular/regular_runtime.h"

namespace {

// Host run of map m / reduce m over the units.
CO host_run(int m, void *input, const std::vector<Offset> &off, void *param)
{
    CO obj;
    for (const Offset &o : off)
        FFCPU::map_function_table[m](&obj, input, o.offset, param, 0, FFCPU::reduce_function_table[m]);
    return obj;
}

// Device run through RegularRuntime: key bytes -> value bytes.
std::map<std::string, std::string> device_run(int m, void *input, size_t bytes, std::vector<Offset> &off, void *param,
                                              int param_size, int lds_keys = 0)
{
    RegularRuntime rr(1, input, bytes, off.data(), off.size(), param, param_size);
    rr.RegularInit();
    if (lds_keys) rr.set_object_capacity(8192, 32, lds_keys);
    rr.set_map_idx(m);
    rr.set_reduce_idx(m);
    rr.RegularStart();
    struct output o = rr.RegularGetOutput();
    std::map<std::string, std::string> r;
    for (unsigned i = 0; i < o.key_num; i++)
        r[std::string(o.output_keys + o.key_index[i], o.key_size)] =
            std::string(o.output_vals + o.val_index[i], o.value_size);
    return r;
}

bool close(double a, double b, double tol) { return std::fabs(a - b) <= tol * std::fmax(std::fabs(b), 1e-300); }

}  // namespace

int main(int argc, char *argv[])
{
    RuntimeInit(argc, argv);
    int bad = 0;

    // 0: velocity statistics over 100,000 Maxwell velocities
    {
        const long n = 100000;
        std::vector<float> v32(3 * n);
        PAT_API_CALL(pat_gen_velocities(n, 1.44, 11, v32.data()));
        std::vector<double> v(v32.begin(), v32.end());
        std::vector<Offset> off(n);
        for (long i = 0; i < n; i++) off[i] = Offset{(size_t)i * 24, 24};
        double s = 0;
        for (long i = 0; i < n; i++) s += std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        float vaverh = (float)(s / n);
        auto dev = device_run(0, v.data(), v.size() * 8, off, &vaverh, sizeof(float));
        CO host = host_run(0, v.data(), off, &vaverh);
        double ds = 0, dc = 0;
        for (auto &e : dev) {
            ds += ((const float *)e.second.data())[0];
            dc += ((const float *)e.second.data())[1];
        }
        const float *h = (const float *)host.value(0);
        const bool ok = dev.size() == 41 && dc == h[1] && close(ds, h[0], 1e-5);
        printf("regular: map 0 velocity statistics: %zu keys, sum |v| %.6e (host %.6e), above %.0f (host %.0f) %s\n",
               dev.size(), ds, (double)h[0], dc, (double)h[1], ok ? "match" : "DIFFER");
        bad += !ok;
    }
    // 1: kmeans at D = 3, K = 40 over 200,000 mixture points
    {
        const long n = 200000;
        std::vector<float> x(3 * n);
        PAT_API_CALL(pat_gen_gmm(n, KMEANS_DIM, KMEANS_K, 1.0, -10.0, 10.0, 12, x.data()));
        std::vector<float> c(x.begin(), x.begin() + KMEANS_K * KMEANS_DIM);
        std::vector<Offset> off(n);
        for (long i = 0; i < n; i++) off[i] = Offset{(size_t)i * 12, 12};
        auto dev = device_run(1, x.data(), x.size() * 4, off, c.data(), (int)(c.size() * 4));
        CO host = host_run(1, x.data(), off, c.data());
        bool ok = dev.size() == host.entries().size();
        for (auto &e : host.entries()) {
            auto it = dev.find(e.first);
            if (it == dev.end()) {
                ok = false;
                continue;
            }
            const float *a = (const float *)it->second.data(), *b = (const float *)e.second.data();
            // counts exact; sums of ~7,500 FP32 values, which the host adds
            // one by one, within 1e-4 of the count times the largest
            // |coordinate| (16) for the coordinate sums (they may nearly
            // cancel) and of the distance sum
            ok = ok && a[3] == b[3];
            for (int d = 0; d < 3; d++) ok = ok && std::fabs(a[d] - b[d]) <= 1e-4 * 16.0 * b[3];
            ok = ok && close(a[4], b[4], 1e-4);
        }
        printf("regular: map 1 kmeans D=3 K=40: %zu keys, counts and sums %s\n", dev.size(), ok ? "match" : "DIFFER");
        if (!ok)
            for (auto &e : host.entries()) {
                auto it = dev.find(e.first);
                if (it == dev.end()) continue;
                const float *a = (const float *)it->second.data(), *b = (const float *)e.second.data();
                printf("  key %u: device %g %g %g %g %g host %g %g %g %g %g\n", *(const unsigned *)e.first.data(),
                       a[0], a[1], a[2], a[3], a[4], b[0], b[1], b[2], b[3], b[4]);
            }
        bad += !ok;
    }
    // 2: miniMD pair energy over an 8^3-cell FCC lattice's pairs within 2.8
    {
        std::vector<float> p(3 * 2048);
        double box[3];
        PAT_API_CALL(pat_gen_fcc(8, 8, 8, 0.8442, 0.05, 13, p.data(), box));
        int64_t ne = 0;
        PAT_API_CALL(pat_gen_pairs(2048, p.data(), box[0], 2.8, 0, 0, nullptr, &ne));
        std::vector<int> e(2 * ne);
        PAT_API_CALL(pat_gen_pairs(2048, p.data(), box[0], 2.8, 0, 0, e.data(), &ne));
        std::vector<double> prm(1 + p.size());
        prm[0] = 6.25;
        for (size_t i = 0; i < p.size(); i++) prm[1 + i] = p[i];
        std::vector<Offset> off(ne);
        for (long i = 0; i < ne; i++) off[i] = Offset{(size_t)i * 8, 8};
        auto dev = device_run(2, e.data(), e.size() * 4, off, prm.data(), (int)(prm.size() * 8));
        CO host = host_run(2, e.data(), off, prm.data());
        double ds = 0;
        for (auto &kv : dev) ds += *(const double *)kv.second.data();
        const double hs = *(const double *)host.value(0);
        const bool ok = close(ds, hs, 1e-10);
        printf("regular: map 2 pair energy: %zu keys, sum %.12e (host %.12e) %s\n", dev.size(), ds, hs,
               ok ? "match" : "DIFFER");
        bad += !ok;
    }
    // 3: histogram of 1,000,000 values placed 12 bytes apart (value at +4)
    {
        const long n = 1000000;
        std::vector<unsigned int> buf(3 * n);
        unsigned long long s = 14;
        for (long i = 0; i < n; i++) {
            s = s * 6364136223846793005ULL + 1442695040888963407ULL;
            buf[3 * i] = 0xdeadbeef;
            buf[3 * i + 1] = (unsigned int)(s >> 33);
            buf[3 * i + 2] = 0xfeedface;
        }
        std::vector<Offset> off(n);
        for (long i = 0; i < n; i++) off[i] = Offset{(size_t)i * 12 + 4, 4};
        CO host = host_run(3, buf.data(), off, nullptr);
        for (int lds : {0, 4}) {
            auto dev = device_run(3, buf.data(), buf.size() * 4, off, nullptr, 0, lds);
            bool ok = dev.size() == host.entries().size() && dev.size() == HIST_KEYS;
            long long total = 0;
            for (auto &e : host.entries()) {
                auto it = dev.find(e.first);
                if (it == dev.end()) {
                    ok = false;
                    continue;
                }
                const FFCPU::HistValue *a = (const FFCPU::HistValue *)it->second.data(),
                                       *b = (const FFCPU::HistValue *)e.second.data();
                ok = ok && a->count == b->count && close(a->sum, b->sum, 1e-12);
                total += a->count;
            }
            ok = ok && total == n;
            printf("regular: map 3 histogram (Offset array with gaps%s): %zu keys, counts %s\n",
                   lds ? ", LDS objects merged every 4 keys" : "", dev.size(), ok ? "match" : "DIFFER");
            bad += !ok;
        }
    }
    {
        // map 4: 6-byte keys, 3-byte values added bytewise modulo 256, so
        // the device object must equal the host one byte for byte
        const long n = 200000;
        std::vector<unsigned int> v(n);
        for (long i = 0; i < n; i++) v[i] = (unsigned int)(i * 2654435761u);
        std::vector<Offset> off(n);
        for (long i = 0; i < n; i++) off[i] = Offset{(size_t)i * 4, 4};
        CO host = host_run(4, v.data(), off, nullptr);
        for (int lds : {0, 3}) {
            auto dev = device_run(4, v.data(), v.size() * 4, off, nullptr, 0, lds);
            const bool ok = dev == host.entries() && dev.size() == 231;
            printf("regular: map 4 odd-sized records (6-byte keys, 3-byte values%s): %zu keys, bytes %s\n",
                   lds ? ", LDS objects merged every 3 keys" : "", dev.size(), ok ? "match" : "DIFFER");
            bad += !ok;
        }
    }
    RuntimeFinalize();
    return bad ? 1 : 0;
}
