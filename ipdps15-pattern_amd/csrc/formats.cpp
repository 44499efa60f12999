// Readers for the reference's input files, so reference-shaped datasets can
// be replayed: moldyn's parameter file (read from stdin by
// kmeans/moldyn.cu:193-201), its graph (.mesh, ReadInputGraph,
// kmeans/misc.c:76-145) and coordinates (.xyz, ReadCoord, misc.c:226-247),
// and miniMD's lj.in (miniMD/input.cpp:39-72).  Host code only.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pat_common.hpp"

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *path) : f(path ? fopen(path, "r") : nullptr) {}
    ~File()
    {
        if (f) fclose(f);
    }
};

// One line without its newline; false at end of file with nothing read.
bool get_line(FILE *f, std::string &s)
{
    s.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        any = true;
        if (c == '\n') break;
        s.push_back((char)c);
    }
    return any;
}

}  // namespace

extern "C" {

int pat_read_mold(const char *path, pat_mold *out)
{
    PAT_REQUIRE(path && out, PAT_ERR_INVALID, "pat_read_mold: null argument");
    File fp(path);
    PAT_REQUIRE(fp.f, PAT_ERR_INVALID, "pat_read_mold: cannot open %s", path);
    pat_mold m;
    memset(&m, 0, sizeof(m));
    // whitespace-separated fields in moldyn's scanf order: nB nTS nU cutoff
    // mesh xyz percent start_time update_freq
    const int got = fscanf(fp.f, "%d %d %d %lf %511s %511s %lf %d %d", &m.box_cells, &m.steps, &m.neigh_update,
                           &m.cutoff, m.mesh, m.xyz, &m.percent, &m.start_time, &m.update_freq);
    PAT_REQUIRE(got == 9, PAT_ERR_INVALID, "pat_read_mold: %s holds %d of the 9 fields", path, got < 0 ? 0 : got);
    PAT_REQUIRE(m.box_cells > 0 && m.steps >= 0 && m.cutoff > 0, PAT_ERR_INVALID,
                "pat_read_mold: %s: box cells %d, steps %d, cutoff %g out of range", path, m.box_cells, m.steps,
                m.cutoff);
    *out = m;
    return PAT_OK;
}

int pat_read_ljin(const char *path, pat_ljin *out)
{
    PAT_REQUIRE(path && out, PAT_ERR_INVALID, "pat_read_ljin: null argument");
    File fp(path);
    PAT_REQUIRE(fp.f, PAT_ERR_INVALID, "pat_read_ljin: cannot open %s", path);
    pat_ljin in;
    memset(&in, 0, sizeof(in));
    std::string line;
    // two header lines, then one value group per line (input.cpp's order)
    for (int i = 0; i < 2; i++) PAT_REQUIRE(get_line(fp.f, line), PAT_ERR_INVALID, "pat_read_ljin: %s is short", path);
    struct Field {
        const char *fmt;
        int want;
        void *a, *b, *c;
    } fields[] = {
        {"%d %d %d", 3, &in.nx, &in.ny, &in.nz},
        {"%d", 1, &in.ntimes, nullptr, nullptr},
        {"%d %d %d", 3, &in.nbinx, &in.nbiny, &in.nbinz},
        {"%lg", 1, &in.dt, nullptr, nullptr},
        {"%lg", 1, &in.t_request, nullptr, nullptr},
        {"%lg", 1, &in.rho, nullptr, nullptr},
        {"%d", 1, &in.every, nullptr, nullptr},
        {"%lg %lg", 2, &in.cutforce, &in.cutneigh, nullptr},
        {"%d", 1, &in.nstat, nullptr, nullptr},
    };
    int ln = 3;
    for (const Field &f : fields) {
        PAT_REQUIRE(get_line(fp.f, line), PAT_ERR_INVALID, "pat_read_ljin: %s ends before line %d", path, ln);
        const int got = sscanf(line.c_str(), f.fmt, f.a, f.b, f.c);
        PAT_REQUIRE(got == f.want, PAT_ERR_INVALID, "pat_read_ljin: %s line %d: expected %d value(s)", path, ln,
                    f.want);
        ln++;
    }
    PAT_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0 && in.dt > 0 && in.rho > 0 && in.cutforce > 0 &&
                    in.cutneigh >= in.cutforce,
                PAT_ERR_INVALID, "pat_read_ljin: %s: values out of range", path);
    *out = in;
    return PAT_OK;
}

int pat_read_mesh(const char *path, int zero_based, int64_t *num_nodes, int64_t *num_edges, int32_t *edges,
                  int64_t capacity)
{
    PAT_REQUIRE(path && num_nodes && num_edges, PAT_ERR_INVALID, "pat_read_mesh: null argument");
    File fp(path);
    PAT_REQUIRE(fp.f, PAT_ERR_INVALID, "pat_read_mesh: cannot open %s", path);
    std::string line;
    // leading comment lines start with '%'
    bool have = false;
    while ((have = get_line(fp.f, line)) && !line.empty() && line[0] == '%') {
    }
    if (!have) {
        *num_nodes = *num_edges = 0;
        return PAT_OK;
    }
    long long nn = 0, ne = 0;
    PAT_REQUIRE(sscanf(line.c_str(), "%lld %lld", &nn, &ne) == 2 && nn >= 0 && ne >= 0 && nn < (1LL << 31) &&
                    ne < (1LL << 31),
                PAT_ERR_INVALID, "pat_read_mesh: %s: bad header \"%s\"", path, line.c_str());
    PAT_REQUIRE(!edges || capacity >= ne, PAT_ERR_INVALID, "pat_read_mesh: capacity %lld below %lld edges",
                (long long)capacity, ne);
    int64_t k = 0;
    for (long long i = 0; i < nn; i++) {
        // node i's line lists its neighbours (1-based) until a non-positive
        // value or the end of the line; each edge is kept once, from the
        // node with the smaller number (neighbour >= i+1)
        if (!get_line(fp.f, line)) line.clear();
        const char *p = line.c_str();
        for (;;) {
            char *q = nullptr;
            errno = 0;
            const long v = strtol(p, &q, 10);
            if (q == p || v <= 0) break;
            PAT_REQUIRE(v <= nn, PAT_ERR_INVALID, "pat_read_mesh: %s: node %lld lists neighbour %ld > %lld", path,
                        i + 1, v, nn);
            p = q;
            if (v >= i + 1) {
                PAT_REQUIRE(k < ne, PAT_ERR_INVALID, "pat_read_mesh: %s holds more than the %lld edges of its header",
                            path, ne);
                if (edges) {
                    edges[2 * k] = (int32_t)(zero_based ? i : i + 1);
                    edges[2 * k + 1] = (int32_t)(zero_based ? v - 1 : v);
                }
                k++;
            }
        }
    }
    PAT_REQUIRE(k == ne, PAT_ERR_INVALID, "pat_read_mesh: %s: %lld edges read, header says %lld", path,
                (long long)k, ne);
    *num_nodes = nn;
    *num_edges = ne;
    return PAT_OK;
}

int pat_read_xyz(const char *path, int64_t num_nodes, double *xyz)
{
    PAT_REQUIRE(path && (xyz || num_nodes == 0), PAT_ERR_INVALID, "pat_read_xyz: null argument");
    File fp(path);
    PAT_REQUIRE(fp.f, PAT_ERR_INVALID, "pat_read_xyz: cannot open %s", path);
    int64_t i = 0;
    double a, b, c;
    while (fscanf(fp.f, "%lf %lf %lf", &a, &b, &c) == 3) {
        PAT_REQUIRE(i < num_nodes, PAT_ERR_INVALID, "pat_read_xyz: %s holds more than %lld coordinates", path,
                    (long long)num_nodes);
        xyz[3 * i] = a;
        xyz[3 * i + 1] = b;
        xyz[3 * i + 2] = c;
        i++;
    }
    PAT_REQUIRE(i == num_nodes, PAT_ERR_INVALID, "pat_read_xyz: %s holds %lld coordinates, expected %lld", path,
                (long long)i, (long long)num_nodes);
    return PAT_OK;
}

}  // extern "C"
