// Device memory of the library: the runtime allocator, or a caller-bound
// arena.  An arena is one hipMalloc'd region handed out first-fit in
// 256-byte units from an offset-ordered free list, with neighbours merged on
// free, under one mutex.
#include <map>
#include <mutex>

#include "pat_common.hpp"

namespace pat {

struct Arena {
    char *base = nullptr;
    size_t size = 0;
    std::mutex mu;
    std::map<size_t, size_t> free_;   // offset -> length
    std::map<size_t, size_t> used_;   // offset -> length
    size_t in_use = 0;
};

namespace {

constexpr size_t kAlign = 256;
thread_local Arena *t_arena = nullptr;
std::mutex g_reg_mu;
std::vector<Arena *> g_arenas;  // live arenas, for dev_free's owner lookup

Arena *owner_of(const void *p)
{
    std::lock_guard<std::mutex> g(g_reg_mu);
    for (Arena *a : g_arenas)
        if ((const char *)p >= a->base && (const char *)p < a->base + a->size) return a;
    return nullptr;
}

void *arena_alloc(Arena *a, size_t bytes)
{
    const size_t n = (bytes + kAlign - 1) / kAlign * kAlign;
    std::lock_guard<std::mutex> g(a->mu);
    for (auto it = a->free_.begin(); it != a->free_.end(); ++it) {
        if (it->second < n) continue;
        const size_t off = it->first, len = it->second;
        a->free_.erase(it);
        if (len > n) a->free_.emplace(off + n, len - n);
        a->used_.emplace(off, n);
        a->in_use += n;
        return a->base + off;
    }
    return nullptr;
}

bool arena_release(Arena *a, void *p)
{
    std::lock_guard<std::mutex> g(a->mu);
    const size_t off = (size_t)((char *)p - a->base);
    auto u = a->used_.find(off);
    if (u == a->used_.end()) return false;
    size_t o = off, len = u->second;
    a->used_.erase(u);
    a->in_use -= len;
    auto next = a->free_.lower_bound(o);
    if (next != a->free_.end() && o + len == next->first) {
        len += next->second;
        next = a->free_.erase(next);
    }
    if (next != a->free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == o) {
            o = prev->first;
            len += prev->second;
            a->free_.erase(prev);
        }
    }
    a->free_.emplace(o, len);
    return true;
}

}  // namespace

hipError_t dev_malloc(void **p, size_t bytes)
{
    if (Arena *a = t_arena) {
        *p = arena_alloc(a, bytes ? bytes : 1);
        return *p ? hipSuccess : hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

hipError_t dev_free(void *p)
{
    if (!p) return hipSuccess;
    if (Arena *a = owner_of(p)) return arena_release(a, p) ? hipSuccess : hipErrorInvalidValue;
    return hipFree(p);
}

Arena *arena_create(size_t bytes)
{
    auto *a = new Arena();
    a->size = (bytes + kAlign - 1) / kAlign * kAlign;
    if (hipMalloc((void **)&a->base, a->size) != hipSuccess) {
        (void)hipGetLastError();
        set_error("arena: hipMalloc of %zu bytes failed", a->size);
        delete a;
        return nullptr;
    }
    a->free_.emplace(0, a->size);
    std::lock_guard<std::mutex> g(g_reg_mu);
    g_arenas.push_back(a);
    return a;
}

void arena_destroy(Arena *a)
{
    if (!a) return;
    {
        std::lock_guard<std::mutex> g(g_reg_mu);
        for (size_t i = 0; i < g_arenas.size(); i++)
            if (g_arenas[i] == a) {
                g_arenas.erase(g_arenas.begin() + (long)i);
                break;
            }
    }
    (void)hipFree(a->base);
    delete a;
}

void arena_bind(Arena *a) { t_arena = a; }

size_t arena_in_use(const Arena *a) { return a ? a->in_use : 0; }

size_t arena_largest_free(const Arena *a)
{
    if (!a) return 0;
    std::lock_guard<std::mutex> g(const_cast<Arena *>(a)->mu);
    size_t m = 0;
    for (auto &f : a->free_) m = std::max(m, f.second);
    return m;
}

}  // namespace pat
