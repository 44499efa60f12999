// Device execution of an application's generalized-reduction functor table
// with keyed reduction objects.
//
// The reference runs map over 5,000-unit tasks, inserting (key, value) pairs
// into an LDS hash per thread group; when an insert finds the object full the
// map returns false, the group merges its LDS object into the device object
// and runs the unit again (kmeans/regular/gpu_kernel.cu:39-183, ros.cu:15-170,
// rog.cu:17-310), with per-bucket spin locks throughout.  Here:
//
//  * each wavefront takes a contiguous run of units and owns an LDS hash
//    table (SO, api/regular/objects.h) whose inserts the lanes make in turn,
//    in lane order -- no locks, a fixed fold order;
//  * a refused insert (table full) makes the wavefront merge its LDS table
//    into its own table in device memory, clear it, and run the refused
//    units' maps again; a map that returns false is run again the same way
//    (the reference's contract).  A map whose refused insert followed an
//    accepted one cannot be run again without inserting twice: that is an
//    error, as is a unit whose values do not fit an empty LDS table;
//  * one workgroup then folds the wavefronts' tables, in wavefront order,
//    into the final table (threads claim buckets with compare-and-swap; the
//    keys of one wavefront's table are distinct, so no two threads fold into
//    one key at once).
//
// Every key's value is therefore the same fold over the units every run.
// The library (pat_regular_*) sizes the tables and reads the final one.
//
//     #define GPU_KERNEL
//     #include "regular.h"
//     #include "gpu/regular_table.hpp"
//     PAT_GR_TABLE(FFGPU::map_function_table, FFGPU::reduce_function_table)
#ifndef PAT_GPU_REGULAR_TABLE_HPP_
#define PAT_GPU_REGULAR_TABLE_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "pattern.h"
#include "regular/objects.h"

namespace pat_gpu {

constexpr int kRoWaves = 4;          // wavefronts per workgroup of the map kernel
constexpr int kRoLdsBuckets = 128;   // largest LDS table per wavefront: buckets
constexpr int kRoLdsPool = 2048;     // and pool words (8 KiB)
constexpr uint32_t kRoPending = 0xffffffffu;

enum RoError { kRoWaveFull = 1, kRoUnitTooLarge = 2, kRoPartialRefusal = 4, kRoFinalFull = 8 };

// Single-writer merge of table s into table d (records in bucket order).
template <typename R>
__device__ bool ro_merge_into(RoTable &d, const RoTable &s, R reduce)
{
    for (int b = 0; b < s.nb; b++) {
        const uint32_t st = s.bk[b];
        if (st == 0u) continue;
        const ro_word *r = s.pool + (st - 1u);
        const int ks = (int)(r[0] >> 16), vs = (int)(r[0] & 0xffffu);
        if (!d.insert(r + 1, ks, r + 1 + RoTable::key_words(ks), vs, reduce)) return false;
    }
    return true;
}

template <class Tab, int MI, int RI>
__global__ __launch_bounds__(256) void ro_map(pat_gr_object a)
{
    __shared__ ro_word lds[kRoWaves * (kRoLdsBuckets + 1 + kRoLdsPool)];
    const int wl = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
    const int64_t w = (int64_t)blockIdx.x * kRoWaves + wl;
    if (w >= a.waves) return;  // whole wavefronts; no workgroup barrier below
    SO so;
    so.t = RoTable::at(lds + wl * (kRoLdsBuckets + 1 + kRoLdsPool), a.lds_buckets, a.lds_pool_words);
    so.t.clear(lane, 64);
    RoTable wt = RoTable::at((ro_word *)a.wave_tables + (size_t)w * RoTable::words(a.table_buckets, a.table_pool_words),
                             a.table_buckets, a.table_pool_words);
    wt.clear(lane, 64);
    __builtin_amdgcn_wave_barrier();
    const auto reduce = Tab::template reduce<RI>();
    const int64_t per = (a.num_units + a.waves - 1) / a.waves;
    const int64_t u0 = w * per, u1 = u0 + per < a.num_units ? u0 + per : a.num_units;
    int err = 0, flushes = 0, reruns = 0;
    for (int64_t base = u0; base < u1; base += 64) {
        const int64_t u = base + lane;
        bool pending = u < u1;
        // every round after a merge, the lowest pending lane inserts first
        // into an empty table, so each round finishes at least one unit
        // unless that unit cannot fit an empty table: 65 rounds suffice
        for (int round = 0; round < 66 && __ballot(pending); round++) {
            so.full = false;
            so.accepted = 0;
            bool ok = true;
            if (pending) {
                const unsigned long long idx =
                    a.offsets ? (unsigned long long)a.offsets[u] : (unsigned long long)(u * a.unit_bytes);
                ok = Tab::template map<MI, RI>(&so, const_cast<void *>(a.data), idx, const_cast<void *>(a.param));
            }
            if (pending && so.full && so.accepted > 0) {
                err |= kRoPartialRefusal;  // part of the unit went in: it cannot run again
                pending = false;
            }
            const bool again = pending && (!ok || so.full);
            if (__ballot(again)) {
                if (lane == 0 && !ro_merge_into(wt, so.t, reduce)) err |= kRoWaveFull;
                __builtin_amdgcn_wave_barrier();
                so.t.clear(lane, 64);
                __builtin_amdgcn_wave_barrier();
                flushes++;
                reruns += __popcll(__ballot(again));
            }
            pending = again;
        }
        if (__ballot(pending)) err |= kRoUnitTooLarge;
    }
    if (lane == 0 && !ro_merge_into(wt, so.t, reduce)) err |= kRoWaveFull;
    int werr = 0;
    for (int bit = 1; bit <= kRoFinalFull; bit <<= 1)
        if (__ballot((err & bit) != 0)) werr |= bit;
    err = werr;
    if (lane == 0) {
        if (err) atomicOr(&a.status[0], err);
        atomicAdd(&a.status[1], flushes + 1);
        atomicAdd(&a.status[2], reruns);
    }
}

// Concurrent insert into the final table: a bucket is claimed with a
// compare-and-swap; the inserts of one round carry distinct keys, so a
// bucket another thread is filling (kRoPending) never holds this key.
template <typename R>
__device__ bool ro_insert_shared(RoTable &d, const ro_word *rec, R reduce)
{
    const int ks = (int)(rec[0] >> 16), vs = (int)(rec[0] & 0xffffu);
    const ro_word *key = rec + 1;
    const ro_word *val = rec + 1 + RoTable::key_words(ks);
    const uint32_t h = ro_hash((const unsigned char *)key, ks);
    for (int p = 0; p < d.nb; p++) {
        const int b = (int)((h + (uint32_t)p) & (uint32_t)(d.nb - 1));
        uint32_t st = __atomic_load_n((uint32_t *)&d.bk[b], __ATOMIC_RELAXED);
        if (st == 0u) {
            st = atomicCAS((uint32_t *)&d.bk[b], 0u, kRoPending);
            if (st == 0u) {
                const int need = 1 + RoTable::key_words(ks) + ((vs + 3) >> 2);
                const uint32_t at = atomicAdd((uint32_t *)d.top, (uint32_t)need);
                if ((int)at + need > d.pw) return false;  // the bucket stays claimed: the result is void anyway
                ro_word *r = d.pool + at;
                for (int i = 0; i < need; i++) r[i] = rec[i];
                __threadfence_block();
                atomicExch((uint32_t *)&d.bk[b], at + 1u);
                return true;
            }
        }
        if (st == kRoPending) continue;
        ro_word *r = d.pool + (st - 1u);
        if (r[0] == rec[0] && ro_equal(r + 1, key, ks)) {
            reduce((void *)(r + 1 + RoTable::key_words(ks)), (unsigned short)vs, (void *)val, (unsigned short)vs);
            return true;
        }
    }
    return false;
}

template <class Tab, int RI>
__global__ __launch_bounds__(256) void ro_fold(pat_gr_object a)
{
    RoTable ft = RoTable::at((ro_word *)a.final_table, a.table_buckets, a.table_pool_words);
    ft.clear((int)threadIdx.x, (int)blockDim.x);
    __syncthreads();
    const auto reduce = Tab::template reduce<RI>();
    int err = 0;
    for (int w = 0; w < a.waves; w++) {
        const RoTable wt = RoTable::at((ro_word *)a.wave_tables +
                                           (size_t)w * RoTable::words(a.table_buckets, a.table_pool_words),
                                       a.table_buckets, a.table_pool_words);
        for (int b = (int)threadIdx.x; b < wt.nb; b += (int)blockDim.x) {
            const uint32_t st = wt.bk[b];
            if (st != 0u && !ro_insert_shared(ft, wt.pool + (st - 1u), reduce)) err |= kRoFinalFull;
        }
        __syncthreads();
    }
    if (err) atomicOr(&a.status[0], err);
}

template <class Tab, int MI, int RI>
int ro_launch(const pat_gr_object &a, hipStream_t st)
{
    if constexpr (!Tab::template valid<MI, RI>()) {
        return PAT_ERR_UNSUPPORTED;
    } else {
        hipLaunchKernelGGL((ro_map<Tab, MI, RI>), dim3((unsigned)((a.waves + kRoWaves - 1) / kRoWaves)),
                           dim3(64 * kRoWaves), 0, st, a);
        if (hipGetLastError() != hipSuccess) return PAT_ERR_HIP;
        hipLaunchKernelGGL((ro_fold<Tab, RI>), dim3(1), dim3(256), 0, st, a);
        return hipGetLastError() == hipSuccess ? PAT_OK : PAT_ERR_HIP;
    }
}

template <class Tab, int MI, std::size_t... RI>
int ro_dispatch_reduce(int ri, const pat_gr_object &a, hipStream_t st, std::index_sequence<RI...>)
{
    int rc = PAT_ERR_UNSUPPORTED;
    ((ri == (int)RI ? (void)(rc = ro_launch<Tab, MI, (int)RI>(a, st)) : (void)0), ...);
    return rc;
}

template <class Tab, std::size_t... MI>
int ro_dispatch(int mi, int ri, const pat_gr_object &a, hipStream_t st, std::index_sequence<MI...>)
{
    int rc = PAT_ERR_UNSUPPORTED;
    ((mi == (int)MI ? (void)(rc = ro_dispatch_reduce<Tab, (int)MI>(
                                 ri, a, st, std::make_index_sequence<(std::size_t)Tab::rcount>{}))
                    : (void)0),
     ...);
    return rc;
}

// pat_gr_table::map_reduce for the tables of Tab.
template <class Tab>
int ro_entry(int map_idx, int reduce_idx, const pat_gr_object *obj, void *stream)
{
    if (!obj || obj->waves <= 0 || obj->lds_buckets <= 0 || obj->lds_buckets > kRoLdsBuckets ||
        (obj->lds_buckets & (obj->lds_buckets - 1)) || obj->lds_pool_words <= 0 ||
        obj->lds_pool_words > kRoLdsPool || obj->table_buckets <= 0 ||
        (obj->table_buckets & (obj->table_buckets - 1)))
        return PAT_ERR_INVALID;
    return ro_dispatch<Tab>(map_idx, reduce_idx, *obj, (hipStream_t)stream,
                            std::make_index_sequence<(std::size_t)Tab::mcount>{});
}

}  // namespace pat_gpu

// Defines pat_app_gr_table() for the application's map and reduce tables
// (__device__ constexpr arrays of map_fp / reduce_fp).
#define PAT_GR_TABLE(MAP_TABLE, REDUCE_TABLE)                                                                     \
    namespace {                                                                                                   \
    struct PatGrTab {                                                                                             \
        static constexpr int mcount = (int)(sizeof(MAP_TABLE) / sizeof(MAP_TABLE[0]));                            \
        static constexpr int rcount = (int)(sizeof(REDUCE_TABLE) / sizeof(REDUCE_TABLE[0]));                      \
        template <int MI, int RI>                                                                                 \
        static constexpr bool valid()                                                                             \
        {                                                                                                         \
            return MAP_TABLE[MI] != nullptr && REDUCE_TABLE[RI] != nullptr;                                       \
        }                                                                                                         \
        template <int RI>                                                                                         \
        __device__ static constexpr auto reduce()                                                                 \
        {                                                                                                         \
            return REDUCE_TABLE[RI];                                                                              \
        }                                                                                                         \
        template <int MI, int RI>                                                                                 \
        __device__ static bool map(SO *o, void *data, unsigned long long index, void *param)                      \
        {                                                                                                         \
            return MAP_TABLE[MI](o, data, index, param, 1, REDUCE_TABLE[RI]);                                     \
        }                                                                                                         \
    };                                                                                                            \
    }                                                                                                             \
    extern "C" const pat_gr_table *pat_app_gr_table(void)                                                         \
    {                                                                                                             \
        static const pat_gr_table t = {PatGrTab::mcount, PatGrTab::rcount, pat_gpu::kRoLdsBuckets,                \
                                       pat_gpu::kRoLdsPool, &pat_gpu::ro_entry<PatGrTab>};                        \
        return &t;                                                                                                \
    }

#endif
