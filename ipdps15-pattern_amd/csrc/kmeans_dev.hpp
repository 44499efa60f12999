// Device state of one kmeans problem, shared by the kernels and the handle.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace pat {

struct KmDev {
    const float *x;  // N x D points (device)
    float *c;        // K x D centres (FP32)
    int64_t N;
    int D, K, K_pad, KS;  // K_pad: K rounded up to 32; KS: 16-wide K-steps of the data
    int NT;               // MFMA data terms: 3 (split BF16), 1 (BF16), 0 (no MFMA path)
    int gc;               // MFMA rows read from global memory (K x D too large for LDS)
    int shift, dshift;    // fixed-point scales 2^shift (coordinates), 2^dshift (distances)
    int32_t *assign;      // N
    float *mindist;       // N
    unsigned long long *gsum, *gcnt, *gdsum, *csnap;  // reduction object (K x D, K, K), count snapshot
    void *cb;             // BF16 MFMA rows, K_pad x 16*(NT*KS+1)
    float *cnorm;         // ||c_k|| (rounded up)
    unsigned long long *rescored;  // points with more than one candidate
    int cus;              // compute units
    int bucketed;         // reduction object too large for LDS: assignments first, then bucket sums
    int *bidx, *bcnt, *boff, *bcur;  // N bucketed point indices; K counts, K+1 offsets, K cursors
};

int kmeans_fixed_shift(int64_t n, float maxabs);
int kmeans_ks_for(int D);
bool kmeans_mfma_fits(int K, int D);
int kmeans_nt_for(int K, int D);
bool kmeans_gc_for(int K, int D);
bool kmeans_object_in_lds(int K, int D);
int kmeans_maxabs(const float *x, int64_t n, unsigned int *out_bits, hipStream_t st);
int kmeans_finalize(const KmDev &d, int update, hipStream_t st);
int kmeans_assign_plain(const KmDev &d, hipStream_t st);
int kmeans_assign_mfma(const KmDev &d, hipStream_t st);
int kmeans_zero(unsigned long long *p, int64_t n, hipStream_t st);
int kmeans_prepare_kernels();

}  // namespace pat
