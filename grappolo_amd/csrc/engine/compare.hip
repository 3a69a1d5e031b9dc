// compare.hip -- cluster-comparison metrics on the device:
// computeCommunityComparisons (Utility/utilityClusterComparisonMetrics.cpp:50-304)
// and computeGiniCoefficient (:309-330).
//
// The reference enumerates every pair inside every community twice (once per
// labelling), O(sum of squared community sizes).  The same four counts follow
// from the contingency table n_ij of the two labellings:
//   same-same = agree = sum_ij C(n_ij, 2)
//   same-diff = sum_i C(a_i, 2) - same-same      (a_i: community sizes of C1)
//   diff-same = sum_j C(b_j, 2) - same-same      (b_j: community sizes of C2)
// The table's non-zero cells are the runs of the sorted keys i * nC2 + j, one
// key per vertex: a histogram, a radix sort and a run-length pass, all
// HBM-bound streaming work.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "compare.hpp"
#include "kernels.hpp"

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

namespace glv {

namespace {

constexpr int kCB = 256;

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// one atomic per workgroup: the block's sum of v into *acc (modular u64)
__device__ inline void block_add(unsigned long long v, unsigned long long *acc) {
    __shared__ unsigned long long sh[kCB / kWave];
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int q = 0; q < kCB / kWave; q++) t += sh[q];
        if (t) atomicAdd(acc, t);
    }
}

// max label + 1 (labels >= -1 make it non-negative) and flags:
// bit 0 = a label 0, bit 1 = a label -1, bit 2 = a label below -1
__global__ __launch_bounds__(kCB) void k_label_stats(const i64 *C, i64 n, unsigned long long *mx, int *flags) {
    __shared__ i64 sh[kCB / kWave];
    __shared__ int shf;
    if (threadIdx.x == 0) shf = 0;
    __syncthreads();
    i64 m = -1;
    int f = 0;
    for (i64 i = blockIdx.x * (i64)kCB + threadIdx.x; i < n; i += (i64)gridDim.x * kCB) {
        const i64 c = C[i];
        m = c > m ? c : m;
        f |= c == 0 ? 1 : 0;
        f |= c == -1 ? 2 : 0;
        f |= c < -1 ? 4 : 0;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const i64 t = __shfl_xor(m, o);
        m = t > m ? t : m;
    }
    if (f) atomicOr(&shf, f);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < kCB / kWave; q++) m = sh[q] > m ? sh[q] : m;
        if (m >= 0) atomicMax(mx, (unsigned long long)(m + 1));
        if (shf) atomicOr(flags, shf);
    }
}

__global__ void k_remap(i64 *C, i64 n, i64 from, i64 to) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        if (C[i] == from) C[i] = to;
}

// community sizes: label c counts at c - base (base 1 for one-based labels)
__global__ void k_sizes(const i64 *C, i64 n, i64 base, i64 nc, unsigned long long *sizes, int *err) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const i64 c = C[i] - base;
        if (c < 0 || c >= nc) {
            atomicOr(err, 1);
            continue;
        }
        atomicAdd(sizes + c, 1ull);
    }
}

__global__ void k_pair_keys(const i64 *C1, const i64 *C2, i64 n, i64 base1, i64 base2, i64 nc1, i64 nc2, u64 *keys,
                            int *err) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        const i64 a = C1[i] - base1, b = C2[i] - base2;
        if (a < 0 || a >= nc1 || b < 0 || b >= nc2) {
            atomicOr(err, 1);
            keys[i] = 0;
            continue;
        }
        keys[i] = (u64)a * (u64)nc2 + (u64)b;
    }
}

// sum of c (c - 1) / 2 over counts
template <typename T>
__global__ __launch_bounds__(kCB) void k_choose2(const T *cnt, i64 n, unsigned long long *acc) {
    unsigned long long s = 0;
    for (i64 i = blockIdx.x * (i64)kCB + threadIdx.x; i < n; i += (i64)gridDim.x * kCB) {
        const unsigned long long c = (unsigned long long)cnt[i];
        s += c * (c - (c > 0 ? 1 : 0)) / 2;
    }
    block_add(s, acc);
}

// Gini sums over sorted sizes: acc[0] += (i + 1) * s_i, acc[1] += s_i (two's
// complement, so negative sizes of a caller's array add up correctly)
__global__ __launch_bounds__(kCB) void k_gini_sums(const i64 *s, i64 n, unsigned long long *acc) {
    unsigned long long a = 0, b = 0;
    for (i64 i = blockIdx.x * (i64)kCB + threadIdx.x; i < n; i += (i64)gridDim.x * kCB) {
        a += (unsigned long long)((i + 1) * s[i]);
        b += (unsigned long long)s[i];
    }
    block_add(a, acc);
    __syncthreads();
    block_add(b, acc + 1);
}

template <typename T>
T fetch1(const T *d, hipStream_t s) {
    T h;
    GLV_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    return h;
}

struct LabelStats {
    i64 max;  // -1 when every label is -1
    int flags;
};

LabelStats label_stats(const i64 *C, i64 n, hipStream_t s) {
    DBuf<unsigned long long> mx(1, s);
    DBuf<int> fl(1, s);
    GLV_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned long long), s));
    GLV_HIP(hipMemsetAsync(fl.p, 0, sizeof(int), s));
    k_label_stats<<<grid_for(n, kCB, 2048), kCB, 0, s>>>(C, n, mx.p, fl.p);
    GLV_LAUNCH_CHECK();
    LabelStats r;
    unsigned long long m = 0;
    GLV_HIP(hipMemcpyAsync(&m, mx.p, sizeof(m), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipMemcpyAsync(&r.flags, fl.p, sizeof(int), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    r.max = (i64)m - 1;
    return r;
}

void remap(i64 *C, i64 n, i64 from, i64 to, hipStream_t s) {
    k_remap<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(C, n, from, to);
    GLV_LAUNCH_CHECK();
}

int bits_of(unsigned __int128 v) {
    int b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

}  // namespace

double gini_coefficient(i64 *d_sizes, i64 n, int numColors, hipStream_t s) {
    i64 num = 0, den = 0;
    if (n > 0) {
        if (n > 1) {
            DBuf<i64> tmp(n, s);
            size_t bytes = 0;
            GLV_HIP(rocprim::radix_sort_keys(nullptr, bytes, d_sizes, tmp.p, (size_t)n, 0, 64, s));
            DBuf<char> st(bytes, s);
            GLV_HIP(rocprim::radix_sort_keys(st.p, bytes, d_sizes, tmp.p, (size_t)n, 0, 64, s));
            GLV_HIP(hipMemcpyAsync(d_sizes, tmp.p, n * sizeof(i64), hipMemcpyDeviceToDevice, s));
        }
        DBuf<unsigned long long> acc(2, s);
        GLV_HIP(hipMemsetAsync(acc.p, 0, 2 * sizeof(unsigned long long), s));
        k_gini_sums<<<grid_for(n, kCB, 2048), kCB, 0, s>>>(d_sizes, n, acc.p);
        GLV_LAUNCH_CHECK();
        unsigned long long h[2];
        GLV_HIP(hipMemcpyAsync(h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
        GLV_HIP(hipStreamSynchronize(s));
        num = (i64)h[0];
        den = (i64)h[1];
    }
    // the reference's expression (utilityClusterComparisonMetrics.cpp:322);
    // its sums are taken in double in index order, equal to these integer
    // sums converted once while every partial sum stays below 2^53
    const double numFunc = (double)num, denFunc = (double)den;
    return ((2 * numFunc) / (numColors * denFunc)) - ((double)(numColors + 1) / (double)numColors);
}

void compare_communities(i64 *C1, i64 *C2, i64 N, CompareResult &r, hipStream_t s) {
    if (N <= 0) fail(GLV_E_INVALID_ARG, "comparison needs at least one vertex");
    if (!C1 || !C2) fail(GLV_E_INVALID_ARG, "null labels");
    r = CompareResult();
    // community counts and the -1 relabelling, :62-107
    const LabelStats s1 = label_stats(C1, N, s), s2 = label_stats(C2, N, s);
    if ((s1.flags | s2.flags) & 4) fail(GLV_E_INVALID_ARG, "label below -1");
    const bool zero1 = s1.flags & 1, zero2 = s2.flags & 1;
    i64 nC1 = s1.max + (zero1 ? 1 : 0);
    if (nC1 <= 0) fail(GLV_E_INVALID_ARG, "truth labelling has no community");
    if (s1.flags & 2) {
        remap(C1, N, -1, nC1, s);
        nC1++;
    }
    i64 max2 = s2.max;
    if (s2.flags & 2) {
        remap(C2, N, -1, N + 1, s);
        max2 = std::max<i64>(max2, N + 1);
    }
    const i64 nC2 = max2 + (zero2 ? 1 : 0);
    if (nC2 <= 0) fail(GLV_E_INVALID_ARG, "output labelling has no community");
    if (nC1 > INT32_MAX || nC2 > INT32_MAX)
        fail(GLV_E_UNSUPPORTED, "more than 2^31-1 communities (the Gini count is an int)");
    const i64 base1 = zero1 ? 0 : 1, base2 = zero2 ? 0 : 1;
    r.nC1 = nC1;
    r.nC2 = nC2;

    DBuf<int> err(1, s);
    GLV_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    // community sizes (the reference's clusterDist arrays) and sum C(size, 2)
    DBuf<i64> sz1(nC1, s), sz2(nC2, s);
    launch_fill_i64(sz1.p, nC1, 0, s);
    launch_fill_i64(sz2.p, nC2, 0, s);
    k_sizes<<<grid_for(N, kBlock, 8192), kBlock, 0, s>>>(C1, N, base1, nC1, (unsigned long long *)sz1.p, err.p);
    GLV_LAUNCH_CHECK();
    k_sizes<<<grid_for(N, kBlock, 8192), kBlock, 0, s>>>(C2, N, base2, nC2, (unsigned long long *)sz2.p, err.p);
    GLV_LAUNCH_CHECK();
    DBuf<unsigned long long> acc(3, s);
    GLV_HIP(hipMemsetAsync(acc.p, 0, 3 * sizeof(unsigned long long), s));
    k_choose2<i64><<<grid_for(nC1, kCB, 2048), kCB, 0, s>>>(sz1.p, nC1, acc.p + 0);
    GLV_LAUNCH_CHECK();
    k_choose2<i64><<<grid_for(nC2, kCB, 2048), kCB, 0, s>>>(sz2.p, nC2, acc.p + 1);
    GLV_LAUNCH_CHECK();

    // contingency cells: runs of the sorted pair keys
    const int kbits = bits_of((unsigned __int128)nC1 * (unsigned __int128)nC2 - 1);
    if (kbits > 64) fail(GLV_E_UNSUPPORTED, "community counts too large for 64-bit pair keys");
    {
        DBuf<u64> keys(N, s), sorted(N, s);
        k_pair_keys<<<grid_for(N, kBlock, 8192), kBlock, 0, s>>>(C1, C2, N, base1, base2, nC1, nC2, keys.p, err.p);
        GLV_LAUNCH_CHECK();
        size_t bytes = 0;
        GLV_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys.p, sorted.p, (size_t)N, 0, std::max(kbits, 1), s));
        {
            DBuf<char> st(bytes, s);
            GLV_HIP(rocprim::radix_sort_keys(st.p, bytes, keys.p, sorted.p, (size_t)N, 0, std::max(kbits, 1), s));
        }
        DBuf<i64> counts(N, s), nruns(1, s);
        u64 *uniq = keys.p;  // the unsorted keys are no longer needed
        bytes = 0;
        GLV_HIP(rocprim::run_length_encode(nullptr, bytes, sorted.p, (size_t)N, uniq, counts.p, nruns.p, s));
        DBuf<char> st(bytes, s);
        GLV_HIP(rocprim::run_length_encode(st.p, bytes, sorted.p, (size_t)N, uniq, counts.p, nruns.p, s));
        const i64 runs = fetch1(nruns.p, s);
        k_choose2<i64><<<grid_for(runs, kCB, 2048), kCB, 0, s>>>(counts.p, runs, acc.p + 2);
        GLV_LAUNCH_CHECK();
    }
    unsigned long long h[3];
    int e = 0;
    GLV_HIP(hipMemcpyAsync(h, acc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    if (e) fail(GLV_E_INVALID_ARG, "label outside the community range");
    r.SameSame = (i64)h[2];
    r.Agree = r.SameSame;
    r.SameDiff = (i64)h[0] - r.SameSame;
    r.DiffSame = (i64)h[1] - r.SameSame;
    // :239-245
    r.precision = (double)r.SameSame / (double)(r.SameSame + r.DiffSame);
    r.recall = (double)r.SameSame / (double)(r.SameSame + r.SameDiff);
    r.fScore = 2 * ((r.precision * r.recall) / (r.precision + r.recall));
    r.gini1 = gini_coefficient(sz1.p, nC1, (int)nC1, s);
    r.gini2 = gini_coefficient(sz2.p, nC2, (int)nC2, s);
}

}  // namespace glv
