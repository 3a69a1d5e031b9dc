// sweep_device.hpp -- device helpers shared by the sweep kernels
// (kernels.hip) and the whole-GPU persistent sweep (wide.hip): the gain
// expression, the argmax order, community moves and the grid barrier.
#pragma once

#include "kernels.hpp"

namespace glv {
namespace {

constexpr u64 kFib = 0x9E3779B97F4A7C15ULL;

__device__ inline i64 ld_relaxed(const i64 *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
// Loads of tables that other lanes fill with atomics (performed in L2): agent
// scope, so a line cached in the CU's L1 before the atomics cannot be served.
__device__ inline double ld_agent(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline i64 ld_agent(const i64 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Gain of moving a vertex of degree kv from community s (internal weight
// eix excluding its self-loop, degree ax excluding the vertex) to community y
// holding weight eiy towards it and degree ay: max() at
// utilityClusteringFunctions.cpp:243, same operation order, no contraction.
__device__ inline double gain_of(double eiy, double eix, double kv, double ay, double ax,
                                 double constant) {
    return 2 * (eiy - eix) - 2 * kv * (ay - ax) * constant;
}

// 1/(2m) of the graph vertex v belongs to (batched graphs carry one per graph)
__device__ inline double vconst(const SweepArgs &A, i64 v) {
    return A.gid ? A.gconst[A.gid[v]] : A.constant;
}

// "better" of two candidates: larger gain, then smaller community id.
__device__ inline bool better(bool oc, double og, i64 oid, bool c, double g, i64 id) {
    return oc && (!c || og > g || (og == g && oid < id));
}

// A candidate (cand, gain) as one unsigned key whose order is better()'s
// order on gains: 0 for no candidate, otherwise the gain's bits mapped so
// that unsigned order is the doubles' order (-0 taken as +0; never 0 for a
// non-NaN gain).
__device__ inline u64 argmax_key(bool cand, double gain) {
    const u64 b = (u64)__double_as_longlong(gain + 0.0);
    return cand ? ((b >> 63) ? ~b : (b | (1ULL << 63))) : 0;
}
__device__ inline double key_gain(u64 k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ULL << 63)) : ~k));
}

// 64-bit value of another lane in the same row of 16 (DPP control CTRL)
template <int CTRL>
__device__ inline u64 dpp64(u64 x) {
    const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)x, CTRL, 0xf, 0xf, false);
    const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(x >> 32), CTRL, 0xf, 0xf, false);
    return ((u64)hi << 32) | lo;
}
template <int CTRL, bool N>
__device__ inline void argmax_dpp(u64 &key, i64 &id, i64 &nid) {
    const u64 ok = dpp64<CTRL>(key);
    const i64 oid = (i64)dpp64<CTRL>((u64)id);
    const bool t = ok > key || (ok == key && oid < id);
    key = t ? ok : key;
    id = t ? oid : id;
    if (N) {
        const i64 on = (i64)dpp64<CTRL>((u64)nid);
        nid = t ? on : nid;
    }
}
template <bool N>
__device__ inline void argmax_xor(u64 &key, i64 &id, i64 &nid, int o) {
    const u64 ok = (u64)__shfl_xor((long long)key, o);
    const i64 oid = __shfl_xor((long long)id, o);
    const bool t = ok > key || (ok == key && oid < id);
    key = t ? ok : key;
    id = t ? oid : id;
    if (N) {
        const i64 on = __shfl_xor((long long)nid, o);
        nid = t ? on : nid;
    }
}

// Maximum of (key, id) over aligned groups of W lanes (W a power of two up
// to 64), with `nid` following the winner when N; the result in every lane
// of the group.  Within a row of 16 lanes by DPP (lane pairs i^1 and i^2 by
// quad permutes, then the half-row and row mirrors pair each lane with the
// other half of its group of 8 and of 16), across rows by ds_bpermute.  It is
// the maximum of a total order, so the pairing does not change the result.
template <int W, bool N = false>
__device__ inline void group_argmax_key(u64 &key, i64 &id, i64 &nid) {
    if (W >= 2) argmax_dpp<0xB1, N>(key, id, nid);   // quad_perm [1,0,3,2]
    if (W >= 4) argmax_dpp<0x4E, N>(key, id, nid);   // quad_perm [2,3,0,1]
    if (W >= 8) argmax_dpp<0x141, N>(key, id, nid);  // row_half_mirror
    if (W >= 16) argmax_dpp<0x140, N>(key, id, nid); // row_mirror
    if (W >= 32) argmax_xor<N>(key, id, nid, 16);
    if (W >= 64) argmax_xor<N>(key, id, nid, 32);
}

// better()'s maximum over aligned groups of W lanes, in every lane of the
// group; a group without a candidate leaves each lane's (gain, id) as it was.
template <int W>
__device__ inline void group_argmax(bool &cand, double &gain, i64 &id) {
    u64 key = argmax_key(cand, gain);
    i64 rid = id, none = 0;
    group_argmax_key<W>(key, rid, none);
    cand = key != 0;
    if (cand) {
        gain = key_gain(key);
        id = rid;
    }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline void move_vertex(i64 t, i64 s, double kv, double *a, i64 *n) {
    atomicAdd(&a[t], kv);
    atomicAdd(&a[s], -kv);
    atomicAdd((unsigned long long *)&n[t], 1ULL);
    atomicAdd((unsigned long long *)&n[s], ~0ULL);  // -1
}

// Spins of a grid barrier before it gives up (an error flag, never a hang).
constexpr unsigned kSpinLimit = 1u << 22;

// Grid barrier over nb workgroups without a shared counter: workgroup
// b publishes the barrier's sequence number in its own flag slot (one cache
// line each) with an agent-scope release, and the first wavefront of every
// workgroup polls all nb slots at once with relaxed agent-scope loads, then
// takes one agent-scope acquire (one L2 write-back per arrival and one
// invalidate per departure, instead of an invalidate per poll: 2.2 us per
// barrier at 64 workgroups against 4.3 us with system-scope polling).  Above
// 64 workgroups each lane polls every 64th slot.  bar holds 16 + 16 * nb words.
// Bounded wait; a workgroup that gives up raises bar[0] (abort), which
// releases every other wait.  Returns false when the barrier was abandoned.
__device__ bool grid_barrier(unsigned *bar, unsigned nb, unsigned seq) {
    __shared__ int s_ok;
    unsigned *flags = bar + 16;  // flag of workgroup b at flags[16 * b]
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(&flags[16 * blockIdx.x], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        unsigned spins = 0;
        int ok = 1;
        while (true) {
            bool done = true;
            for (unsigned b = lane; b < nb; b += kWave)
                done = done &&
                       __hip_atomic_load(&flags[16 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= seq;
            if (__ballot(!done) == 0) break;
            if (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > kSpinLimit) {
                if (lane == 0) atomicOr(&bar[0], 1u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (lane == 0) s_ok = ok && !__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace
}  // namespace glv
