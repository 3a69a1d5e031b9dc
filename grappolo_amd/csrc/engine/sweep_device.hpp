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

template <int W>
__device__ inline void group_argmax(bool &cand, double &gain, i64 &id) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        double og = __shfl_xor(gain, o, W);
        long long oid = __shfl_xor((long long)id, o, W);
        int ocd = __shfl_xor((int)cand, o, W);
        if (better(ocd != 0, og, oid, cand, gain, id)) {
            cand = true;
            gain = og;
            id = oid;
        }
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
