// batch.hip -- kernels of the batched plain Louvain: many independent graphs
// held as one block-diagonal union CSR (graph g owns the vertex range
// [vb[g], vb[g+1])), swept together, with per-graph 1/(2m), modularity,
// termination and phase progress.
#include <algorithm>

#include "active_set.hpp"
#include "batch.hpp"
#include "kernels.hpp"

namespace glv {

namespace {

constexpr int kBB = 256;

__device__ inline double bsum(double v) {
    __shared__ double sh[kBB / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    for (int q = 0; q < kBB / kWave; q++) r += sh[q];
    return r;
}

__global__ void k_fill_gid(const i64 *vb, i64 ng, int32_t *gid) {
    for (i64 g = blockIdx.x; g < ng; g += gridDim.x)
        for (i64 v = vb[g] + threadIdx.x; v < vb[g + 1]; v += blockDim.x) gid[v] = (int32_t)g;
}

// out[g] = 1 / sum of k over graph g (0 when the sum is 0)
__global__ __launch_bounds__(kBB) void k_graph_const(const i64 *vb, i64 ng, const double *k, double *gconst) {
    for (i64 g = blockIdx.x; g < ng; g += gridDim.x) {
        double t = 0;
        for (i64 v = vb[g] + threadIdx.x; v < vb[g + 1]; v += kBB) t += k[v];
        t = bsum(t);
        if (threadIdx.x == 0) gconst[g] = t > 0 ? (double)1 / t : 0.0;
    }
}

// per active graph: internal weight and sum of squared community degrees
__global__ __launch_bounds__(kBB) void k_graph_mod(const i64 *vb, i64 ng, const uint8_t *active,
                                                   const double *eii, const double *a, double *exx,
                                                   double *a2, const int32_t *skip) {
    if (skip && *skip) return;
    for (i64 g = blockIdx.x; g < ng; g += gridDim.x) {
        if (!active[g]) continue;
        double e = 0, q = 0;
        for (i64 v = vb[g] + threadIdx.x; v < vb[g + 1]; v += kBB) {
            e += eii[v];
            q += a[v] * a[v];
        }
        e = bsum(e);
        q = bsum(q);
        if (threadIdx.x == 0) {
            exx[g] = e;
            a2[g] = q;
        }
    }
}

// The termination test of parallelLouvianMethod (parallelLouvainMethod.cpp:
// 266-273) per graph; apply[g] = 1 when graph g commits this iteration's moves.
__global__ void k_batch_decide(i64 ng, const double *exx, const double *a2, const double *gconst,
                               const double *lower, double thresh, int maxitr, double *prev, double *curr,
                               int32_t *itrs, uint8_t *active, uint8_t *apply) {
    for (i64 g = blockIdx.x * (i64)blockDim.x + threadIdx.x; g < ng; g += (i64)gridDim.x * blockDim.x) {
        apply[g] = 0;
        if (!active[g]) continue;
        itrs[g] += 1;
        const double c = gconst[g];
        const double q = (exx[g] * (double)c) - (a2[g] * (double)c * (double)c);
        curr[g] = q;
        if ((q - prev[g]) < thresh || itrs[g] >= maxitr) {
            active[g] = 0;
            continue;
        }
        double p = q;
        if (p < lower[g]) p = lower[g];
        prev[g] = p;
        apply[g] = 1;
    }
}

// After an iteration: live = vertices of the graphs still iterating, done =
// (none is); one workgroup, so a captured run of iterations can test it.
__global__ __launch_bounds__(kBB) void k_level_status(const i64 *vb, i64 ng, const uint8_t *active,
                                                      int32_t *done, i64 *live) {
    __shared__ i64 s_part[kBB / kWave];
    i64 t = 0;
    for (i64 g = threadIdx.x; g < ng; g += kBB)
        if (active[g]) t += vb[g + 1] - vb[g];
    for (int o = kWave / 2; o > 0; o >>= 1) t += __shfl_xor(t, o);
    if (threadIdx.x % kWave == 0) s_part[threadIdx.x / kWave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        i64 sum = 0;
        for (int i = 0; i < kBB / kWave; i++) sum += s_part[i];
        *live = sum;
        *done = sum == 0 ? 1 : 0;
    }
}

// commit of one Jacobi iteration of the graphs with apply set
// (parallelLouvainMethod.cpp:275-283): community updates, then past <- curr <- tgt
__global__ void k_batch_apply(i64 nv, const int32_t *gid, const uint8_t *apply, const i64 *tgt, i64 *curr,
                              i64 *past, const double *k, double *a, i64 *n, const int32_t *skip) {
    if (skip && *skip) return;
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        if (!apply[gid[v]]) continue;
        const i64 t = tgt[v], s = curr[v];
        if (t != s && t != -1) {
            atomicAdd(&a[t], k[v]);
            atomicAdd(&a[s], -k[v]);
            atomicAdd((unsigned long long *)&n[t], 1ULL);
            atomicAdd((unsigned long long *)&n[s], ~0ULL);
        }
        past[v] = s;
        curr[v] = t;
    }
}

__global__ void k_flag_graphs(const i64 *list, i64 n, const int32_t *gid, const uint8_t *active, i64 *flags) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        flags[i] = active[gid[list[i]]] ? 1 : 0;
}

// per graph: smallest and largest non-negative label (renumbered labels of a
// graph form one contiguous range)
__global__ __launch_bounds__(kBB) void k_graph_label_range(const i64 *vb, i64 ng, const i64 *C, i64 *lo,
                                                           i64 *hi) {
    __shared__ i64 smin, smax;
    for (i64 g = blockIdx.x; g < ng; g += gridDim.x) {
        if (threadIdx.x == 0) {
            smin = INT64_MAX;
            smax = -1;
        }
        __syncthreads();
        i64 mn = INT64_MAX, mx = -1;
        for (i64 v = vb[g] + threadIdx.x; v < vb[g + 1]; v += kBB) {
            const i64 c = C[v];
            if (c >= 0) {
                mn = c < mn ? c : mn;
                mx = c > mx ? c : mx;
            }
        }
        if (mx >= 0) {
            atomicMin((unsigned long long *)&smin, (unsigned long long)mn);
            atomicMax((long long *)&smax, (long long)mx);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            lo[g] = smax >= 0 ? smin : -1;
            hi[g] = smax;
        }
        __syncthreads();
    }
}

// C_orig[i] = C[C_orig[i]] (or C[i] in the first phase) for the original
// vertices of graphs still running
__global__ void k_batch_compose(i64 n0, const int32_t *gid0, const uint8_t *alive, i64 *C_orig, const i64 *C,
                                bool first) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n0; i += (i64)gridDim.x * blockDim.x) {
        if (!alive[gid0[i]]) continue;
        if (first) {
            C_orig[i] = C[i];
        } else {
            const i64 c = C_orig[i];
            C_orig[i] = c >= 0 ? C[c] : -1;
        }
    }
}

// labels of graphs that stop now become graph-local
__global__ void k_batch_finalize(i64 n0, const int32_t *gid0, const uint8_t *stop, const i64 *base,
                                 i64 *C_orig) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n0; i += (i64)gridDim.x * blockDim.x) {
        const int32_t g = gid0[i];
        if (!stop[g]) continue;
        const i64 c = C_orig[i];
        C_orig[i] = c >= 0 ? c - base[g] : -1;
    }
}

__global__ void k_gather_i64(const i64 *src, const i64 *idx, i64 n, i64 *dst) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

// C[v] = -1 for the vertices of graphs with keep[g] == 0
__global__ void k_mask_graphs(const i64 *vb, i64 ng, const uint8_t *keep, i64 *C) {
    for (i64 g = blockIdx.x; g < ng; g += gridDim.x) {
        if (keep[g]) continue;
        for (i64 v = vb[g] + threadIdx.x; v < vb[g + 1]; v += blockDim.x) C[v] = -1;
    }
}

// err |= 1 when an entry of graph g's rows points outside [vb[g], vb[g+1])
__global__ void k_check_blocks(const i64 *off, const i64 *tails, const i64 *vb, i64 ng, int32_t *err) {
    for (i64 g = blockIdx.x; g < ng; g += gridDim.x) {
        const i64 lo = vb[g], hi = vb[g + 1];
        for (i64 j = off[lo] + threadIdx.x; j < off[hi]; j += blockDim.x) {
            const i64 t = tails[j];
            if (t < lo || t >= hi) atomicOr(err, 1);
        }
    }
}

}  // namespace

#define GLV_LAUNCH_CHECK() GLV_HIP(hipGetLastError())

void launch_check_blocks(const i64 *off, const i64 *tails, const i64 *vb, i64 ng, int32_t *err, hipStream_t s) {
    if (ng <= 0) return;
    k_check_blocks<<<(unsigned)std::max<i64>(1, std::min<i64>(ng, 65535)), kBB, 0, s>>>(off, tails, vb, ng, err);
    GLV_LAUNCH_CHECK();
}

void launch_gather_i64(const i64 *src, const i64 *idx, i64 n, i64 *dst, hipStream_t s) {
    if (n <= 0) return;
    k_gather_i64<<<grid_for(n, kBB, 4096), kBB, 0, s>>>(src, idx, n, dst);
    GLV_LAUNCH_CHECK();
}
void launch_mask_graphs(const i64 *vb, i64 ng, const uint8_t *keep, i64 *C, hipStream_t s) {
    if (ng <= 0) return;
    k_mask_graphs<<<(unsigned)std::max<i64>(1, std::min<i64>(ng, 65535)), kBB, 0, s>>>(vb, ng, keep, C);
    GLV_LAUNCH_CHECK();
}

static unsigned graph_grid(i64 ng) { return (unsigned)std::max<i64>(1, std::min<i64>(ng, 65535)); }

void launch_fill_gid(const i64 *vb, i64 ng, int32_t *gid, hipStream_t s) {
    if (ng <= 0) return;
    k_fill_gid<<<graph_grid(ng), kBB, 0, s>>>(vb, ng, gid);
    GLV_LAUNCH_CHECK();
}
void launch_graph_const(const i64 *vb, i64 ng, const double *k, double *gconst, hipStream_t s) {
    if (ng <= 0) return;
    k_graph_const<<<graph_grid(ng), kBB, 0, s>>>(vb, ng, k, gconst);
    GLV_LAUNCH_CHECK();
}
void launch_graph_mod(const i64 *vb, i64 ng, const uint8_t *active, const double *eii, const double *a,
                      double *exx, double *a2, hipStream_t s, const int32_t *skip) {
    if (ng <= 0) return;
    k_graph_mod<<<graph_grid(ng), kBB, 0, s>>>(vb, ng, active, eii, a, exx, a2, skip);
    GLV_LAUNCH_CHECK();
}
void launch_batch_decide(i64 ng, const double *exx, const double *a2, const double *gconst, const double *lower,
                         double thresh, int maxitr, double *prev, double *curr, int32_t *itrs, uint8_t *active,
                         uint8_t *apply, hipStream_t s) {
    if (ng <= 0) return;
    k_batch_decide<<<grid_for(ng, kBB, 1024), kBB, 0, s>>>(ng, exx, a2, gconst, lower, thresh, maxitr, prev,
                                                          curr, itrs, active, apply);
    GLV_LAUNCH_CHECK();
}
void launch_level_status(const i64 *vb, i64 ng, const uint8_t *active, int32_t *done, i64 *live, hipStream_t s) {
    k_level_status<<<1, kBB, 0, s>>>(vb, ng, active, done, live);
    GLV_LAUNCH_CHECK();
}
void launch_batch_apply(i64 nv, const int32_t *gid, const uint8_t *apply, const i64 *tgt, i64 *curr, i64 *past,
                        const double *k, double *a, i64 *n, hipStream_t s, const int32_t *skip) {
    if (nv <= 0) return;
    k_batch_apply<<<grid_for(nv, kBB, 8192), kBB, 0, s>>>(nv, gid, apply, tgt, curr, past, k, a, n, skip);
    GLV_LAUNCH_CHECK();
}
void launch_compact_graphs(const i64 *list, i64 n, const int32_t *gid, const uint8_t *active, const i64 *start,
                           i64 nkeys, i64 *flags, i64 *pos, i64 *out, i64 *out_start, hipStream_t s) {
    if (n <= 0) return;
    k_flag_graphs<<<grid_for(n, kBB, 8192), kBB, 0, s>>>(list, n, gid, active, flags);
    GLV_LAUNCH_CHECK();
    compact_flagged(list, n, flags, pos, start, nkeys, out, out_start, s);
}
void launch_graph_label_range(const i64 *vb, i64 ng, const i64 *C, i64 *lo, i64 *hi, hipStream_t s) {
    if (ng <= 0) return;
    k_graph_label_range<<<graph_grid(ng), kBB, 0, s>>>(vb, ng, C, lo, hi);
    GLV_LAUNCH_CHECK();
}
void launch_batch_compose(i64 n0, const int32_t *gid0, const uint8_t *alive, i64 *C_orig, const i64 *C,
                          bool first, hipStream_t s) {
    if (n0 <= 0) return;
    k_batch_compose<<<grid_for(n0, kBB, 8192), kBB, 0, s>>>(n0, gid0, alive, C_orig, C, first);
    GLV_LAUNCH_CHECK();
}
void launch_batch_finalize(i64 n0, const int32_t *gid0, const uint8_t *stop, const i64 *base, i64 *C_orig,
                           hipStream_t s) {
    if (n0 <= 0) return;
    k_batch_finalize<<<grid_for(n0, kBB, 8192), kBB, 0, s>>>(n0, gid0, stop, base, C_orig);
    GLV_LAUNCH_CHECK();
}

}  // namespace glv
