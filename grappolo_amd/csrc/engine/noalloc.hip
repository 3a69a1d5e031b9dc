// noalloc.hip -- modularity of a labelling without any allocation: the
// variant of glv_modularity_dev for callers that cannot allocate device
// memory (code under stream capture, serving loops).  Its scratch memory is a
// file-scope device array (one copy per GPU, held until the process ends); the
// device address of each GPU's copy is looked up once, with that GPU current.
//
// Same quantity as the in-sweep modularity of the reference
// (Coloring/parallelLouvainWithColoring.cpp:226-249): Q = e_xx / (2m) -
// sum_c a_c^2 / (2m)^2 with k_i the row sums (sumVertexDegree,
// Utility/utilityClusteringFunctions.cpp:70-87) and a_c the sum of k_i over the
// community.
#include <mutex>

#include "kernels.hpp"
#include "noalloc.hpp"

namespace glv {
namespace {

constexpr int kNaBlock = 256;
constexpr int kNaBlocks = 1024;  // fixed grid of the partial sums: a fixed summation order
constexpr size_t kNaHeader = 64 << 10;  // flags and partial sums ahead of k[] and a[]

// The scratch memory: [0, 4) error flags, [256, 256 + 24 * kNaBlocks) partial
// sums, [kNaHeader, + 8 nv) vertex degrees k, then community degrees a.
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static_assert(256 + 3 * 8 * kNaBlocks + 3 * 8 <= kNaHeader, "header overflow");

__global__ void k_na_degrees(const i64 *off, const double *w, i64 nv, double *k, double *a) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        double s = 0;
        for (i64 j = off[v]; j < off[v + 1]; j++) s += w[j];  // row order
        k[v] = s;
        a[v] = 0.0;
    }
}

__global__ void k_na_comm(const i64 *C, const double *k, i64 nv, double *a) {
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const i64 c = C[v];
        if (c >= 0) atomicAdd(&a[c], k[v]);
    }
}

__device__ inline double na_block_sum(double v, double *sh) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    for (int q = 0; q < kNaBlock / kWave; q++) r += sh[q];
    return r;
}

// per block: internal weight, sum of degrees, sum of a^2 (grid of kNaBlocks)
__global__ __launch_bounds__(kNaBlock) void k_na_partials(const i64 *off, const i64 *tails, const double *w,
                                                          const i64 *C, const double *k, const double *a, i64 nv,
                                                          double *part) {
    __shared__ double sh[kNaBlock / kWave];
    double exx = 0, m2 = 0, a2 = 0;
    for (i64 v = blockIdx.x * (i64)blockDim.x + threadIdx.x; v < nv; v += (i64)gridDim.x * blockDim.x) {
        const i64 s = C[v];
        if (s >= 0)
            for (i64 j = off[v]; j < off[v + 1]; j++)
                if (C[tails[j]] == s) exx += w[j];
        m2 += k[v];
        a2 += a[v] * a[v];
    }
    exx = na_block_sum(exx, sh);
    m2 = na_block_sum(m2, sh);
    a2 = na_block_sum(a2, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = exx;
        part[3 * blockIdx.x + 1] = m2;
        part[3 * blockIdx.x + 2] = a2;
    }
}

__global__ void k_na_final(const double *part, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double e = 0, m = 0, q = 0;
        for (int b = 0; b < kNaBlocks; b++) {
            e += part[3 * b];
            m += part[3 * b + 1];
            q += part[3 * b + 2];
        }
        out[0] = e;
        out[1] = m;
        out[2] = q;
    }
}

__global__ void k_na_clear_flags(int32_t *err) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *err = 0;
}

std::mutex g_na_mu;
void *g_na_ptr[64] = {};

// the device address of this GPU's scratch (looked up once per device)
unsigned char *scratch_ptr() {
    int dev = 0;
    GLV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) fail(GLV_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(GLV_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_na_ptr[dev] = p;
    }
    return static_cast<unsigned char *>(g_na_ptr[dev]);
}

}  // namespace

i64 noalloc_max_nv() { return (i64)((kNoAllocScratchBytes - kNaHeader) / 16); }

double modularity_noalloc(const DevGraph &g, const i64 *C, hipStream_t s) {
    if (g.nv > noalloc_max_nv())
        fail(GLV_E_INVALID_ARG, "graph has more vertices than the no-allocation scratch holds (" +
                                    std::to_string(noalloc_max_nv()) + ")");
    unsigned char *base = scratch_ptr();
    int32_t *err = reinterpret_cast<int32_t *>(base);
    double *part = reinterpret_cast<double *>(base + 256);
    double *out = part + 3 * kNaBlocks;
    double *k = reinterpret_cast<double *>(base + kNaHeader);
    double *a = k + g.nv;
    // validation with the scratch's flag word (cleared by a kernel, no memset)
    i64 ends[2] = {0, 0};
    GLV_HIP(hipMemcpyAsync(&ends[0], g.off, sizeof(i64), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipMemcpyAsync(&ends[1], g.off + g.nv, sizeof(i64), hipMemcpyDeviceToHost, s));
    k_na_clear_flags<<<1, kWave, 0, s>>>(err);
    GLV_HIP(hipGetLastError());
    launch_check_graph(g.off, g.tails, g.nv, g.nnz, err, s);
    int32_t e = 0;
    GLV_HIP(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    if (ends[0] != 0 || ends[1] != g.nnz) fail(GLV_E_INVALID_GRAPH, "offsets do not span [0, nnz]");
    if (e & 1) fail(GLV_E_INVALID_GRAPH, "offsets not monotone");
    if (e & 2) fail(GLV_E_INVALID_GRAPH, "tail id out of range");
    k_na_clear_flags<<<1, kWave, 0, s>>>(err);
    GLV_HIP(hipGetLastError());
    launch_check_labels(g.off, C, g.nv, g.nv, err, s);
    GLV_HIP(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    if (e & 1) fail(GLV_E_INVALID_ARG, "cluster id out of range");
    if (e & 2) fail(GLV_E_INVALID_ARG, "vertex with edges has cluster id -1");
    // degrees, community degrees, partial sums, total
    const unsigned grid = grid_for(g.nv, kNaBlock, 8192);
    k_na_degrees<<<grid, kNaBlock, 0, s>>>(g.off, g.w, g.nv, k, a);
    GLV_HIP(hipGetLastError());
    k_na_comm<<<grid, kNaBlock, 0, s>>>(C, k, g.nv, a);
    GLV_HIP(hipGetLastError());
    k_na_partials<<<kNaBlocks, kNaBlock, 0, s>>>(g.off, g.tails, g.w, C, k, a, g.nv, part);
    GLV_HIP(hipGetLastError());
    k_na_final<<<1, kWave, 0, s>>>(part, out);
    GLV_HIP(hipGetLastError());
    double h[3];
    GLV_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    GLV_HIP(hipStreamSynchronize(s));
    const double c = (double)1 / h[1];
    return h[0] * c - h[2] * c * c;
}

}  // namespace glv
