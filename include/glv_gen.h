/*
 * glv_gen.h -- seeded synthetic graphs and the on-disk CSR formats.
 *
 * Host-only C ABI of libglv_host.so.  Generators have no counterpart in the
 * reference (they build the benchmark inputs); the readers/writers follow the
 * reference's formats:
 *   glv_write_binary / glv_read_binary  <->  writeGraphBinaryFormatNew
 *       (InputsOutput/writeBinary.cpp:2-29) / parse_EdgeListBinaryNew
 *       (InputsOutput/loadBinary.cpp:6-48), the "-f 9" file:
 *       int64 NV, int64 NE, int64 weighted, int64 ptrs[NV+1], edge[2*NE]
 *   glv_load_metis  <->  loadMetisFileFormat (InputsOutput/loadMetis.cpp:19-232)
 *
 * All graphs are symmetric CSR (structure of arrays), rows sorted by tail,
 * no duplicate entries.  Generated graphs have no self-loops.  Every array in
 * glv_csr is malloc()ed by the library; release it with glv_csr_free.
 * Return value: 0 on success, a GLV_E_* code otherwise.
 */
#ifndef GLV_GEN_H
#define GLV_GEN_H
#include <stdint.h>

#include "glv_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glv_csr {
    int64_t nv;
    int64_t nnz;       /* entries = off[nv] */
    int64_t num_edges; /* the reference's numEdges field (nnz/2 for simple graphs) */
    int64_t *off;      /* [nv+1] */
    int64_t *tails;    /* [nnz]  */
    double *w;         /* [nnz]  */
} glv_csr;

/* Planted partition: nblocks equal blocks, block(i) = perm(i) mod nblocks
 * for a seeded permutation; nv*mean_degree/2 distinct undirected edges drawn
 * one at a time (u uniform; v uniform inside u's block with probability
 * 1-mu, else uniform outside it; self-loops and repeats rejected).
 * Weights: 1.0, or U[wlo, whi] when whi > wlo.  block_out (optional, nv). */
int glv_gen_planted(int64_t nv, int64_t nblocks, double mean_degree, double mu, double wlo,
                    double whi, uint64_t seed, int64_t *block_out, glv_csr *out);

/* LFR-style benchmark: degrees from a discrete power law (exponent deg_exp on
 * [dmin, dmax]), community sizes from a power law (comm_exp on [cmin, cmax]),
 * a fraction mu of each vertex's stubs wired outside its community by a
 * configuration model, the rest inside; self-loops and repeats dropped.
 * Weights U[wlo, whi], the same in both directions.  community_out optional. */
int glv_gen_lfr(int64_t nv, double deg_exp, int64_t dmin, int64_t dmax, double comm_exp,
                int64_t cmin, int64_t cmax, double mu, double wlo, double whi, uint64_t seed,
                int64_t *community_out, glv_csr *out);

/* R-MAT: 2^scale vertices, edge_factor * 2^scale draws with quadrant
 * probabilities a, b, c (d = 1-a-b-c); self-loops and repeats removed, then
 * symmetrised.  Unit weights. */
int glv_gen_rmat(int scale, int64_t edge_factor, double a, double b, double c, uint64_t seed,
                 glv_csr *out);

void glv_csr_free(glv_csr *g);

int glv_write_binary(const char *path, const glv_csr *g);
int glv_read_binary(const char *path, glv_csr *out);
int glv_load_metis(const char *path, glv_csr *out);

/* threads used by the generators (default: hardware concurrency, at most 16) */
void glv_gen_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
