// gen.cpp -- seeded synthetic graphs and CSR file formats (libglv_host.so).
//
// Random numbers come from a counter-based generator (the splitmix64
// finaliser applied to seed/stream/index), so every parallel stage gives the
// same graph for any thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "glv_gen.h"

namespace {

using i64 = int64_t;
using u64 = uint64_t;

int g_threads = 0;

int nthreads() {
    if (g_threads > 0) return g_threads;
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    return (int)std::min(h, 16u);
}

template <class F>
void parallel_for(i64 n, F fn) {
    int T = nthreads();
    if (n < 4096 || T == 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    i64 chunk = (n + T - 1) / T;
    for (int t = 0; t < T; t++) {
        i64 lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] { fn(lo, hi); });
    }
    for (auto &x : th) x.join();
}

inline u64 mix64(u64 z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline u64 crng(u64 seed, u64 stream, u64 idx) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) ^ mix64(idx + 0x632BE59BD9B4E019ULL));
}
inline double u01(u64 r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// sequential stream (for inherently serial draws)
struct Seq {
    u64 s;
    explicit Seq(u64 seed, u64 stream) : s(mix64(seed ^ (stream * 0xA24BAED4963EE407ULL))) {}
    u64 next() { return mix64(s += 0x9E3779B97F4A7C15ULL); }
    double uni() { return u01(next()); }
    i64 below(i64 n) { return (i64)(uni() * (double)n) % n; }
};

double pair_weight(u64 seed, i64 u, i64 v, double wlo, double whi) {
    if (!(whi > wlo)) return 1.0;
    i64 a = std::min(u, v), b = std::max(u, v);
    return wlo + (whi - wlo) * u01(crng(seed, 0x77, (u64)a * 0x9E3779B97F4A7C15ULL ^ (u64)b));
}

// Inverse-CDF table of a discrete power law P(k) ~ k^-gamma on [lo, hi].
struct PowerLaw {
    i64 lo;
    std::vector<double> cdf;
    PowerLaw(double gamma, i64 lo_, i64 hi) : lo(lo_) {
        double s = 0;
        for (i64 k = lo; k <= hi; k++) {
            s += std::pow((double)k, -gamma);
            cdf.push_back(s);
        }
        for (auto &x : cdf) x /= s;
        cdf.back() = 1.0;
    }
    i64 draw(double u) const {
        auto it = std::lower_bound(cdf.begin(), cdf.end(), u);
        return lo + (i64)(it - cdf.begin());
    }
};

// Symmetric CSR from undirected pairs (u != v); duplicates are removed, rows
// sorted by tail, weights given by wfn(u, v) (must be symmetric).
template <class WF>
int build_symmetric(i64 nv, const std::vector<i64> &pu, const std::vector<i64> &pv, WF wfn,
                    glv_csr *out) {
    i64 m = (i64)pu.size();
    std::vector<i64> deg(nv + 1, 0);
    for (i64 e = 0; e < m; e++) {
        deg[pu[e] + 1]++;
        deg[pv[e] + 1]++;
    }
    for (i64 i = 0; i < nv; i++) deg[i + 1] += deg[i];
    std::vector<i64> tmp(2 * m > 0 ? 2 * m : 1);
    {
        std::vector<i64> fill(deg.begin(), deg.end() - 1);
        for (i64 e = 0; e < m; e++) {
            tmp[fill[pu[e]]++] = pv[e];
            tmp[fill[pv[e]]++] = pu[e];
        }
    }
    std::vector<i64> nd(nv + 1, 0);
    parallel_for(nv, [&](i64 lo, i64 hi) {
        for (i64 i = lo; i < hi; i++) {
            auto b = tmp.begin() + deg[i], e = tmp.begin() + deg[i + 1];
            std::sort(b, e);
            nd[i + 1] = std::unique(b, e) - b;
        }
    });
    for (i64 i = 0; i < nv; i++) nd[i + 1] += nd[i];
    i64 nnz = nd[nv];
    out->nv = nv;
    out->nnz = nnz;
    out->num_edges = nnz / 2;
    out->off = (i64 *)malloc((nv + 1) * sizeof(i64));
    out->tails = (i64 *)malloc((nnz > 0 ? nnz : 1) * sizeof(i64));
    out->w = (double *)malloc((nnz > 0 ? nnz : 1) * sizeof(double));
    if (!out->off || !out->tails || !out->w) return GLV_E_OOM;
    memcpy(out->off, nd.data(), (nv + 1) * sizeof(i64));
    parallel_for(nv, [&](i64 lo, i64 hi) {
        for (i64 i = lo; i < hi; i++) {
            i64 n = nd[i + 1] - nd[i];
            for (i64 q = 0; q < n; q++) {
                i64 t = tmp[deg[i] + q];
                out->tails[nd[i] + q] = t;
                out->w[nd[i] + q] = wfn(i, t);
            }
        }
    });
    return GLV_OK;
}

void clear(glv_csr *g) { memset(g, 0, sizeof(*g)); }

}  // namespace

extern "C" {

void glv_gen_set_threads(int n) { g_threads = n; }

void glv_csr_free(glv_csr *g) {
    if (!g) return;
    free(g->off);
    free(g->tails);
    free(g->w);
    clear(g);
}

int glv_gen_planted(i64 nv, i64 nblocks, double mean_degree, double mu, double wlo, double whi,
                    u64 seed, i64 *block_out, glv_csr *out) {
    if (!out || nv < 2 || nblocks < 1 || nblocks > nv || mean_degree <= 0 || mu < 0 || mu > 1)
        return GLV_E_INVALID_ARG;
    clear(out);
    // block(i) = perm(i) mod nblocks, perm a seeded Fisher-Yates permutation
    std::vector<i64> perm(nv);
    for (i64 i = 0; i < nv; i++) perm[i] = i;
    Seq rp(seed, 1);
    for (i64 i = nv - 1; i > 0; i--) std::swap(perm[i], perm[rp.below(i + 1)]);
    std::vector<i64> block(nv), members(nv), bstart(nblocks + 1, 0);
    for (i64 i = 0; i < nv; i++) {
        block[i] = perm[i] % nblocks;
        bstart[block[i] + 1]++;
    }
    for (i64 b = 0; b < nblocks; b++) bstart[b + 1] += bstart[b];
    {
        std::vector<i64> f(bstart.begin(), bstart.end() - 1);
        for (i64 i = 0; i < nv; i++) members[f[block[i]]++] = i;
    }
    if (block_out)
        for (i64 i = 0; i < nv; i++) block_out[i] = block[i];
    i64 target = (i64)std::llround((double)nv * mean_degree / 2.0);
    i64 max_pairs = nv * (nv - 1) / 2;
    if (target > max_pairs / 2) return GLV_E_INVALID_ARG;
    std::vector<i64> pu, pv;
    pu.reserve(target);
    pv.reserve(target);
    std::unordered_set<u64> seen;
    seen.reserve((size_t)target * 2);
    Seq r(seed, 2);
    i64 attempts = 0, limit = target * 64 + 1000000;
    while ((i64)pu.size() < target && attempts < limit) {
        attempts++;
        i64 u = r.below(nv);
        i64 bu = block[u];
        i64 v;
        i64 bsz = bstart[bu + 1] - bstart[bu];
        if (r.uni() < 1.0 - mu || nblocks == 1) {
            v = members[bstart[bu] + r.below(bsz)];
        } else {
            if (bsz == nv) continue;
            do {
                v = r.below(nv);
            } while (block[v] == bu);
        }
        if (u == v) continue;
        i64 a = std::min(u, v), b = std::max(u, v);
        u64 key = (u64)a * (u64)nv + (u64)b;
        if (!seen.insert(key).second) continue;
        pu.push_back(a);
        pv.push_back(b);
    }
    if ((i64)pu.size() < target) return GLV_E_INVALID_ARG;
    return build_symmetric(nv, pu, pv,
                           [&](i64 a, i64 b) { return pair_weight(seed, a, b, wlo, whi); }, out);
}

int glv_gen_lfr(i64 nv, double deg_exp, i64 dmin, i64 dmax, double comm_exp, i64 cmin, i64 cmax,
                double mu, double wlo, double whi, u64 seed, i64 *community_out, glv_csr *out) {
    if (!out || nv < 2 || dmin < 1 || dmax < dmin || cmin < 2 || cmax < cmin || mu < 0 || mu > 1)
        return GLV_E_INVALID_ARG;
    clear(out);
    // 1. degrees
    PowerLaw dl(deg_exp, dmin, dmax);
    std::vector<i64> deg(nv);
    parallel_for(nv, [&](i64 lo, i64 hi) {
        for (i64 i = lo; i < hi; i++) deg[i] = dl.draw(u01(crng(seed, 11, (u64)i)));
    });
    // 2. community sizes (sequential), the last one folded into its predecessor
    //    when it would fall below cmin
    PowerLaw cl(comm_exp, cmin, cmax);
    std::vector<i64> csize;
    {
        Seq r(seed, 12);
        i64 s = 0;
        while (s < nv) {
            i64 c = cl.draw(r.uni());
            if (s + c > nv) c = nv - s;
            csize.push_back(c);
            s += c;
        }
        if (csize.size() > 1 && csize.back() < cmin) {
            i64 last = csize.back();
            csize.pop_back();
            csize.back() += last;
        }
    }
    i64 ncomm = (i64)csize.size();
    std::vector<i64> cstart(ncomm + 1, 0);
    for (i64 c = 0; c < ncomm; c++) cstart[c + 1] = cstart[c] + csize[c];
    // 3. assign vertices: seeded permutation filled community by community
    std::vector<i64> perm(nv);
    for (i64 i = 0; i < nv; i++) perm[i] = i;
    {
        Seq r(seed, 13);
        for (i64 i = nv - 1; i > 0; i--) std::swap(perm[i], perm[r.below(i + 1)]);
    }
    std::vector<i64> comm(nv);
    for (i64 c = 0; c < ncomm; c++)
        for (i64 q = cstart[c]; q < cstart[c + 1]; q++) comm[perm[q]] = c;
    if (community_out)
        for (i64 i = 0; i < nv; i++) community_out[i] = comm[i];
    std::vector<i64> din(nv), dout(nv);
    for (i64 i = 0; i < nv; i++) {
        i64 in = (i64)std::llround((1.0 - mu) * (double)deg[i]);
        in = std::min(in, csize[comm[i]] - 1);
        din[i] = in;
        dout[i] = deg[i] - in;
    }
    // 4. internal stubs, wired per community by a configuration model
    std::vector<std::vector<i64>> cu(ncomm), cv(ncomm);
    parallel_for(ncomm, [&](i64 lo, i64 hi) {
        std::vector<i64> stubs;
        for (i64 c = lo; c < hi; c++) {
            stubs.clear();
            for (i64 q = cstart[c]; q < cstart[c + 1]; q++)
                for (i64 t = 0; t < din[perm[q]]; t++) stubs.push_back(perm[q]);
            Seq r(seed, 1000 + (u64)c);
            for (i64 i = (i64)stubs.size() - 1; i > 0; i--) std::swap(stubs[i], stubs[r.below(i + 1)]);
            for (size_t i = 0; i + 1 < stubs.size(); i += 2)
                if (stubs[i] != stubs[i + 1]) {
                    cu[c].push_back(std::min(stubs[i], stubs[i + 1]));
                    cv[c].push_back(std::max(stubs[i], stubs[i + 1]));
                }
        }
    });
    // 5. external stubs, wired globally
    std::vector<i64> pu, pv;
    {
        std::vector<i64> stubs;
        i64 tot = 0;
        for (i64 i = 0; i < nv; i++) tot += dout[i];
        stubs.reserve(tot);
        for (i64 i = 0; i < nv; i++)
            for (i64 t = 0; t < dout[i]; t++) stubs.push_back(i);
        Seq r(seed, 14);
        for (i64 i = (i64)stubs.size() - 1; i > 0; i--) std::swap(stubs[i], stubs[r.below(i + 1)]);
        size_t internal = 0;
        for (i64 c = 0; c < ncomm; c++) internal += cu[c].size();
        pu.reserve(internal + tot / 2);
        pv.reserve(internal + tot / 2);
        for (i64 c = 0; c < ncomm; c++) {
            pu.insert(pu.end(), cu[c].begin(), cu[c].end());
            pv.insert(pv.end(), cv[c].begin(), cv[c].end());
            std::vector<i64>().swap(cu[c]);
            std::vector<i64>().swap(cv[c]);
        }
        for (size_t i = 0; i + 1 < stubs.size(); i += 2)
            if (stubs[i] != stubs[i + 1]) {
                pu.push_back(std::min(stubs[i], stubs[i + 1]));
                pv.push_back(std::max(stubs[i], stubs[i + 1]));
            }
    }
    return build_symmetric(nv, pu, pv,
                           [&](i64 a, i64 b) { return pair_weight(seed, a, b, wlo, whi); }, out);
}

int glv_gen_rmat(int scale, i64 edge_factor, double a, double b, double c, u64 seed, glv_csr *out) {
    if (!out || scale < 1 || scale > 31 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0)
        return GLV_E_INVALID_ARG;
    clear(out);
    i64 nv = (i64)1 << scale;
    i64 m = edge_factor * nv;
    std::vector<i64> pu(m), pv(m);
    std::atomic<i64> dummy{0};
    parallel_for(m, [&](i64 lo, i64 hi) {
        for (i64 e = lo; e < hi; e++) {
            i64 u = 0, v = 0;
            for (int l = 0; l < scale; l++) {
                double r = u01(crng(seed, 21 + (u64)l, (u64)e));
                int bu, bv;
                if (r < a) {
                    bu = 0;
                    bv = 0;
                } else if (r < a + b) {
                    bu = 0;
                    bv = 1;
                } else if (r < a + b + c) {
                    bu = 1;
                    bv = 0;
                } else {
                    bu = 1;
                    bv = 1;
                }
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            pu[e] = std::min(u, v);
            pv[e] = std::max(u, v);
        }
    });
    (void)dummy;
    // drop self-loops (repeats are removed while building the CSR)
    i64 k = 0;
    for (i64 e = 0; e < m; e++)
        if (pu[e] != pv[e]) {
            pu[k] = pu[e];
            pv[k] = pv[e];
            k++;
        }
    pu.resize(k);
    pv.resize(k);
    return build_symmetric(nv, pu, pv, [](i64, i64) { return 1.0; }, out);
}

int glv_write_binary(const char *path, const glv_csr *g) {
    if (!path || !g || !g->off) return GLV_E_INVALID_ARG;
    if (g->nnz != 2 * g->num_edges) return GLV_E_INVALID_GRAPH;  // the format stores 2*NE entries
    FILE *f = fopen(path, "wb");
    if (!f) return GLV_E_IO;
    i64 hdr[3] = {g->nv, g->num_edges, 1};
    bool ok = fwrite(hdr, 8, 3, f) == 3 && fwrite(g->off, 8, g->nv + 1, f) == (size_t)(g->nv + 1);
    const i64 chunk = 1 << 20;
    std::vector<i64> buf(3 * chunk);
    i64 n = 0;
    for (i64 i = 0; ok && i < g->nv; i++)
        for (i64 j = g->off[i]; ok && j < g->off[i + 1]; j++) {
            buf[3 * n] = i;
            buf[3 * n + 1] = g->tails[j];
            memcpy(&buf[3 * n + 2], &g->w[j], 8);
            if (++n == chunk) {
                ok = fwrite(buf.data(), 24, n, f) == (size_t)n;
                n = 0;
            }
        }
    if (ok && n) ok = fwrite(buf.data(), 24, n, f) == (size_t)n;
    ok = (fclose(f) == 0) && ok;
    return ok ? GLV_OK : GLV_E_IO;
}

int glv_read_binary(const char *path, glv_csr *out) {
    if (!path || !out) return GLV_E_INVALID_ARG;
    clear(out);
    FILE *f = fopen(path, "rb");
    if (!f) return GLV_E_IO;
    i64 hdr[3];
    if (fread(hdr, 8, 3, f) != 3 || hdr[0] < 0 || hdr[1] < 0) {
        fclose(f);
        return GLV_E_IO;
    }
    i64 nv = hdr[0], nnz = 2 * hdr[1];
    out->nv = nv;
    out->num_edges = hdr[1];
    out->nnz = nnz;
    out->off = (i64 *)malloc((nv + 1) * 8);
    out->tails = (i64 *)malloc((nnz > 0 ? nnz : 1) * 8);
    out->w = (double *)malloc((nnz > 0 ? nnz : 1) * 8);
    if (!out->off || !out->tails || !out->w) {
        fclose(f);
        glv_csr_free(out);
        return GLV_E_OOM;
    }
    bool ok = fread(out->off, 8, nv + 1, f) == (size_t)(nv + 1);
    const i64 chunk = 1 << 20;
    std::vector<i64> buf(3 * chunk);
    for (i64 j = 0; ok && j < nnz;) {
        i64 n = std::min(chunk, nnz - j);
        ok = fread(buf.data(), 24, n, f) == (size_t)n;
        for (i64 q = 0; ok && q < n; q++, j++) {
            out->tails[j] = buf[3 * q + 1];
            memcpy(&out->w[j], &buf[3 * q + 2], 8);
        }
    }
    fclose(f);
    if (ok && (out->off[0] != 0 || out->off[nv] != nnz)) ok = false;
    for (i64 j = 0; ok && j < nnz; j++)
        if (out->tails[j] < 0 || out->tails[j] >= nv) ok = false;
    if (!ok) {
        glv_csr_free(out);
        return GLV_E_INVALID_GRAPH;
    }
    return GLV_OK;
}

// Metis format: header "NV NE [fmt]"; line i lists the 1-based neighbours of
// vertex i (fmt 1: "neighbour weight" pairs).  Self-edges are skipped and
// unweighted edges get weight 1, as in loadMetis.cpp:86-110.
int glv_load_metis(const char *path, glv_csr *out) {
    if (!path || !out) return GLV_E_INVALID_ARG;
    clear(out);
    FILE *f = fopen(path, "r");
    if (!f) return GLV_E_IO;
    std::string line;
    auto getline = [&](std::string &s) {
        s.clear();
        int ch;
        bool any = false;
        while ((ch = fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            s.push_back((char)ch);
        }
        return any;
    };
    do {
        if (!getline(line)) {
            fclose(f);
            return GLV_E_IO;
        }
    } while (!line.empty() && line[0] == '%');
    long long nv = 0, ne = 0, fmt = 0;
    int nread = sscanf(line.c_str(), "%lld %lld %lld", &nv, &ne, &fmt);
    if (nread < 2 || nv < 0 || (fmt != 0 && fmt != 1)) {
        fclose(f);
        return GLV_E_IO;
    }
    std::vector<i64> off(nv + 1, 0), tails;
    std::vector<double> ws;
    for (i64 i = 0; i < nv; i++) {
        if (!getline(line)) break;
        const char *p = line.c_str();
        char *end;
        while (true) {
            long long nb = strtoll(p, &end, 10);
            if (end == p) break;
            p = end;
            double wt = 1.0;
            if (fmt == 1) {
                wt = strtod(p, &end);
                if (end == p) break;
                p = end;
            }
            if (nb - 1 == i) continue;
            if (nb < 1 || nb > nv) {
                fclose(f);
                return GLV_E_INVALID_GRAPH;
            }
            tails.push_back(nb - 1);
            ws.push_back(wt);
        }
        off[i + 1] = (i64)tails.size();
    }
    fclose(f);
    i64 nnz = (i64)tails.size();
    out->nv = nv;
    out->nnz = nnz;
    out->num_edges = ne;
    out->off = (i64 *)malloc((nv + 1) * 8);
    out->tails = (i64 *)malloc((nnz > 0 ? nnz : 1) * 8);
    out->w = (double *)malloc((nnz > 0 ? nnz : 1) * 8);
    if (!out->off || !out->tails || !out->w) return GLV_E_OOM;
    memcpy(out->off, off.data(), (nv + 1) * 8);
    if (nnz) {
        memcpy(out->tails, tails.data(), nnz * 8);
        memcpy(out->w, ws.data(), nnz * 8);
    }
    return GLV_OK;
}

}  // extern "C"
